// The C ABI, part 5: node radiosity: the sparse Jacobi on host-given tables, the node cache on the uploaded scene
// (gather table, direct light, Jacobi, eAvg, the interior tree; gpu/renderer.js:456-852), the dense solver.
#include "vt_ctx.h"

using namespace vt;

// ====================================================================== node radiosity
extern "C" int vt_jacobi_sparse(vt_ctx *c, const vt_gather_tables *t, float *node_state, uint32_t iters,
                                float *aux_out) {
  VT_LIVE(c);
  if (!t || !node_state || !t->active || !t->gtable || !t->normals || !t->node_mat || !t->albedo || !t->hh)
    return fail(c, VT_ERR_ARG, "jacobi: null table");
  if (t->nbr_k && (!t->nbr || !t->nbr_w)) return fail(c, VT_ERR_ARG, "jacobi: null neighbour table");
  if (t->nbr_k > 32) return fail(c, VT_ERR_ARG, "jacobi: at most 32 smoothing neighbours");
  if (t->gather_k == 0 || t->albedo_bins == 0 || t->n_mat == 0) return fail(c, VT_ERR_ARG, "jacobi: empty sizes");
  const size_t N = t->n_static, A = t->n_active, K = t->gather_k;
  for (size_t k = 0; k < A; k++)
    if (t->active[k] >= N) return fail(c, VT_ERR_ARG, "jacobi: active node out of range");
  for (size_t k = 0; k < N * K; k++)
    if (t->gtable[k] != 0xffffffffu && t->gtable[k] >= N) return fail(c, VT_ERR_ARG, "jacobi: gather entry out of range");
  for (size_t k = 0; k < N; k++)
    if (t->node_mat[k] >= t->n_mat) return fail(c, VT_ERR_ARG, "jacobi: node material out of range");
  for (size_t k = 0; k < A * t->nbr_k; k++)
    if (t->nbr_w[k] > 0 && t->nbr[k] >= A) return fail(c, VT_ERR_ARG, "jacobi: neighbour out of range");
  VT_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = upload(c, c->s_active, t->active, A * 4))) return rc;
  if ((rc = upload(c, c->s_table, t->gtable, N * K * 4))) return rc;
  if ((rc = upload(c, c->s_normals, t->normals, N * 16))) return rc;
  if ((rc = upload(c, c->s_mat, t->node_mat, N * 4))) return rc;
  if ((rc = upload(c, c->s_alb, t->albedo, (size_t)t->n_mat * t->albedo_bins * 32))) return rc;
  if ((rc = upload(c, c->s_hh, t->hh, (size_t)t->n_mat * 16))) return rc;
  if (t->nbr_k) {
    if ((rc = upload(c, c->s_nbr, t->nbr, A * t->nbr_k * 4))) return rc;
    if ((rc = upload(c, c->s_nbrw, t->nbr_w, A * t->nbr_k * 4))) return rc;
  }
  if ((rc = upload(c, c->s_state, node_state, N * 64))) return rc;
  if ((rc = ensure(c, c->s_aux, std::max<size_t>(A, 1) * 16))) return rc;
  GatherDev G{};
  G.n_active = (uint32_t)A; G.gather_k = (uint32_t)K; G.nbr_k = t->nbr_k; G.bins = t->albedo_bins;
  G.active = (const uint32_t *)c->s_active.p; G.gtable = (const uint32_t *)c->s_table.p;
  G.node_mat = (const uint32_t *)c->s_mat.p; G.nbr = (const uint32_t *)c->s_nbr.p;
  G.normals = (const float4 *)c->s_normals.p; G.albedo = (const float4 *)c->s_alb.p; G.hh = (const float4 *)c->s_hh.p;
  G.nbr_w = (const float *)c->s_nbrw.p; G.node_state = (float4 *)c->s_state.p; G.aux = (float4 *)c->s_aux.p;
  // ping-pong slots 1 and 2 (gpu/renderer.js:729-749): step k reads slot_in and writes the other
  uint32_t slot_in = 1;
  for (uint32_t it = 0; it < iters; it++) {
    G.slot_in = slot_in;
    G.slot_out = 3 - slot_in;
    VT_HIP(c, launch_jacobi_gather(G, c->stream));
    VT_HIP(c, launch_jacobi_update(G, c->stream));
    slot_in = 3 - slot_in;
  }
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(node_state, c->s_state.p, N * 64, hipMemcpyDeviceToHost));
  if (aux_out) VT_HIP(c, hipMemcpy(aux_out, c->s_aux.p, A * 16, hipMemcpyDeviceToHost));
  return VT_OK;
}

// ====================================================================== node radiosity on the uploaded scene
namespace {

// smoothing neighbours of every node in a list (_neighbours, gpu/renderer.js:574-616): the same-facing
// (n.n' >= 0.9) nodes of the list within R, Gaussian weights exp(-d^2 / (2 sigma^2)), the K largest (ties: lower
// index first), sorted by weight; a uniform grid of cell R finds the candidates. sigma <= 0: each node alone.
void node_neighbours(const float *pos, const float *nrm, size_t n, uint32_t K, double R, double sg,
                     std::vector<uint32_t> &nb, std::vector<float> &nw) {
  nb.assign(n * K, 0u);
  nw.assign(n * K, 0.0f);
  if (K == 0) return;
  if (!(sg > 0)) {
    for (size_t k = 0; k < n; k++) {
      nb[k * K] = (uint32_t)k;
      nw[k * K] = 1.0f;
    }
    return;
  }
  const int64_t OFF = 1 << 20;
  auto cell = [&](double x) { return (int64_t)std::floor(x / R) + OFF; };
  auto key = [](int64_t x, int64_t y, int64_t z) {
    return ((uint64_t)(x & 0x1fffff) << 42) | ((uint64_t)(y & 0x1fffff) << 21) | (uint64_t)(z & 0x1fffff);
  };
  std::vector<std::pair<uint64_t, uint32_t>> order(n);
  for (size_t k = 0; k < n; k++)
    order[k] = {key(cell(pos[3 * k]), cell(pos[3 * k + 1]), cell(pos[3 * k + 2])), (uint32_t)k};
  std::sort(order.begin(), order.end());
  std::vector<double> cw(K);
  std::vector<uint32_t> cm(K);
  for (size_t k = 0; k < n; k++) {
    const double px = pos[3 * k], py = pos[3 * k + 1], pz = pos[3 * k + 2];
    const double nx = nrm[3 * k], ny = nrm[3 * k + 1], nz = nrm[3 * k + 2];
    const int64_t cx = cell(px), cy = cell(py), cz = cell(pz);
    uint32_t cnt = 0;
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -1; dz <= 1; dz++) {
          const uint64_t kk = key(cx + dx, cy + dy, cz + dz);
          auto lo = std::lower_bound(order.begin(), order.end(), std::make_pair(kk, 0u));
          for (auto it = lo; it != order.end() && it->first == kk; ++it) {
            const uint32_t m = it->second;
            const double ex = pos[3 * m] - px, ey = pos[3 * m + 1] - py, ez = pos[3 * m + 2] - pz;
            const double d2 = ex * ex + ey * ey + ez * ez;
            if (d2 > R * R) continue;
            if (nx * nrm[3 * m] + ny * nrm[3 * m + 1] + nz * nrm[3 * m + 2] < 0.9) continue;
            const double w = std::exp(-0.5 * d2 / (sg * sg));
            uint32_t at = cnt < K ? cnt : K;
            while (at > 0 && (cw[at - 1] < w || (cw[at - 1] == w && cm[at - 1] > m))) at--;
            if (at >= K) continue;
            for (uint32_t q = std::min(cnt, K - 1); q > at; q--) {
              cw[q] = cw[q - 1];
              cm[q] = cm[q - 1];
            }
            cw[at] = w;
            cm[at] = m;
            if (cnt < K) cnt++;
          }
        }
    for (uint32_t q = 0; q < cnt; q++) {
      nb[k * K + q] = cm[q];
      nw[k * K + q] = (float)cw[q];
    }
  }
}

hipError_t jacobi_weights(vt_ctx *c) {
  GatherDev G{};
  G.n_active = c->rad_active; G.gather_k = c->rad_K; G.bins = ALB_BINS;
  G.active = (const uint32_t *)c->g_active.p; G.gtable = (const uint32_t *)c->g_table.p;
  G.node_mat = (const uint32_t *)c->g_mat.p; G.normals = (const float4 *)c->g_normals.p;
  G.albedo = (const float4 *)c->g_alb.p; G.gw = (float4 *)c->g_w.p;
  return launch_jacobi_weights(G, c->stream);
}

NodeLaunch node_launch(vt_ctx *c) {
  NodeLaunch L;
  std::memset(&L, 0, sizeof L);
  L.S = c->S;
  L.npos = (const float4 *)c->g_pos.p;
  L.nnorm = (const float4 *)c->g_normals.p;
  L.nmat = (const uint32_t *)c->g_mat.p;
  L.n_nodes = (uint32_t)(c->h_verts.size() / 8);
  L.K = c->rad_K;
  L.node_samples = c->rad_samples;
  L.n_active = c->rad_active;
  L.n_mat = c->n_mat;
  L.gtable = (uint32_t *)c->g_table.p;
  L.miss = (float *)c->g_miss.p;
  L.active = (const uint32_t *)c->g_active.p;
  L.state = (float4 *)c->g_state.p;
  L.alb = (float4 *)c->g_alb.p;
  L.hh = (float4 *)c->g_hh.p;
  L.overflow = (uint32_t *)c->overflow.p;
  L.spill = (uint2 *)c->spill.p;
  L.spill_stride = c->spill_lanes;
  L.max_blocks = (uint32_t)(c->spill_lanes / 256);
  return L;
}


}  // namespace

extern "C" int vt_node_neighbours(const float *pos, const float *nrm, uint32_t n, uint32_t k, float radius, float sigma,
                                  uint32_t *idx, float *w) {
  if (n && (!pos || !nrm || !idx || !w)) return VT_ERR_ARG;
  if (k == 0 || k > 32 || !(radius > 0)) return VT_ERR_ARG;
  std::vector<uint32_t> nb;
  std::vector<float> nw;
  node_neighbours(pos, nrm, n, k, radius, sigma, nb, nw);
  std::memcpy(idx, nb.data(), nb.size() * 4);
  std::memcpy(w, nw.data(), nw.size() * 4);
  return VT_OK;
}

extern "C" int vt_radiosity_build(vt_ctx *c, const vt_radiosity_opts *o, vt_radiosity_info *info) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "radiosity: no scene uploaded");
  vt_radiosity_opts op{64, 4, 16, 0.006f, 0.003f};
  if (o) op = *o;
  if (op.gather_k < 8 || op.gather_k > 1024 || op.node_samples < 1 || op.node_samples > 1024 || op.nbr_k < 1 ||
      op.nbr_k > 32 || !(op.smooth_radius > 0) || !(op.smooth_sigma >= 0))
    return fail(c, VT_ERR_ARG, "radiosity: gather_k 8..1024, node_samples 1..1024, nbr_k 1..32, radius > 0, sigma >= 0");
  const auto t0 = std::chrono::steady_clock::now();
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  // the gather rays walk the whole static scene; an interior tree of an earlier build is dropped (its slots stay
  // unused until vt_interior_bvh builds the next one from the new miss counts)
  if (c->S.root) c->scene_gen++;
  c->S.root = 0;
  c->interior_root = 0;
  c->interior_tris = 0;
  const size_t V = c->h_verts.size() / 8;
  // static nodes: the vertices of static-flagged triangles of the static scene, with a third of every such
  // triangle's area; a node's material is that of its own mesh (the vertex record's word 3, as the reference's
  // prepared vertices carry it, gpu/scene_gpu.js:384-387)
  std::vector<float4> npos(V), nnorm(V);
  std::vector<uint32_t> nmat(V, 0u);
  std::vector<float> area(V, 0.0f);
  std::vector<uint8_t> stat(V, 0);
  for (uint32_t sl = 0; sl < c->static_slots; sl++) {
    const uint4 t = c->h_tvid[sl];
    if (!(c->h_meshes[4 * (size_t)t.w + 1] & VT_MESH_STATIC)) continue;
    const uint32_t vv[3] = {t.x, t.y, t.z};
    const float *a = &c->h_verts[8 * (size_t)t.x], *b = &c->h_verts[8 * (size_t)t.y], *q = &c->h_verts[8 * (size_t)t.z];
    const double e1[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
    const double e2[3] = {(double)q[0] - a[0], (double)q[1] - a[1], (double)q[2] - a[2]};
    const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
    const double ar = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
    for (uint32_t v : vv) {
      stat[v] = 1;
      area[v] += (float)(ar / 3.0);
    }
  }
  const uint32_t n_mesh = (uint32_t)(c->h_meshes.size() / 4);
  for (size_t v = 0; v < V; v++) {
    if (!stat[v]) continue;
    uint32_t mesh;
    std::memcpy(&mesh, &c->h_verts[8 * v + 3], 4);
    if (mesh >= n_mesh) return fail(c, VT_ERR_ARG, "radiosity: vertex %zu mesh id %u out of range", v, mesh);
    nmat[v] = c->h_meshes[4 * (size_t)mesh];
  }
  uint32_t n_static = 0;
  for (size_t v = 0; v < V; v++) {
    const float *p = &c->h_verts[8 * v];
    double nx = p[4], ny = p[5], nz = p[6];
    const double l = std::sqrt(nx * nx + ny * ny + nz * nz);
    if (l > 0) { nx /= l; ny /= l; nz /= l; } else { nx = 0; ny = 0; nz = 1; }
    npos[v] = make_float4(p[0], p[1], p[2], stat[v] ? 1.0f : 0.0f);
    nnorm[v] = make_float4((float)nx, (float)ny, (float)nz, 0.0f);
    n_static += stat[v];
  }
  int rc;
  if ((rc = upload(c, c->g_pos, npos.data(), V * sizeof(float4)))) return rc;
  if ((rc = upload(c, c->g_normals, nnorm.data(), V * sizeof(float4)))) return rc;
  if ((rc = upload(c, c->g_mat, nmat.data(), V * 4))) return rc;
  if ((rc = upload(c, c->g_area, area.data(), V * 4))) return rc;
  if ((rc = ensure(c, c->g_table, V * op.gather_k * 4))) return rc;
  if ((rc = ensure(c, c->g_miss, std::max<size_t>(V, 1) * 4))) return rc;
  if ((rc = ensure(c, c->g_alb, (size_t)c->n_mat * ALB_BINS * 32))) return rc;
  if ((rc = ensure(c, c->g_hh, (size_t)c->n_mat * 16))) return rc;
  if ((rc = ensure(c, c->g_state, std::max<size_t>(V, 1) * 64))) return rc;
  if ((rc = ensure(c, c->radu, 2 * sizeof(float4)))) return rc;
  VT_HIP(c, launch_zero_f4((float4 *)c->g_state.p, (uint64_t)V * 4, c->stream));
  c->rad_K = op.gather_k;
  c->rad_samples = op.node_samples;
  c->rad_nbr_k = op.nbr_k;
  c->rad_active = 0;
  NodeLaunch L = node_launch(c);
  VT_HIP(c, launch_albedo_tables(L, c->stream));
  VT_HIP(c, launch_gather_build(L, c->stream));
  VT_HIP(c, launch_gather_count(L, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<float> miss(V);
  if (V) VT_HIP(c, hipMemcpy(miss.data(), c->g_miss.p, V * 4, hipMemcpyDeviceToHost));
  std::vector<uint32_t> act;
  for (size_t v = 0; v < V; v++)
    if (stat[v] && miss[v] <= 0.5f * (float)op.gather_k) act.push_back((uint32_t)v);
  const size_t A = act.size();
  if ((rc = upload(c, c->g_active, act.empty() ? nullptr : act.data(), std::max<size_t>(A, 1) * 4))) return rc;
  c->rad_active = (uint32_t)A;
  L = node_launch(c);
  VT_HIP(c, launch_gather_remap(L, c->stream));
  // smoothing neighbours among the active nodes, by active index
  std::vector<float> ap(3 * A), an(3 * A);
  for (size_t k = 0; k < A; k++)
    for (int j = 0; j < 3; j++) {
      ap[3 * k + j] = (&npos[act[k]].x)[j];
      an[3 * k + j] = (&nnorm[act[k]].x)[j];
    }
  std::vector<uint32_t> nb;
  std::vector<float> nw;
  node_neighbours(ap.data(), an.data(), A, op.nbr_k, op.smooth_radius, op.smooth_sigma, nb, nw);
  if ((rc = upload(c, c->g_nbr, nb.empty() ? nullptr : nb.data(), std::max<size_t>(nb.size(), 1) * 4))) return rc;
  if ((rc = upload(c, c->g_nbrw, nw.empty() ? nullptr : nw.data(), std::max<size_t>(nw.size(), 1) * 4))) return rc;
  if ((rc = ensure(c, c->g_aux, std::max<size_t>(A, 1) * 16))) return rc;
  // the emitters' centroid and radius in the emitter's frame, for the shadow culling (gpu/scene_gpu.js:418-422)
  double cen[3] = {0, 0, 0}, tot = 0, rad = 0;
  const size_t E = c->h_emit.size() / 16;
  for (size_t e = 0; e < E; e++) {
    const float *r = &c->h_emit[16 * e];
    for (int k = 0; k < 3; k++) cen[k] += ((double)r[k] + ((double)r[4 + k] + r[8 + k]) / 3.0) * r[7];
    tot += r[7];
  }
  for (int k = 0; k < 3; k++) cen[k] /= tot > 0 ? tot : 1.0;
  for (size_t e = 0; e < E; e++) {
    const float *r = &c->h_emit[16 * e];
    for (int q = 0; q < 3; q++) {
      double d2 = 0;
      for (int k = 0; k < 3; k++) {
        const double x = (double)r[k] + (q == 1 ? r[4 + k] : 0.0) + (q == 2 ? r[8 + k] : 0.0) - cen[k];
        d2 += x * x;
      }
      rad = std::max(rad, std::sqrt(d2));
    }
  }
  // the gather entries' static weights, once per build, when the table stays within 4 GiB (else each step
  // computes them)
  c->rad_weights = false;
  if (A && (uint64_t)A * op.gather_k * 16 <= (4ull << 30)) {
    if ((rc = ensure(c, c->g_w, (size_t)A * op.gather_k * 16))) return rc;
    VT_HIP(c, jacobi_weights(c));
    c->rad_weights = true;
  }
  c->rad_direct_stale = false;
  c->rad_pcaps_valid = false;
  c->rad_slot = 1;
  VT_HIP(c, launch_fill_f4((float4 *)c->radu.p, 1, make_float4(0.f, 0.f, 0.f, bits_f(1u)), c->stream));
  VT_HIP(c, launch_fill_f4((float4 *)c->radu.p + 1, 1, make_float4((float)cen[0], (float)cen[1], (float)cen[2], (float)rad),
                           c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  c->S.nstate = (const float4 *)c->g_state.p;
  c->S.alb = (const float4 *)c->g_alb.p;
  c->S.hh = (const float4 *)c->g_hh.p;
  c->S.radu = (const float4 *)c->radu.p;
  c->rad_built = true;
  drop_graph(c);
  if (info) {
    info->nodes = (uint32_t)V;
    info->static_nodes = n_static;
    info->active = (uint32_t)A;
    info->ms_build = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return VT_OK;
}

// _buildInteriorBVH (gpu/renderer.js:533-571): a second static tree without the triangles whose three nodes see
// the sky in >= 90 % of their gather rays (the shell's outer skin; bench triangles always stay). Its triangles take
// new slots past the scene's (copies of their records, so a hit's slot still names its vertices and mesh) and its
// nodes follow the scene's; bodies keep their trees.
static int build_interior(vt_ctx *c) {
  const size_t V = c->h_verts.size() / 8;
  const uint32_t T = c->tri_count, K = c->rad_K;
  std::vector<float> miss(V);
  if (V) VT_HIP(c, hipMemcpy(miss.data(), c->g_miss.p, V * 4, hipMemcpyDeviceToHost));
  std::vector<uint32_t> keep;
  for (uint32_t k = 0; k < c->static_slots; k++) {
    const uint4 t = c->h_tvid[k];
    const bool bench = (c->h_meshes[4 * (size_t)t.w + 1] & 64u) != 0;    // MESH_FLAG.BENCH
    const float lim = 0.9f * (float)K;
    if (!bench && miss[t.x] >= lim && miss[t.y] >= lim && miss[t.z] >= lim) continue;
    keep.push_back(k);
  }
  const uint32_t n = (uint32_t)keep.size();
  c->interior_root = 0;
  c->interior_tris = 0;
  if (n == 0 || n == c->static_slots) return VT_OK;     // nothing to drop
  if ((uint64_t)T + n >= (1u << 26)) return fail(c, VT_ERR_SIZE, "interior_bvh: more than 2^26 slots");
  std::vector<float> pos(V * 3);
  for (size_t v = 0; v < V; v++)
    for (int k = 0; k < 3; k++) pos[3 * v + k] = c->h_verts[8 * v + k];
  std::vector<uint32_t> idx((size_t)n * 3), order(n);
  for (uint32_t m = 0; m < n; m++) {
    const uint4 t = c->h_tvid[keep[m]];
    idx[3 * m] = t.x;
    idx[3 * m + 1] = t.y;
    idx[3 * m + 2] = t.z;
  }
  std::vector<float> ng2((size_t)(n + 2) * 16);
  const int64_t nn = build_bvh2(pos.data(), idx.data(), n, 4, 16, ng2.data(), n + 2, order.data());
  if (nn < 0) return fail(c, VT_ERR_ARG, "interior_bvh: BVH build failed");
  std::vector<Node4> g4;
  try {
    g4 = collapse_bvh4(ng2.data(), (uint32_t)nn, n);
  } catch (const std::exception &ex) {
    return fail(c, VT_ERR_ARG, "interior_bvh: %s", ex.what());
  }
  // the scene's nodes, then the interior tree's, relocated
  std::vector<Node4> nodes(c->node4_count + g4.size());
  VT_HIP(c, hipMemcpy(nodes.data(), c->nodes.p, (size_t)c->node4_count * sizeof(Node4), hipMemcpyDeviceToHost));
  for (size_t q = 0; q < g4.size(); q++) {
    Node4 nd = g4[q];
    for (int ch = 0; ch < 4; ch++) {
      const uint32_t r = nd.ref[ch];
      if (r == EMPTY_REF) continue;
      if (r & LEAF_BIT) nd.ref[ch] = (r & ~0x03ffffffu) | ((r & 0x03ffffffu) + T);
      else nd.ref[ch] = r + c->node4_count;
    }
    nodes[c->node4_count + q] = nd;
  }
  // slot records for T + n slots: the scene's, then copies of the kept triangles in the interior tree's order
  const size_t S2 = (size_t)T + n;
  std::vector<uint4> attr(S2);
  for (uint32_t k = 0; k < T; k++) attr[k] = c->h_tvid[k];
  for (uint32_t m = 0; m < n; m++) attr[T + m] = c->h_tvid[keep[order[m]]];
  std::vector<float4> tri(S2 * 3), vn(V);
  for (size_t k = 0; k < S2; k++) {
    const uint32_t vid[3] = {attr[k].x, attr[k].y, attr[k].z};
    for (int j = 0; j < 3; j++) {
      const float *p = &c->h_verts[8 * (size_t)vid[j]];
      tri[3 * k + j] = make_float4(p[0], p[1], p[2], 0.0f);
    }
  }
  for (size_t v = 0; v < V; v++) {
    const float *p = &c->h_verts[8 * v];
    vn[v] = make_float4(p[4], p[5], p[6], 0.0f);
  }
  int rc;
  if ((rc = upload(c, c->nodes, nodes.data(), nodes.size() * sizeof(Node4)))) return rc;
  if ((rc = upload(c, c->tri, tri.data(), tri.size() * sizeof(float4)))) return rc;
  if ((rc = upload(c, c->tvid, attr.data(), attr.size() * sizeof(uint4)))) return rc;
  DevBuf d_vn;
  rc = upload(c, d_vn, vn.data(), vn.size() * sizeof(float4));
  if (!rc) rc = ensure(c, c->tshade, S2 * 64);
  if (!rc) {
    hipError_t e = launch_tri_shading((const float4 *)c->tri.p, (const uint4 *)c->tvid.p, (const float4 *)d_vn.p,
                                      (uint32_t)S2, (float4 *)c->tshade.p, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_fail(c, e, "interior triangle shading records");
  }
  release(d_vn);
  if (rc) return rc;
  c->S.nodes = (const Node4 *)c->nodes.p;
  c->S.tri = (const float4 *)c->tri.p;
  c->S.tvid = (const uint4 *)c->tvid.p;
  c->S.tshade = (const float4 *)c->tshade.p;
  c->interior_root = c->node4_count;
  c->interior_tris = n;
  return VT_OK;
}

extern "C" int vt_interior_bvh(vt_ctx *c, uint32_t enable, uint32_t *interior_tris) {
  VT_LIVE(c);
  if (!c->rad_built) return fail(c, VT_ERR_STATE, "interior_bvh: the node cache is not built");
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  if (enable && !c->interior_tris) {
    if (int rc = build_interior(c)) return rc;
  }
  c->S.root = enable ? c->interior_root : 0u;
  c->scene_gen++;                 // the fixed cameras' caches hold slots and hits of the other tree
  drop_graph(c);
  if (interior_tris) *interior_tris = c->S.root ? c->interior_tris : c->static_slots;
  return VT_OK;
}

int vt::refresh_node_materials(vt_ctx *c) {
  VT_HIP(c, hipSetDevice(c->device));
  NodeLaunch L = node_launch(c);
  VT_HIP(c, launch_albedo_tables(L, c->stream));
  if (c->rad_weights) VT_HIP(c, jacobi_weights(c));
  return VT_OK;
}

namespace {

// nodeDirect over every active node (moved = 0) or over those near the moved bodies, then the bodies' bounding
// capsules become the snapshot the next incremental pass compares with (_solvedCaps, gpu/renderer.js:786)
int direct_pass(vt_ctx *c, uint32_t moved) {
  const uint32_t B = (uint32_t)c->body_root.size();
  if (int rc = ensure(c, c->rad_pcaps, std::max<size_t>(B, 1) * 2 * sizeof(float4))) return rc;
  NodeLaunch L = node_launch(c);
  L.moved = moved;
  L.pcaps = (const float4 *)c->rad_pcaps.p;
  VT_HIP(c, launch_node_direct(L, c->stream));
  // instance words 6 and 7 of each body: its bounding capsule
  if (B)
    VT_HIP(c, hipMemcpy2DAsync(c->rad_pcaps.p, 2 * sizeof(float4), (const char *)c->inst.p + 6 * sizeof(float4),
                               INST_WORDS * sizeof(float4), 2 * sizeof(float4), B, hipMemcpyDeviceToDevice, c->stream));
  c->rad_pcaps_valid = true;
  c->rad_direct_stale = false;
  return VT_OK;
}

// `iters` Jacobi steps from the current source slot, then eAvg and the slot on the device
int jacobi_steps(vt_ctx *c, uint32_t iters) {
  GatherDev G{};
  G.n_active = c->rad_active; G.gather_k = c->rad_K; G.nbr_k = c->rad_nbr_k; G.bins = ALB_BINS;
  G.active = (const uint32_t *)c->g_active.p; G.gtable = (const uint32_t *)c->g_table.p;
  G.node_mat = (const uint32_t *)c->g_mat.p; G.nbr = (const uint32_t *)c->g_nbr.p;
  G.normals = (const float4 *)c->g_normals.p; G.albedo = (const float4 *)c->g_alb.p; G.hh = (const float4 *)c->g_hh.p;
  G.nbr_w = (const float *)c->g_nbrw.p; G.node_state = (float4 *)c->g_state.p; G.aux = (float4 *)c->g_aux.p;
  G.gw = c->rad_weights ? (float4 *)c->g_w.p : nullptr;
  // ping-pong source slots 1 and 2 (gpu/renderer.js:729-749): a step reads slot_in and writes the other
  for (uint32_t it = 0; it < iters; it++) {
    G.slot_in = c->rad_slot;
    G.slot_out = 3 - c->rad_slot;
    VT_HIP(c, launch_jacobi_gather(G, c->stream));
    VT_HIP(c, launch_jacobi_update(G, c->stream));
    c->rad_slot = 3 - c->rad_slot;
  }
  // eAvg (the mean irradiance a body-to-body bounce reads) and the source slot, on the device
  if (int erc = ensure(c, c->eavg_part, 256 * 4 * sizeof(double))) return erc;
  const NodeLaunch L = node_launch(c);
  VT_HIP(c, launch_eavg(L, (const float *)c->g_area.p, (float4 *)c->radu.p, c->rad_slot, (double *)c->eavg_part.p,
                        c->stream));
  return VT_OK;
}

}  // namespace

extern "C" int vt_radiosity_update(vt_ctx *c, uint32_t direct, uint32_t iters) {
  VT_LIVE(c);
  if (!c->rad_built) return fail(c, VT_ERR_STATE, "radiosity: not built");
  VT_HIP(c, hipSetDevice(c->device));
  c->S.nstate = (const float4 *)c->g_state.p;     // a scene's instances may have changed S since the build
  if (direct || c->rad_direct_stale)
    if (int rc = direct_pass(c, 0)) return rc;
  return jacobi_steps(c, iters);
}

extern "C" int vt_radiosity_update_moved(vt_ctx *c, uint32_t moved, uint32_t iters, uint32_t *ran) {
  VT_LIVE(c);
  if (!c->rad_built) return fail(c, VT_ERR_STATE, "radiosity: not built");
  const uint32_t B = (uint32_t)c->body_root.size();
  if (B < 32 && (moved >> B)) return fail(c, VT_ERR_ARG, "radiosity_update_moved: mask 0x%x names bodies past %u", moved, B);
  VT_HIP(c, hipSetDevice(c->device));
  c->S.nstate = (const float4 *)c->g_state.p;
  uint32_t kind = VT_DIRECT_NONE;
  // the light's own body moved: every node sees the light from elsewhere (_movedSinceSolve's scope case)
  const bool light_moved = c->S.emit_body < 32 && ((moved >> c->S.emit_body) & 1u);
  if (c->rad_direct_stale || (moved && (!c->rad_pcaps_valid || light_moved))) kind = VT_DIRECT_FULL;
  else if (moved) kind = VT_DIRECT_INCREMENTAL;
  if (kind != VT_DIRECT_NONE)
    if (int rc = direct_pass(c, kind == VT_DIRECT_FULL ? 0u : moved)) return rc;
  if (ran) *ran = kind;
  return jacobi_steps(c, iters);
}

extern "C" int vt_radiosity_tables(vt_ctx *c, uint32_t which, void *dst, size_t bytes) {
  VT_LIVE(c);
  if (!c->rad_built) return fail(c, VT_ERR_STATE, "radiosity: not built");
  const size_t V = c->h_verts.size() / 8, A = c->rad_active;
  const DevBuf *src = nullptr;
  size_t need = 0;
  switch (which) {
    case VT_RAD_GATHER: src = &c->g_table; need = V * c->rad_K * 4; break;
    case VT_RAD_MISSES: src = &c->g_miss; need = V * 4; break;
    case VT_RAD_ACTIVE: src = &c->g_active; need = A * 4; break;
    case VT_RAD_NEIGHBOURS: src = &c->g_nbr; need = A * c->rad_nbr_k * 4; break;
    case VT_RAD_WEIGHTS: src = &c->g_nbrw; need = A * c->rad_nbr_k * 4; break;
    case VT_RAD_ALBEDO: src = &c->g_alb; need = (size_t)c->n_mat * ALB_BINS * 32; break;
    case VT_RAD_HH: src = &c->g_hh; need = (size_t)c->n_mat * 16; break;
    default: return fail(c, VT_ERR_ARG, "radiosity_tables: table %u", which);
  }
  if (bytes != need || (need && !dst)) return fail(c, VT_ERR_SIZE, "radiosity_tables: table %u has %zu bytes", which, need);
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  if (need) VT_HIP(c, hipMemcpy(dst, src->p, need, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_radiosity_state(vt_ctx *c, float *node_state, size_t bytes, float e_avg[3]) {
  VT_LIVE(c);
  if (!c->rad_built) return fail(c, VT_ERR_STATE, "radiosity: not built");
  const size_t V = c->h_verts.size() / 8;
  if (node_state && bytes != V * 64) return fail(c, VT_ERR_SIZE, "radiosity_state: %zu bytes for %zu nodes", V * 64, V);
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  if (node_state && V) VT_HIP(c, hipMemcpy(node_state, c->g_state.p, V * 64, hipMemcpyDeviceToHost));
  float4 ru;
  VT_HIP(c, hipMemcpy(&ru, c->radu.p, sizeof ru, hipMemcpyDeviceToHost));
  if (e_avg) {
    e_avg[0] = ru.x;
    e_avg[1] = ru.y;
    e_avg[2] = ru.z;
  }
  return VT_OK;
}

static int dense_alloc(vt_ctx *c, uint32_t n, uint32_t r) {
  const size_t nn = (size_t)n * n, m = (size_t)n * r;
  int rc;
  if ((rc = ensure(c, c->dF32, nn * 4))) return rc;
  if ((rc = ensure(c, c->dE, m * 4))) return rc;
  if ((rc = ensure(c, c->dX, m * 4))) return rc;
  if ((rc = ensure(c, c->dY, m * 4))) return rc;
  if ((rc = ensure(c, c->dR, m * 4))) return rc;
  if ((rc = ensure(c, c->dXd, m * 16))) return rc;      // x64, y64
  if ((rc = ensure(c, c->dPart, RES_BLOCKS_H * 4 * 8))) return rc;
  c->dn_n = n;
  c->dn_r = r;
  c->dn_rp = r;
  release(c->dF16);
  return VT_OK;
}

extern "C" int vt_dense_upload(vt_ctx *c, uint32_t n, uint32_t r, const float *F, const float *e) {
  VT_LIVE(c);
  if (!F || !e || n == 0 || r == 0) return fail(c, VT_ERR_ARG, "dense: null or empty input");
  if (r > 48) return fail(c, VT_ERR_ARG, "dense: at most 48 right-hand sides");
  VT_HIP(c, hipSetDevice(c->device));
  int rc = dense_alloc(c, n, r);
  if (rc) return rc;
  VT_HIP(c, hipMemcpy(c->dF32.p, F, (size_t)n * n * 4, hipMemcpyHostToDevice));
  VT_HIP(c, hipMemcpy(c->dE.p, e, (size_t)n * r * 4, hipMemcpyHostToDevice));
  return VT_OK;
}

extern "C" int vt_dense_formfactors(vt_ctx *c, uint32_t n, const float *nodes, float row_max, uint32_t r,
                                    const float *e) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "dense: form factors need an uploaded scene");
  if (!nodes || !e || n == 0 || r == 0 || r > 48) return fail(c, VT_ERR_ARG, "dense: bad form-factor input");
  VT_HIP(c, hipSetDevice(c->device));
  int rc = dense_alloc(c, n, r);
  if (rc) return rc;
  DevBuf nb;
  if ((rc = upload(c, nb, nodes, (size_t)n * 32))) return rc;
  VT_HIP(c, hipMemcpy(c->dE.p, e, (size_t)n * r * 4, hipMemcpyHostToDevice));
  hipError_t er = launch_formfactors(c->S, (uint32_t *)c->overflow.p, (uint2 *)c->spill.p, c->spill_lanes, n,
                                     (const float4 *)nb.p, (float *)c->dF32.p, c->stream);
  if (er == hipSuccess) er = launch_row_scale((float *)c->dF32.p, n, row_max, c->stream);
  if (er == hipSuccess) er = hipStreamSynchronize(c->stream);
  release(nb);
  if (er != hipSuccess) return hip_fail(c, er, "form factors");
  return VT_OK;
}

extern "C" int vt_dense_download_F(vt_ctx *c, float *F_out, size_t bytes) {
  VT_LIVE(c);
  if (!c->dn_n) return fail(c, VT_ERR_STATE, "dense: nothing uploaded");
  if (!F_out || bytes != (size_t)c->dn_n * c->dn_n * 4) return fail(c, VT_ERR_SIZE, "dense: F is n*n*4 bytes");
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipMemcpy(F_out, c->dF32.p, bytes, hipMemcpyDeviceToHost));
  return VT_OK;
}

namespace {
struct Norms {
  double res2, e2, err2, ref2;
};
int read_norms(vt_ctx *c, Norms &o) {
  double p[RES_BLOCKS_H * 4];
  VT_HIP(c, hipMemcpyAsync(p, c->dPart.p, sizeof p, hipMemcpyDeviceToHost, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  o = {0, 0, 0, 0};
  for (int b = 0; b < RES_BLOCKS_H; b++) {
    o.res2 += p[4 * b];
    o.e2 += p[4 * b + 1];
    o.err2 += p[4 * b + 2];
    o.ref2 += p[4 * b + 3];
  }
  return VT_OK;
}
}  // namespace

extern "C" int vt_dense_solve(vt_ctx *c, const vt_dense_opts *o, float *x_out, vt_dense_result *res) {
  VT_LIVE(c);
  if (!o || !res) return VT_ERR_ARG;
  if (!c->dn_n) return fail(c, VT_ERR_STATE, "dense: nothing uploaded");
  if (o->n != c->dn_n || o->r != c->dn_r) return fail(c, VT_ERR_ARG, "dense: n, r differ from the upload");
  const uint32_t n = c->dn_n, rp = c->dn_rp;
  const uint64_t m = (uint64_t)n * rp;
  const uint32_t maxs = o->max_sweeps ? o->max_sweeps : 1000;
  VT_HIP(c, hipSetDevice(c->device));
  std::memset(res, 0, sizeof *res);
  float *F = (float *)c->dF32.p, *E = (float *)c->dE.p, *X = (float *)c->dX.p, *Y = (float *)c->dY.p,
        *R = (float *)c->dR.p;
  double *X64 = (double *)c->dXd.p, *Y64 = X64 + m, *part = (double *)c->dPart.p;
  int rc;
  Norms nm;
  VT_HIP(c, hipEventRecord(c->ev_s0, c->stream));
  // x transposed (RHS-major) for the MFMA sweeps
  DevBuf xT32;
  if ((rc = ensure(c, xT32, (size_t)rp * n * 4))) return rc;
  float *XT = (float *)xT32.p;
  struct Free {
    DevBuf &b;
    ~Free() { release(b); }
  } free_xT{xT32};
  if (o->method == VT_DENSE_FP32) {
    VT_HIP(c, hipMemcpyAsync(X, E, m * 4, hipMemcpyDeviceToDevice, c->stream));
    for (uint32_t s = 0; s < maxs; s++) {
      VT_HIP(c, launch_sweep_fp32_mfma(F, n, rp, X, XT, E, Y, c->stream));
      VT_HIP(c, launch_resid_f32(Y, X, E, nullptr, m, part, c->stream));
      std::swap(X, Y);
      res->sweeps++;
      if ((rc = read_norms(c, nm))) return rc;
      if (std::sqrt(nm.res2) <= o->tol * std::sqrt(nm.e2)) break;
    }
  } else if (o->method == VT_DENSE_BF16_REFINE) {
    if (n % 128) return fail(c, VT_ERR_ARG, "dense bf16: n must be a multiple of 128");
    if (!c->dF16.p) {
      if ((rc = ensure(c, c->dF16, (size_t)n * n * 2))) return rc;
      VT_HIP(c, launch_convert_bf16(F, (uint16_t *)c->dF16.p, (uint64_t)n * n, c->stream));
    }
    DevBuf xT, D0, D1;
    if ((rc = ensure(c, xT, m * 2))) return rc;
    if ((rc = ensure(c, D0, m * 4))) return rc;
    if ((rc = ensure(c, D1, m * 4))) return rc;
    float *Da = (float *)D0.p, *Db = (float *)D1.p;
    // inner sweeps: a fixed count, or (0) until the correction changes by less than BF16's resolution
    const uint32_t inner = o->inner_sweeps ? o->inner_sweeps : 64;
    const bool adaptive = o->inner_sweeps == 0;
    VT_HIP(c, hipMemsetAsync(X, 0, m * 4, c->stream));
    bool ok = true;
    hipError_t er = hipSuccess;
    while (res->sweeps < maxs) {
      // FP32 residual r = e + F x - x
      if ((er = launch_sweep_fp32_mfma(F, n, rp, X, XT, E, Y, c->stream)) != hipSuccess) break;
      if ((er = launch_sub(Y, X, R, m, c->stream)) != hipSuccess) break;
      if ((er = launch_resid_f32(Y, X, E, nullptr, m, part, c->stream)) != hipSuccess) break;
      res->sweeps++;
      if ((rc = read_norms(c, nm))) { ok = false; break; }
      if (std::sqrt(nm.res2) <= o->tol * std::sqrt(nm.e2)) break;
      // correction d = r + F d by BF16 sweeps from d = r, then x += d
      if ((er = hipMemcpyAsync(Da, R, m * 4, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess) break;
      for (uint32_t k = 0; k < inner && er == hipSuccess; k++) {
        er = launch_x_to_bf16T(Da, n, rp, (uint16_t *)xT.p, c->stream);
        if (er == hipSuccess) er = launch_sweep_bf16_T((const uint16_t *)c->dF16.p, (const uint16_t *)xT.p, n, rp, R, Db, c->stream);
        res->sweeps++;
        std::swap(Da, Db);
        if (adaptive && er == hipSuccess && k >= 1) {
          // |d_new - d_old| against |r|: below 2^-9 (BF16's unit roundoff) the sweeps cannot improve d
          if ((er = launch_resid_f32(Da, Db, R, nullptr, m, part, c->stream)) != hipSuccess) break;
          Norms dn;
          if ((rc = read_norms(c, dn))) { ok = false; break; }
          if (std::sqrt(dn.res2) <= 0.002 * std::sqrt(dn.e2)) break;
        }
      }
      if (!ok || er != hipSuccess) break;
      if ((er = launch_add(X, Da, m, c->stream)) != hipSuccess) break;
      res->refinements++;
    }
    release(xT);
    release(D0);
    release(D1);
    if (er != hipSuccess) return hip_fail(c, er, "bf16 refinement");
    if (!ok) return rc;
  } else if (o->method == VT_DENSE_FP64) {
    VT_HIP(c, launch_widen(E, X64, m, c->stream));
    for (uint32_t s = 0; s < maxs; s++) {
      VT_HIP(c, launch_sweep_fp64(F, n, rp, X64, E, Y64, c->stream));
      VT_HIP(c, launch_resid_f64(Y64, X64, E, nullptr, m, part, c->stream));
      std::swap(X64, Y64);
      res->sweeps++;
      if ((rc = read_norms(c, nm))) return rc;
      if (std::sqrt(nm.res2) <= o->tol * std::sqrt(nm.e2)) break;
    }
  } else {
    return fail(c, VT_ERR_ARG, "dense: unknown method %u", o->method);
  }
  VT_HIP(c, hipEventRecord(c->ev_s1, c->stream));
  VT_HIP(c, hipEventSynchronize(c->ev_s1));
  float ms = 0;
  VT_HIP(c, hipEventElapsedTime(&ms, c->ev_s0, c->ev_s1));
  res->ms = ms;
  res->ms_per_sweep = res->sweeps ? ms / res->sweeps : 0;
  // final residual in FP64: r = e + F x - x with F widened exactly
  double *Xc = X64, *Yc = Y64;
  if (o->method != VT_DENSE_FP64) VT_HIP(c, launch_widen(X, Xc, m, c->stream));
  VT_HIP(c, launch_sweep_fp64(F, n, rp, Xc, E, Yc, c->stream));
  VT_HIP(c, launch_resid_f64(Yc, Xc, E, nullptr, m, part, c->stream));
  if ((rc = read_norms(c, nm))) return rc;
  res->residual = std::sqrt(nm.res2) / std::max(1e-300, std::sqrt(nm.e2));
  if (x_out) {
    if (o->method == VT_DENSE_FP64) {
      std::vector<double> tmp(m);
      VT_HIP(c, hipMemcpy(tmp.data(), Xc, m * 8, hipMemcpyDeviceToHost));
      for (uint64_t i = 0; i < m; i++) x_out[i] = (float)tmp[i];
    } else {
      VT_HIP(c, hipMemcpy(x_out, X, m * 4, hipMemcpyDeviceToHost));
    }
  }
  // keep the solution in dX for later calls
  if (X != (float *)c->dX.p && o->method != VT_DENSE_FP64) VT_HIP(c, hipMemcpy(c->dX.p, X, m * 4, hipMemcpyDeviceToDevice));
  return VT_OK;
}

