// The C ABI (include/vt.h), part 1: context lifetime, the native BVH build, scene upload and re-layout (the
// reference's _buildBuffers, gpu/renderer.js:271-331), material / light / profile updates, cameras and posed
// bodies; and the host helpers every part shares. Device memory is owned by the context; one HIP stream (default
// priority) orders every call's work.
#include "vt_ctx.h"

using namespace vt;

namespace vt {

int fail(vt_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

int hip_fail(vt_ctx *c, hipError_t e, const char *what) {
  c->lost = true;
  return fail(c, VT_ERR_HIP, "%s: %s (%d); the context is lost", what, hipGetErrorString(e), (int)e);
}

int ensure(vt_ctx *c, DevBuf &b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return VT_OK;
  if (b.p) {
    hipError_t e = hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (e != hipSuccess) return hip_fail(c, e, "hipFree");
  }
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      return fail(c, VT_ERR_NOMEM, "device allocation of %zu bytes failed", bytes);
    }
    return hip_fail(c, e, "hipMalloc");
  }
  b.bytes = bytes;
  return VT_OK;
}
void release(DevBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

int upload(vt_ctx *c, DevBuf &b, const void *src, size_t bytes) {
  int rc = ensure(c, b, bytes);
  if (rc) return rc;
  if (bytes) VT_HIP(c, hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
  return VT_OK;
}

float bits_f(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

HostPose host_pose(const vt_instance &I, double t) {
  double a[4], b[4];
  for (int k = 0; k < 4; k++) { a[k] = I.q0[k]; b[k] = I.q1[k]; }
  double c = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
  if (c < 0) {
    for (double &x : b) x = -x;
    c = -c;
  }
  double q[4];
  if (c > 0.9995) {
    double l = 0;
    for (int k = 0; k < 4; k++) { q[k] = a[k] + (b[k] - a[k]) * t; l += q[k] * q[k]; }
    for (double &x : q) x /= std::sqrt(l);
  } else {
    const double th = std::acos(c), sn = std::sin(th);
    for (int k = 0; k < 4; k++) q[k] = (a[k] * std::sin((1 - t) * th) + b[k] * std::sin(t * th)) / sn;
  }
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  HostPose P;
  const double m[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w)},
                          {2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w)},
                          {2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)}};
  std::memcpy(P.c, m, sizeof m);
  for (int k = 0; k < 3; k++) P.p[k] = I.p0[k] + (I.p1[k] - I.p0[k]) * t;
  return P;
}
void host_apply(const HostPose &P, const double v[3], double out[3]) {
  for (int k = 0; k < 3; k++) out[k] = P.c[0][k] * v[0] + P.c[1][k] * v[1] + P.c[2][k] * v[2] + P.p[k];
}


// relative illumination per pixel (gpu/sensor.js:103-124), FP64: cos^q sin / (theta_d dtheta_d/dtheta) at the pixel's
// field angle, times the vignetting table read linearly between its knots and held at its ends
void ri_map_host(const vt_camera &cam, double q, uint32_t vn, const float *vt_deg, const float *vv,
                 std::vector<float> &out) {
  double vt[VT_MAX_VIGNETTING];
  for (uint32_t k = 0; k < vn; k++) vt[k] = (double)vt_deg[k] * M_PI / 180.0;
  auto vig = [&](double th) -> double {
    if (vn == 0) return 1.0;
    if (th <= vt[0]) return vv[0];
    for (uint32_t k = 1; k < vn; k++)
      if (th <= vt[k]) return vv[k - 1] + ((double)vv[k] - vv[k - 1]) * (th - vt[k - 1]) / (vt[k] - vt[k - 1]);
    return vv[vn - 1];
  };
  const double fx = cam.lens[0], fy = cam.lens[1], cx = cam.lens[2], cy = cam.lens[3], al = cam.lens[8];
  const double k[4] = {cam.lens[4], cam.lens[5], cam.lens[6], cam.lens[7]};
  auto dist = [&](double th) {
    double t2 = th * th;
    return th * (1 + t2 * (k[0] + t2 * (k[1] + t2 * (k[2] + t2 * k[3]))));
  };
  auto deriv = [&](double th) {
    double t2 = th * th;
    return 1 + t2 * (3 * k[0] + t2 * (5 * k[1] + t2 * (7 * k[2] + t2 * 9 * k[3])));
  };
  const uint32_t w = cam.width, h = cam.height;
  out.resize((size_t)w * h);
  for (uint32_t y = 0; y < h; y++)
    for (uint32_t x = 0; x < w; x++) {
      double ray_z;
      if (cam.mode == 0) {
        double yd = (y - cy) / fy, xd = (x - cx) / fx - al * yd;
        double thd = std::hypot(xd, yd);
        if (thd < 1e-300) {
          ray_z = 1;
        } else {
          double th = thd;
          for (int i = 0; i < 20; i++) {
            double step = (dist(th) - thd) / deriv(th);
            th -= step;
            if (std::fabs(step) < 1e-15) break;
          }
          ray_z = std::cos(th);
        }
      } else {
        double dx = (x - cx) / fx, dy = (y - cy) / fy;
        ray_z = 1 / std::sqrt(dx * dx + dy * dy + 1);
      }
      double th = std::acos(std::max(-1.0, std::min(1.0, ray_z)));
      double ri = 1;
      if (th > 1e-6) {
        if (cam.mode == 0) ri = std::pow(std::cos(th), q) * std::sin(th) / (dist(th) * deriv(th));
        else ri = std::pow(std::cos(th), q + 3);   // pinhole mapping: cos^3 * cos^q
      }
      out[(size_t)y * w + x] = (float)(std::max(0.0, ri) * vig(th));
    }
}

}  // namespace

// ====================================================================== lifetime
extern "C" int vt_create(const vt_create_info *info, vt_ctx **out) {
  if (!out) return VT_ERR_ARG;
  *out = nullptr;
  vt_ctx *c = new vt_ctx();
  c->device = info ? info->device : 0;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= c->device || c->device < 0) {
    delete c;
    return VT_ERR_HIP;
  }
  e = hipSetDevice(c->device);
  // the no-allocation raw stage's static glare array on this device (vt_sensor_raw_device)
  if (e == hipSuccess) e = sensor_scratch_init(c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  hipDeviceProp_t prop;
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, c->device);
  if (e != hipSuccess) {
    delete c;
    return VT_ERR_HIP;
  }
  c->cu_count = prop.multiProcessorCount;
  c->wf_budget = (uint64_t)prop.totalGlobalMem / 8;
  int occ = 0;
  // persistent grids: up to 8 resident 256-thread blocks per CU (LDS and VGPRs usually allow fewer)
  c->spill_lanes = (uint64_t)c->cu_count * 8 * 256;
  (void)occ;
  int rc = ensure(c, c->counters, 64);
  if (!rc) rc = ensure(c, c->overflow, 64);
  if (!rc) rc = ensure(c, c->spill, c->spill_lanes * SPILL_ENTRIES * sizeof(uint2));
  if (!rc) rc = ensure(c, c->dcams, sizeof(DevCam) * VT_MAX_CAMERAS);
  if (!rc && hipHostMalloc((void **)&c->h_cams, sizeof(DevCam) * VT_MAX_CAMERAS, hipHostMallocDefault) != hipSuccess) rc = VT_ERR_HIP;
  hipEvent_t *evs[] = {&c->ev_stage, &c->ev_r0, &c->ev_r1, &c->ev_s0, &c->ev_s1, &c->ev_b0, &c->ev_b1, &c->ev_i0,
                       &c->ev_i1};
  for (hipEvent_t *ev : evs)
    if (!rc && hipEventCreate(ev) != hipSuccess) rc = VT_ERR_HIP;
  if (!rc && hipMemsetAsync(c->overflow.p, 0, 64, c->stream) != hipSuccess) rc = VT_ERR_HIP;
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = VT_ERR_HIP;
  if (rc) {
    vt_destroy(c);
    return rc;
  }
  *out = c;
  return VT_OK;
}

extern "C" void vt_destroy(vt_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  DevBuf *bufs[] = {&c->nodes, &c->tri, &c->tshade, &c->meshes, &c->mats, &c->lights, &c->emit, &c->inst, &c->lprof, &c->knots, &c->tvid, &c->uv, &c->tex, &c->g_pos, &c->g_miss, &c->radu, &c->g_area,
                    &c->counters, &c->overflow, &c->spill, &c->dcams, &c->test_in, &c->test_out, &c->test_aux,
                    &c->refine_list, &c->eavg_part,
                    &c->g_active, &c->g_table, &c->g_normals, &c->g_mat, &c->g_alb, &c->g_hh, &c->g_nbr, &c->g_w,
                    &c->g_nbrw, &c->g_state, &c->g_aux, &c->s_active, &c->s_table, &c->s_normals, &c->s_mat,
                    &c->s_alb, &c->s_hh, &c->s_nbr, &c->s_nbrw, &c->s_state, &c->s_aux, &c->dF32, &c->dF16, &c->dF64, &c->dE, &c->dX, &c->dY,
                    &c->dXd, &c->dR, &c->dPart, &c->grads, &c->wf, &c->wf_temp, &c->ltot};
  for (DevBuf *b : bufs) release(*b);
  for (auto &cs : c->cams) {
    release(cs.acc);
    release(cs.dn);
    release(cs.ri);
    release(cs.lin);
    release(cs.adj);
    DevBuf *isp[] = {&cs.glare_a, &cs.glare_b, &cs.isp_rgb, &cs.isp_part, &cs.rgba, &cs.yuv, &cs.stats, &cs.rgb8,
                     &cs.keys, &cs.hcache, &cs.scache};
    for (DevBuf *b : isp) release(*b);
  }
  for (hipEvent_t ev : c->ev_launch) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : c->ev_ext) (void)hipEventDestroy(ev);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  hipEvent_t evs[] = {c->ev_stage, c->ev_r0, c->ev_r1, c->ev_s0, c->ev_s1, c->ev_b0, c->ev_b1, c->ev_i0, c->ev_i1};
  for (hipEvent_t ev : evs)
    if (ev) (void)hipEventDestroy(ev);
  if (c->h_cams) (void)hipHostFree(c->h_cams);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char *vt_last_error(const vt_ctx *c) { return c ? c->err.c_str() : "null context"; }

// ====================================================================== BVH
extern "C" int vt_build_bvh(const float *pos, uint32_t nvert, const uint32_t *idx, uint32_t ntri, int leaf_size,
                            int bins, float *nodes_out, uint32_t cap_nodes, uint32_t *node_count,
                            uint32_t *order_out) {
  if ((!pos && nvert) || (!idx && ntri) || !nodes_out || !node_count || (!order_out && ntri)) return VT_ERR_ARG;
  if (leaf_size < 1 || bins < 2 || bins > 256) return VT_ERR_ARG;
  for (uint64_t k = 0; k < 3ull * ntri; k++)
    if (idx[k] >= nvert) return VT_ERR_ARG;
  int64_t n = build_bvh2(pos, idx, ntri, leaf_size, bins, nodes_out, cap_nodes, order_out);
  if (n < 0) return VT_ERR_SIZE;
  *node_count = (uint32_t)n;
  return VT_OK;
}

// ====================================================================== scene
extern "C" int vt_upload_scene(vt_ctx *c, const vt_scene *s) {
  VT_LIVE(c);
  if (!s || !s->tris || !s->verts || !s->meshes || !s->materials) return fail(c, VT_ERR_ARG, "scene: null array");
  if (s->tri_count == 0 || s->vert_count == 0 || s->mesh_count == 0 || s->material_count == 0)
    return fail(c, VT_ERR_ARG, "scene: empty triangle, vertex, mesh or material table");
  if (s->emit_count && (!s->emit || !s->lights || s->light_count == 0))
    return fail(c, VT_ERR_ARG, "scene: emitter triangles without lights");
  if (s->tri_count >= (1u << 26)) return fail(c, VT_ERR_ARG, "scene: more than 2^26 triangles");
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  const uint32_t T = s->tri_count, V = s->vert_count;
  const uint32_t *tw = reinterpret_cast<const uint32_t *>(s->tris);
  // validate the triangle records
  for (uint32_t k = 0; k < T; k++) {
    const uint32_t *r = tw + (size_t)k * 16;
    if (r[3] >= V || r[7] >= V || r[11] >= V) return fail(c, VT_ERR_ARG, "scene: triangle %u vertex id out of range", k);
    if (r[12] >= s->mesh_count) return fail(c, VT_ERR_ARG, "scene: triangle %u mesh id out of range", k);
  }
  for (uint32_t m = 0; m < s->mesh_count; m++) {
    if (s->meshes[4 * m] >= s->material_count) return fail(c, VT_ERR_ARG, "scene: mesh %u material out of range", m);
    if ((s->meshes[4 * m + 1] & VT_MESH_EMITTER) && s->meshes[4 * m + 2] >= std::max(1u, s->light_count))
      return fail(c, VT_ERR_ARG, "scene: mesh %u light out of range", m);
  }
  for (uint32_t e = 0; e < s->emit_count; e++) {
    uint32_t lid;
    std::memcpy(&lid, s->emit + 16 * (size_t)e + 11, 4);
    if (lid >= s->light_count) return fail(c, VT_ERR_ARG, "scene: emitter %u light out of range", e);
  }
  if (s->body_count > VT_MAX_BODIES) return fail(c, VT_ERR_ARG, "scene: more than %d bodies", VT_MAX_BODIES);
  if (s->body_count && (!s->tri_body || s->nodes))
    return fail(c, VT_ERR_ARG, "scene: bodies need tri_body and no caller BVH (nodes = NULL)");
  if (s->emit_posed && s->emit_body >= s->body_count)
    return fail(c, VT_ERR_ARG, "scene: the light rides on body %u of %u", s->emit_body, s->body_count);
  // emitter samples come from the emit table in the emitter's frame, so the emitter triangles that rays hit must
  // lie in that same frame: the static scene, or the light's body
  const uint32_t emit_group = s->emit_posed ? s->emit_body : 0xffffffffu;
  for (uint32_t k = 0; k < T; k++) {
    const uint32_t g = s->body_count ? s->tri_body[k] : 0xffffffffu;
    if (g != 0xffffffffu && g >= s->body_count) return fail(c, VT_ERR_ARG, "scene: triangle %u body out of range", k);
    if ((s->meshes[4 * (size_t)tw[16 * (size_t)k + 12] + 1] & VT_MESH_EMITTER) && g != emit_group)
      return fail(c, VT_ERR_ARG, "scene: emitter triangle %u lies outside the light's %s", k,
                  s->emit_posed ? "body" : "static scene");
  }
  // BVH: the caller's (reference layout) or built here; slot order is the tris order
  std::vector<float> nodes2;
  std::vector<float> tris_sorted;
  const float *tr = s->tris;
  uint32_t n2 = s->node_count;
  const float *np = s->nodes;
  std::vector<Node4> n4;
  std::vector<uint32_t> body_root(s->body_count, 0);
  uint32_t body_slots = 0, static_count = 0;
  std::vector<float> body_box(6 * (size_t)s->body_count, 0.0f);
  if (s->body_count) {
    // one BVH per group (the static scene, then each body) over that group's triangles, built and collapsed
    // separately and concatenated: node refs offset by the group's first node, leaf slots by its first slot
    std::vector<float> pos((size_t)V * 3);
    for (uint32_t v = 0; v < V; v++)
      for (int k = 0; k < 3; k++) pos[3 * v + k] = s->verts[8 * (size_t)v + k];
    tris_sorted.resize((size_t)T * 16);
    uint32_t slot_base = 0;
    for (uint32_t g = 0; g <= s->body_count; g++) {
      const uint32_t want = g == 0 ? 0xffffffffu : g - 1;
      std::vector<uint32_t> members;
      for (uint32_t k = 0; k < T; k++)
        if (s->tri_body[k] == want) members.push_back(k);
      const uint32_t Tg = (uint32_t)members.size();
      if (g == 0) static_count = Tg;
      if (Tg == 0) return fail(c, VT_ERR_ARG, g ? "scene: body %u has no triangles" : "scene: no static triangles", g - 1);
      std::vector<uint32_t> idx((size_t)Tg * 3), order(Tg);
      for (uint32_t m = 0; m < Tg; m++)
        for (int j = 0; j < 3; j++) idx[3 * m + j] = tw[16 * (size_t)members[m] + 4 * j + 3];
      std::vector<float> ng2((size_t)(Tg + 2) * 16);
      const int64_t nn = build_bvh2(pos.data(), idx.data(), Tg, 4, 16, ng2.data(), Tg + 2, order.data());
      if (nn < 0) return fail(c, VT_ERR_ARG, "scene: BVH build failed");
      std::vector<Node4> g4;
      try {
        g4 = collapse_bvh4(ng2.data(), (uint32_t)nn, Tg);
      } catch (const std::exception &ex) {
        return fail(c, VT_ERR_ARG, "scene: %s", ex.what());
      }
      const uint32_t node_base = (uint32_t)n4.size();
      for (Node4 &nd : g4) {
        for (int ch = 0; ch < 4; ch++) {
          const uint32_t r = nd.ref[ch];
          if (r == EMPTY_REF) continue;
          if (r & LEAF_BIT) nd.ref[ch] = (r & ~0x03ffffffu) | ((r & 0x03ffffffu) + slot_base);
          else nd.ref[ch] = r + node_base;
        }
        n4.push_back(nd);
      }
      for (uint32_t m = 0; m < Tg; m++)
        std::memcpy(&tris_sorted[16 * (size_t)(slot_base + m)], s->tris + 16 * (size_t)members[order[m]], 64);
      if (g > 0) {
        body_root[g - 1] = node_base;
        float *bx = &body_box[6 * (size_t)(g - 1)];
        for (int k = 0; k < 3; k++) { bx[k] = INFINITY; bx[3 + k] = -INFINITY; }
        for (uint32_t m = 0; m < Tg; m++)
          for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
              const float x = pos[3 * (size_t)idx[3 * m + j] + k];
              bx[k] = std::min(bx[k], x);
              bx[3 + k] = std::max(bx[3 + k], x);
            }
      }
      slot_base += Tg;
    }
    tr = tris_sorted.data();
    body_slots = T - static_count;
  } else if (!np) {
    std::vector<float> pos((size_t)V * 3);
    for (uint32_t v = 0; v < V; v++)
      for (int k = 0; k < 3; k++) pos[3 * v + k] = s->verts[8 * (size_t)v + k];
    std::vector<uint32_t> idx((size_t)T * 3), order(T);
    for (uint32_t k = 0; k < T; k++)
      for (int j = 0; j < 3; j++) idx[3 * k + j] = tw[16 * (size_t)k + 4 * j + 3];
    nodes2.resize((size_t)(T + 2) * 16);
    int64_t n = build_bvh2(pos.data(), idx.data(), T, 4, 16, nodes2.data(), T + 2, order.data());
    if (n < 0) return fail(c, VT_ERR_ARG, "scene: BVH build failed");
    n2 = (uint32_t)n;
    np = nodes2.data();
    tris_sorted.resize((size_t)T * 16);
    for (uint32_t k = 0; k < T; k++) std::memcpy(&tris_sorted[16 * (size_t)k], s->tris + 16 * (size_t)order[k], 64);
    tr = tris_sorted.data();
  } else if (n2 == 0) {
    return fail(c, VT_ERR_ARG, "scene: node_count is 0");
  }
  if (!s->body_count) {
    try {
      n4 = collapse_bvh4(np, n2, T);
    } catch (const std::exception &ex) {
      return fail(c, VT_ERR_ARG, "scene: %s", ex.what());
    }
  }
  const uint32_t *trw = reinterpret_cast<const uint32_t *>(tr);
  std::vector<float4> tri((size_t)T * 3), vn(V);
  std::vector<uint4> attr(T);
  for (uint32_t k = 0; k < T; k++) {
    const uint32_t *r = trw + (size_t)k * 16;
    for (int j = 0; j < 3; j++) {
      const float *p = s->verts + 8 * (size_t)r[4 * j + 3];
      tri[3 * (size_t)k + j] = make_float4(p[0], p[1], p[2], 0.0f);
    }
    attr[k] = make_uint4(r[3], r[7], r[11], r[12]);
  }
  for (uint32_t v = 0; v < V; v++) {
    const float *p = s->verts + 8 * (size_t)v;
    vn[v] = make_float4(p[4], p[5], p[6], 0.0f);
  }
  int rc;
  if ((rc = upload(c, c->nodes, n4.data(), n4.size() * sizeof(Node4)))) return rc;
  if ((rc = upload(c, c->tri, tri.data(), tri.size() * sizeof(float4)))) return rc;
  {
    // per-slot vertex ids stay on the device (the node cache's lookups); vertex normals only feed the shading
    // records
    DevBuf d_vn;
    rc = upload(c, c->tvid, attr.data(), attr.size() * sizeof(uint4));
    if (!rc) rc = upload(c, d_vn, vn.data(), vn.size() * sizeof(float4));
    if (!rc) rc = ensure(c, c->tshade, (size_t)T * 64);
    if (!rc) {
      hipError_t e = launch_tri_shading((const float4 *)c->tri.p, (const uint4 *)c->tvid.p, (const float4 *)d_vn.p, T,
                                        (float4 *)c->tshade.p, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) rc = hip_fail(c, e, "triangle shading records");
    }
    release(d_vn);
    if (rc) return rc;
  }
  c->h_verts.assign(s->verts, s->verts + 8 * (size_t)V);
  c->h_tvid = attr;
  c->h_meshes.assign(s->meshes, s->meshes + 4 * (size_t)s->mesh_count);
  c->h_emit.assign(s->emit_count ? s->emit : nullptr, s->emit_count ? s->emit + 16 * (size_t)s->emit_count : nullptr);
  c->static_slots = s->body_count ? (uint32_t)(T - body_slots) : T;
  c->rad_built = false;
  c->S.nstate = nullptr;
  c->S.alb = nullptr;
  c->S.hh = nullptr;
  c->S.radu = nullptr;
  if ((rc = upload(c, c->meshes, s->meshes, (size_t)s->mesh_count * 16))) return rc;
  if ((rc = upload(c, c->mats, s->materials, (size_t)s->material_count * 64))) return rc;
  float zero4[4] = {0, 0, 0, 0};
  if ((rc = upload(c, c->lights, s->light_count ? (const void *)s->lights : zero4, std::max<size_t>(16, (size_t)s->light_count * 16)))) return rc;
  float zero16[16] = {0};
  if ((rc = upload(c, c->emit, s->emit_count ? (const void *)s->emit : zero16, std::max<size_t>(64, (size_t)s->emit_count * 64)))) return rc;
  c->S.nodes = (const Node4 *)c->nodes.p;
  c->S.tri = (const float4 *)c->tri.p;
  c->S.tshade = (const float4 *)c->tshade.p;
  c->S.meshes = (const uint4 *)c->meshes.p;
  c->S.tvid = (const uint4 *)c->tvid.p;
  {
    // texture coordinates of textured meshes, and the print mask: one paper texel until vt_set_board_texture
    c->has_textured = false;
    for (uint32_t k = 0; k < s->mesh_count; k++)
      if (s->meshes[4 * (size_t)k + 1] & VT_MESH_TEXTURED) c->has_textured = true;
    if (c->has_textured && !s->uvs)
      return fail(c, VT_ERR_ARG, "scene: a VT_MESH_TEXTURED mesh needs texture coordinates (vt_scene.uvs)");
    if (s->uvs) {
      if ((rc = upload(c, c->uv, s->uvs, (size_t)V * 8))) return rc;
    } else {
      release(c->uv);
    }
    const uint8_t paper = 255;
    if ((rc = upload(c, c->tex, &paper, 1))) return rc;
    c->S.uv = s->uvs ? (const float2 *)c->uv.p : nullptr;
    c->S.tex = (const uint8_t *)c->tex.p;
    c->S.tex_w = c->S.tex_h = 1;
  }
  c->S.mats = (const float4 *)c->mats.p;
  c->S.lights = (const float4 *)c->lights.p;
  {
    // every light Lambertian until vt_set_light_profile gives it a profile
    const size_t nl = std::max<uint32_t>(1, s->light_count);
    std::vector<float4> lp(nl, make_float4(0.f, 0.f, 1.f, 0.f));
    std::vector<float2> kn(nl * MAX_KNOTS, make_float2(0.f, 0.f));
    if ((rc = upload(c, c->lprof, lp.data(), lp.size() * sizeof(float4)))) return rc;
    if ((rc = upload(c, c->knots, kn.data(), kn.size() * sizeof(float2)))) return rc;
    c->S.lprof = (const float4 *)c->lprof.p;
    c->S.knots = (const float2 *)c->knots.p;
  }
  c->S.emit = (const float4 *)c->emit.p;
  c->S.emit_count = s->emit_count;
  c->S.root = 0;
  c->body_root = body_root;
  c->body_box = body_box;
  c->S.inst_count = s->body_count;
  c->S.shadow_skip = s->body_count ? s->shadow_skip : 0u;
  c->S.emit_body = s->emit_posed ? s->emit_body : 0xffffffffu;
  c->emit_rec_set = false;
  if (s->body_count) {
    // every body starts hidden until vt_set_instances gives it a pose
    std::vector<float4> rec((size_t)INST_WORDS * s->body_count, make_float4(0.f, 0.f, 0.f, 0.f));
    for (uint32_t b = 0; b < s->body_count; b++) {
      rec[INST_WORDS * b + 2].w = bits_f(body_root[b]);
      rec[INST_WORDS * b + 3].w = bits_f(INST_HIDDEN);
    }
    if ((rc = upload(c, c->inst, rec.data(), rec.size() * sizeof(float4)))) return rc;
    c->S.inst = (const float4 *)c->inst.p;
  } else {
    c->S.inst = nullptr;
  }
  c->S.emit_area = s->emit_area > 0 ? s->emit_area : 1e-12f;
  for (int k = 0; k < 3; k++) {
    c->bmin[k] = INFINITY;
    c->bmax[k] = -INFINITY;
  }
  for (uint32_t v = 0; v < V; v++)
    for (int k = 0; k < 3; k++) {
      c->bmin[k] = std::min(c->bmin[k], s->verts[8 * (size_t)v + k]);
      c->bmax[k] = std::max(c->bmax[k], s->verts[8 * (size_t)v + k]);
    }
  c->tri_count = T;
  c->node4_count = (uint32_t)n4.size();
  c->interior_root = 0;
  c->interior_tris = 0;
  c->n_mat = s->material_count;
  c->n_light = s->light_count;
  c->has_scene = true;
  c->scene_gen++;
  VT_HIP(c, hipMemsetAsync(c->overflow.p, 0, 64, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  return VT_OK;
}

extern "C" int vt_set_board_texture(vt_ctx *c, const uint8_t *mask, uint32_t width, uint32_t height) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "board texture: no scene uploaded");
  if (!mask || width == 0 || height == 0 || width > 16384 || height > 16384)
    return fail(c, VT_ERR_ARG, "board texture: null mask or size outside 1..16384");
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  if (int rc = upload(c, c->tex, mask, (size_t)width * height)) return rc;
  c->S.tex = (const uint8_t *)c->tex.p;
  c->S.tex_w = width;
  c->S.tex_h = height;
  return VT_OK;
}

extern "C" int vt_update_materials(vt_ctx *c, const float *materials, uint32_t count) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "update_materials: no scene uploaded");
  if (!materials || count != c->n_mat) return fail(c, VT_ERR_SIZE, "update_materials: %u materials expected", c->n_mat);
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(c->mats.p, materials, (size_t)count * 64, hipMemcpyHostToDevice));
  c->shade_gen++;
  if (c->rad_built) {
    // the node cache follows: its albedo tables and gather weights now, its direct light at the next update
    if (int rc = refresh_node_materials(c)) return rc;
    c->rad_direct_stale = true;
  }
  return VT_OK;
}

extern "C" int vt_update_lights(vt_ctx *c, const float *lights, uint32_t count) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "update_lights: no scene uploaded");
  if (!lights || count != c->n_light || count == 0) return fail(c, VT_ERR_SIZE, "update_lights: %u lights expected", c->n_light);
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(c->lights.p, lights, (size_t)count * 16, hipMemcpyHostToDevice));
  c->shade_gen++;
  c->rad_direct_stale = c->rad_built;
  return VT_OK;
}

static_assert(VT_MAX_KNOTS == MAX_KNOTS, "knot table stride");

extern "C" int vt_set_light_profile(vt_ctx *c, uint32_t light, const float axis[3], const float *theta, const float *p,
                                    uint32_t count) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "light_profile: no scene uploaded");
  if (light >= c->n_light) return fail(c, VT_ERR_ARG, "light_profile: light %u of %u", light, c->n_light);
  if (count > VT_MAX_KNOTS) return fail(c, VT_ERR_SIZE, "light_profile: at most %d knots", VT_MAX_KNOTS);
  float4 rec = make_float4(0.f, 0.f, 1.f, 0.f);
  std::vector<float2> kn(MAX_KNOTS, make_float2(0.f, 0.f));
  if (count) {
    if (!axis || !theta || !p) return fail(c, VT_ERR_ARG, "light_profile: null axis or knots");
    const double l = std::sqrt((double)axis[0] * axis[0] + (double)axis[1] * axis[1] + (double)axis[2] * axis[2]);
    if (!(l > 1e-12) || !std::isfinite(l)) return fail(c, VT_ERR_ARG, "light_profile: zero axis");
    for (uint32_t k = 0; k < count; k++) {
      if (!std::isfinite(theta[k]) || !std::isfinite(p[k]) || p[k] < 0.0f || (k && theta[k] < theta[k - 1]))
        return fail(c, VT_ERR_ARG, "light_profile: knots need finite ascending theta and p >= 0 (knot %u)", k);
      kn[k] = make_float2(theta[k], p[k]);
    }
    rec = make_float4((float)(axis[0] / l), (float)(axis[1] / l), (float)(axis[2] / l), bits_f(count));
  }
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy((float4 *)c->lprof.p + light, &rec, sizeof rec, hipMemcpyHostToDevice));
  VT_HIP(c, hipMemcpy((float2 *)c->knots.p + (size_t)light * MAX_KNOTS, kn.data(), kn.size() * sizeof(float2),
                      hipMemcpyHostToDevice));
  c->shade_gen++;
  c->rad_direct_stale = c->rad_built;
  return VT_OK;
}

extern "C" int vt_set_camera(vt_ctx *c, uint32_t cam, const vt_camera *k) {
  VT_LIVE(c);
  if (cam >= VT_MAX_CAMERAS || !k) return fail(c, VT_ERR_ARG, "camera slot %u out of range", cam);
  if (k->width == 0 || k->height == 0 || k->width > 16384 || k->height > 16384)
    return fail(c, VT_ERR_ARG, "camera %u: bad size %ux%u", cam, k->width, k->height);
  if (k->spp == 0) return fail(c, VT_ERR_ARG, "camera %u: spp is 0", cam);
  if (k->mode > 1) return fail(c, VT_ERR_ARG, "camera %u: mode %u is not kb4 (0) or pinhole (1)", cam, k->mode);
  if (!(k->lens[0] > 0) || !(k->lens[1] > 0)) return fail(c, VT_ERR_ARG, "camera %u: focal lengths must be > 0", cam);
  if (!(k->shutter[0] >= 0.0f && k->shutter[0] <= 1.0f && k->shutter[1] >= 0.0f && k->shutter[1] <= 1.0f))
    return fail(c, VT_ERR_ARG, "camera %u: shutter times must lie in [0, 1]", cam);
  if (k->refine > 64 || k->refine_light_samples > 16)
    return fail(c, VT_ERR_ARG, "camera %u: refine %u (<= 64), refine light samples %u (<= 16)", cam, k->refine,
                k->refine_light_samples);
  VT_HIP(c, hipSetDevice(c->device));
  CamSlot &s = c->cams[cam];
  const size_t px = (size_t)k->width * k->height;
  const bool resized = !s.set || s.c.width != k->width || s.c.height != k->height;
  const bool lens_changed = resized || std::memcmp(s.c.lens, k->lens, sizeof k->lens) != 0 || s.c.mode != k->mode;
  int rc;
  if ((rc = ensure(c, s.acc, px * sizeof(float4)))) return rc;
  if ((rc = ensure(c, s.dn, px * sizeof(uint16_t)))) return rc;
  if ((rc = ensure(c, s.ri, px * sizeof(float)))) return rc;
  if ((rc = ensure(c, s.lin, px * sizeof(float4)))) return rc;
  if (lens_changed) s.ri_valid = false;
  s.c = *k;
  s.set = true;
  if (resized) {
    s.rendered = false;
    s.sensed = false;
    s.spp_total = 0;
  }
  return VT_OK;
}

extern "C" int vt_set_instances(vt_ctx *c, const vt_instance *in, uint32_t count) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "set_instances: no scene uploaded");
  const uint32_t B = (uint32_t)c->body_root.size();
  if (count != B || (B && !in)) return fail(c, VT_ERR_SIZE, "set_instances: %u bodies in the scene", B);
  if (!B) return VT_OK;
  std::vector<float4> rec((size_t)INST_WORDS * B);
  for (uint32_t b = 0; b < B; b++) {
    vt_instance I = in[b];
    for (float *q : {I.q0, I.q1}) {
      const double l = std::sqrt((double)q[0] * q[0] + (double)q[1] * q[1] + (double)q[2] * q[2] + (double)q[3] * q[3]);
      if (!(l > 1e-12) || !std::isfinite(l)) return fail(c, VT_ERR_ARG, "set_instances: body %u has no unit quaternion", b);
      for (int k = 0; k < 4; k++) q[k] = (float)(q[k] / l);
    }
    bool moving = false;
    for (int k = 0; k < 4; k++) moving |= I.q0[k] != I.q1[k];
    for (int k = 0; k < 3; k++) moving |= I.p0[k] != I.p1[k];
    const uint32_t flags = (moving ? INST_MOVING : 0u) | (I.hidden ? INST_HIDDEN : 0u);
    float4 *r = &rec[INST_WORDS * (size_t)b];
    r[0] = make_float4(I.q0[0], I.q0[1], I.q0[2], I.q0[3]);
    r[1] = make_float4(I.q1[0], I.q1[1], I.q1[2], I.q1[3]);
    r[2] = make_float4(I.p0[0], I.p0[1], I.p0[2], bits_f(c->body_root[b]));
    r[3] = make_float4(I.p1[0], I.p1[1], I.p1[2], bits_f(flags));
    // Culling bounds that hold at every exposure time. Between shutter open and close a body point y (body frame)
    // moves on R(t) y + p(t): R(t) turns about one axis by at most phi (the angle between q0 and q1), p(t) is
    // linear. The world box is the box of the body's corners posed at t = k/8, padded by the largest gap between
    // a circular arc of angle phi/8 and its chord, rad (1 - cos(phi/16)), rad = |y| bound over the box.
    const float *bx = &c->body_box[6 * (size_t)b];
    double ctr[3], half[3];
    for (int k = 0; k < 3; k++) { ctr[k] = 0.5 * (bx[k] + bx[3 + k]); half[k] = 0.5 * (bx[3 + k] - bx[k]); }
    const double halfdiag = std::sqrt(half[0] * half[0] + half[1] * half[1] + half[2] * half[2]);
    const double rad = std::sqrt(ctr[0] * ctr[0] + ctr[1] * ctr[1] + ctr[2] * ctr[2]) + halfdiag;
    double qd = 0, dp = 0;
    for (int k = 0; k < 4; k++) qd += (double)I.q0[k] * I.q1[k];
    for (int k = 0; k < 3; k++) dp += ((double)I.p1[k] - I.p0[k]) * ((double)I.p1[k] - I.p0[k]);
    const double phi = moving ? 2.0 * std::acos(std::min(1.0, std::fabs(qd))) : 0.0;
    dp = std::sqrt(dp);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const int nt = moving ? 9 : 1;
    for (int ti = 0; ti < nt; ti++) {
      const HostPose P = host_pose(I, moving ? ti / 8.0 : 0.0);
      for (int cn = 0; cn < 8; cn++) {
        const double v[3] = {cn & 1 ? bx[3] : bx[0], cn & 2 ? bx[4] : bx[1], cn & 4 ? bx[5] : bx[2]};
        double w[3];
        host_apply(P, v, w);
        for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], w[k]); hi[k] = std::max(hi[k], w[k]); }
      }
    }
    const double scale = std::max({std::fabs(lo[0]), std::fabs(lo[1]), std::fabs(lo[2]), std::fabs(hi[0]),
                                   std::fabs(hi[1]), std::fabs(hi[2]), 1e-3});
    const double pad = 1.01 * rad * (1.0 - std::cos(phi / 16.0)) + 4e-6 * scale + 1e-6;
    r[4] = make_float4((float)(lo[0] - pad), (float)(lo[1] - pad), (float)(lo[2] - pad), 0.f);
    r[5] = make_float4((float)(hi[0] + pad), (float)(hi[1] + pad), (float)(hi[2] + pad), 0.f);
    // Bounding capsule of the box along its longest axis, posed at mid-exposure; every point stays within
    // 2 sin(phi / 4) rad + |p1 - p0| / 2 of its mid-exposure place, which pads the radius.
    const int ax = half[0] >= half[1] ? (half[0] >= half[2] ? 0 : 2) : (half[1] >= half[2] ? 1 : 2);
    double ea[3], eb[3];
    for (int k = 0; k < 3; k++) { ea[k] = ctr[k]; eb[k] = ctr[k]; }
    ea[ax] -= half[ax];
    eb[ax] += half[ax];
    double rr = 0;
    for (int k = 0; k < 3; k++)
      if (k != ax) rr += half[k] * half[k];
    rr = std::sqrt(rr);
    const HostPose Pm = host_pose(I, moving ? 0.5 : 0.0);
    double wa[3], wb[3];
    host_apply(Pm, ea, wa);
    host_apply(Pm, eb, wb);
    const double motion = moving ? 1.01 * (2.0 * std::sin(phi / 4.0) * rad + 0.5 * dp) : 0.0;
    r[6] = make_float4((float)wa[0], (float)wa[1], (float)wa[2], (float)(rr * 1.0001 + motion + 4e-6 * scale + 1e-6));
    r[7] = make_float4((float)wb[0], (float)wb[1], (float)wb[2], 0.f);
    // occlusion capsule of the reference estimator's indirect light, posed at mid-exposure
    if (!(I.occ_r >= 0.0f) || !std::isfinite(I.occ_r) || !(I.rho_eff >= 0.0f && I.rho_eff <= 1.0f))
      return fail(c, VT_ERR_ARG, "set_instances: body %u needs occ_r >= 0 and rho_eff in [0, 1]", b);
    const double oa[3] = {I.occ_a[0], I.occ_a[1], I.occ_a[2]}, ob[3] = {I.occ_b[0], I.occ_b[1], I.occ_b[2]};
    double woa[3], wob[3];
    host_apply(Pm, oa, woa);
    host_apply(Pm, ob, wob);
    double occ_r = I.occ_r;
    if (I.cavity) {
      // the part inside the cavity: s >= s0 along the inward axis (cavityCapsule, gpu/scene_gpu.js:486-503)
      const double ax[3] = {I.cavity_axis[0], I.cavity_axis[1], I.cavity_axis[2]};
      const double al = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
      if (!(al > 1e-12) || !std::isfinite(al) || !std::isfinite(I.cavity_s0) || !(I.cavity_r >= 0.0f))
        return fail(c, VT_ERR_ARG, "set_instances: body %u has a cavity without a finite axis, s0 and r >= 0", b);
      if (I.cavity_r > 0.0f) occ_r = I.cavity_r;
      const auto s_of = [&](const double X[3]) {
        double d = 0;
        for (int k = 0; k < 3; k++) d += (X[k] - I.cavity_origin[k]) * ax[k];
        return d / al;
      };
      const double s0 = I.cavity_s0, sa = s_of(woa), sb = s_of(wob);
      if (sa < s0 && sb < s0) {
        occ_r = 0.0;
      } else if (sa < s0) {
        const double t = (s0 - sa) / (sb - sa);
        for (int k = 0; k < 3; k++) woa[k] += (wob[k] - woa[k]) * t;
      } else if (sb < s0) {
        const double t = (s0 - sb) / (sa - sb);
        for (int k = 0; k < 3; k++) wob[k] += (woa[k] - wob[k]) * t;
      }
    }
    r[8] = make_float4((float)woa[0], (float)woa[1], (float)woa[2], (float)occ_r);
    r[9] = make_float4((float)wob[0], (float)wob[1], (float)wob[2], I.rho_eff);
  }
  // a light that rides on a body lights the static scene differently once that body's pose record changes: the
  // fixed cameras' occlusion caches start over
  if (c->S.emit_body != 0xffffffffu) {
    const float4 *r = &rec[INST_WORDS * (size_t)c->S.emit_body];
    if (!c->emit_rec_set || std::memcmp(c->emit_rec, r, sizeof c->emit_rec) != 0) c->shade_gen++;
    std::memcpy(c->emit_rec, r, sizeof c->emit_rec);
    c->emit_rec_set = true;
  }
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(c->inst.p, rec.data(), rec.size() * sizeof(float4), hipMemcpyHostToDevice));
  return VT_OK;
}

