// The node radiosity cache on gfx950 (the reference's gather table, node direct light and albedo tables,
// gpu/wgsl/kernels.wgsl.js:283-406, gpu/bsdf.js:210-254):
//  * k_albedo: one wave per material, every lane a 16-sample slice of the 1024-sample quadrature of each bin,
//    summed by a butterfly; lane 0 writes the bins and the hemispherical albedo.
//  * k_gather_build / k_gather_remap: one lane per gather ray (node-major), persistent over the rays, the static
//    BVH walked with the LDS short stack of the trace kernels.
//  * k_node_direct: four lanes per active node sharing its node_samples emitter samples with shadow rays.
#include "vt_internal.h"
#include "vt_nodes.h"

namespace vt {

namespace {

constexpr uint32_t ALB_SAMPLES = 1024;    // scene_gpu.js:197

__device__ __forceinline__ float radinv2(uint32_t i) { return (float)__brev(i) * (1.0f / 4294967296.0f); }

__device__ __forceinline__ float wave_sum(float v) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// directional albedo tables (albedoTable, gpu/bsdf.js:210-254): specular part by VNDF sampling with a Hammersley
// set (coated: the material's own sampler), diffuse part analytic; kind 3 (Lambert) is rho, all diffuse
__global__ __launch_bounds__(64) void k_albedo(NodeLaunch L) {
  const uint32_t m = blockIdx.x, lane = threadIdx.x;
  if (m >= L.n_mat) return;
  const Mat M = load_mat(L.S.mats, m);
  const float fbar_c = L.S.mats[4 * (size_t)m + 3].x;
  const v3 n = mk(0.f, 0.f, 1.f);
  v3 hh = mk(0.f, 0.f, 0.f);
  for (uint32_t k = 0; k < ALB_BINS; k++) {
    const float mu = ((float)k + 0.5f) / (float)ALB_BINS;
    const v3 v = mk(sqrtf(1.0f - mu * mu), 0.0f, mu);
    v3 acc = mk(0.f, 0.f, 0.f);
    if (M.kind == 2) {
      for (uint32_t i = lane; i < ALB_SAMPLES; i += 64) {
        v3 l;
        const float pdf = sample_bsdf(M, n, v, ((float)i + 0.5f) / (float)ALB_SAMPLES, radinv2(i), radinv2(i ^ 0x5555u), l);
        if (!(pdf > 0.0f)) continue;
        acc = acc + eval_bsdf(M, n, l, v) * (l.z / pdf);
      }
    } else if (M.kind != 3) {
      for (uint32_t i = lane; i < ALB_SAMPLES; i += 64) {
        const v3 h = sample_vndf(v, M.alpha, ((float)i + 0.5f) / (float)ALB_SAMPLES, radinv2(i));
        const float vh = dot(v, h);
        const v3 l = h * (2.0f * vh) - v;
        if (l.z <= 0.0f) continue;
        const float w = ggx_g2(l.z, mu, M.alpha) / ggx_g1(mu, M.alpha);
        const v3 F = M.kind == 1 ? schlick3(M.c, vh) : [&] { const float f = fresnel_dielectric(vh, M.ior); return mk(f, f, f); }();
        acc = acc + F * w;
      }
    }
    acc = mk(wave_sum(acc.x), wave_sum(acc.y), wave_sum(acc.z)) * (1.0f / (float)ALB_SAMPLES);
    v3 diff = mk(0.f, 0.f, 0.f);
    if (M.kind == 3) {
      acc = mk(0.f, 0.f, 0.f);
      diff = M.c;
    } else if (M.kind == 0) {
      diff = M.c * ((1.0f - fresnel_dielectric(mu, M.ior)) * (1.0f - fbar_c) / (1.0f - M.fbar));
    }
    if (lane == 0) {
      float4 *b = L.alb + ((size_t)m * ALB_BINS + k) * 2;
      b[0] = make_float4(acc.x, acc.y, acc.z, 0.0f);
      b[1] = make_float4(diff.x, diff.y, diff.z, 0.0f);
    }
    hh = hh + (acc + diff) * (2.0f * mu / (float)ALB_BINS);
  }
  if (lane == 0) L.hh[m] = make_float4(hh.x, hh.y, hh.z, 0.0f);
}

// the vertex of the hit triangle a gather ray lands on: the largest barycentric, if the ray meets the triangle's
// front (interpolated vertex normals against the ray); MISS for an escape or a back face (gatherBuild :291-306)
__device__ __forceinline__ uint32_t gather_hit(const NodeLaunch &L, uint32_t i, uint32_t j, Stack &st, v3 &bb,
                                               uint4 &vi) {
  const v3 p = ld3(L.npos[i]), n = ld3(L.nnorm[i]);
  const v3 dir = gather_dir(i, j, L.K, n);
  Hit h;
  if (!traverse<false>(L.S, L.S.root, p + n * EPS_OFF, dir, 0.0f, INFINITY, h, st)) return MISS;
  bb = mk(1.0f - h.u - h.v, h.u, h.v);
  vi = L.S.tvid[h.tri];
  const v3 nv = ld3(L.nnorm[vi.x]) * bb.x + ld3(L.nnorm[vi.y]) * bb.y + ld3(L.nnorm[vi.z]) * bb.z;
  if (!(dot(nv, dir) < 0.0f)) return MISS;
  if (bb.x >= bb.y && bb.x >= bb.z) return vi.x;
  return bb.y >= bb.z ? vi.y : vi.z;
}

#define NODE_STACK(L)                                                                                          \
  __shared__ uint32_t s_node[STACK_LDS * 256];                                                                 \
  __shared__ float s_t[STACK_LDS * 256];                                                                       \
  Stack st{(lds_u32 *)s_node, (lds_f32 *)s_t, L.spill, threadIdx.x, 256, STACK_LDS, STACK_SPILL, L.spill_stride,                     \
           (uint64_t)blockIdx.x * 256 + threadIdx.x, L.overflow}

__global__ __launch_bounds__(256) void k_gather_build(NodeLaunch L) {
  NODE_STACK(L);
  const uint64_t rays = (uint64_t)L.n_nodes * L.K;
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < rays; r += (uint64_t)gridDim.x * 256) {
    const uint32_t i = (uint32_t)(r / L.K), j = (uint32_t)(r - (uint64_t)i * L.K);
    uint32_t v = MISS;
    if (L.npos[i].w > 0.5f) {
      v3 bb;
      uint4 vi;
      v = gather_hit(L, i, j, st, bb, vi);
    }
    L.gtable[r] = v;
  }
}

// escapes per node (gatherCount :326-333); a vertex that is no static node counts every ray as an escape
__global__ __launch_bounds__(256) void k_gather_count(NodeLaunch L) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= L.n_nodes) return;
  uint32_t misses = 0;
  for (uint32_t j = 0; j < L.K; j++) misses += L.gtable[(size_t)i * L.K + j] == MISS ? 1u : 0u;
  L.miss[i] = L.npos[i].w > 0.5f ? (float)misses : (float)L.K;
}

__device__ __forceinline__ bool node_active(const NodeLaunch &L, uint32_t v) { return L.miss[v] <= 0.5f * (float)L.K; }

// entries at an inactive node move to the hit triangle's highest-barycentric active vertex (gatherRemap :308-324)
__global__ __launch_bounds__(256) void k_gather_remap(NodeLaunch L) {
  NODE_STACK(L);
  const uint64_t rays = (uint64_t)L.n_nodes * L.K;
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < rays; r += (uint64_t)gridDim.x * 256) {
    const uint32_t v0 = L.gtable[r];
    if (v0 == MISS || node_active(L, v0)) continue;
    const uint32_t i = (uint32_t)(r / L.K), j = (uint32_t)(r - (uint64_t)i * L.K);
    const v3 p = ld3(L.npos[i]), n = ld3(L.nnorm[i]);
    const v3 dir = gather_dir(i, j, L.K, n);
    Hit h;
    if (!traverse<false>(L.S, L.S.root, p + n * EPS_OFF, dir, 0.0f, INFINITY, h, st)) continue;
    const v3 bb = mk(1.0f - h.u - h.v, h.u, h.v);
    const uint4 vi = L.S.tvid[h.tri];
    uint32_t best = v0;
    float bw = -1.0f;
    if (node_active(L, vi.x) && bb.x > bw) { best = vi.x; bw = bb.x; }
    if (node_active(L, vi.y) && bb.y > bw) { best = vi.y; bw = bb.y; }
    if (node_active(L, vi.z) && bb.z > bw) { best = vi.z; bw = bb.z; }
    L.gtable[r] = best;
  }
}

// the first reflection of the direct light at each active node, B_dir = sum_s E_s A(mu_s) / samples over
// node_samples fixed emitter samples with shadow rays through the static scene and the posed bodies at
// mid-exposure; the visibility code of the pixels' shadow culling; the capsule visibility (nodeDirect :348-406)
// Four lanes per node, lane q taking samples q, q + 4, ...: four times the shadow rays in flight of a lane per node;
// the group's sums meet in a fixed butterfly and its first lane writes the node.
constexpr uint32_t NODE_GROUP = 4;

// The incremental pass (kernels.wgsl.js:357-366): with a moved-body mask, a node keeps its direct light unless
// the segment from it to the emitter's centre passes within (capsule radius + emitter radius + 1 mm) of a moved
// body's bounding capsule, where the body is now or where it was at the last direct pass. Every shadow segment
// from the node to a point of the emitter lies within the emitter radius of that segment, so a node kept this
// way sends no shadow ray near the moved bodies, and its samples come out as a full pass would compute them.
__device__ __forceinline__ bool near_moved(const NodeLaunch &L, v3 p) {
  const float4 ec = L.S.radu[1];
  const v3 dE = emit_centre(L.S, emitter_pose(L.S, 0.5f)) - p;
  const float rr = ec.w + 1e-3f;
  for (uint32_t b = 0; b < L.S.inst_count; b++) {
    if (!((L.moved >> b) & 1u)) continue;
    const float4 ca = L.S.inst[INST_WORDS * (size_t)b + 6], cb = L.S.inst[INST_WORDS * (size_t)b + 7];
    const float4 pa = L.pcaps[2 * b], pb = L.pcaps[2 * b + 1];
    if (seg_near_capsule(p, dE, 0.0f, 1.0f, ld3(ca), ld3(cb), ca.w + rr) ||
        seg_near_capsule(p, dE, 0.0f, 1.0f, ld3(pa), ld3(pb), pa.w + rr))
      return true;
  }
  return false;
}

template <bool INST>
__global__ __launch_bounds__(256) void k_node_direct(NodeLaunch L) {
  NODE_STACK(L);
  // a group's lanes are consecutive and share every loop trip: lanes is a multiple of NODE_GROUP, strides of 256
  const uint64_t lanes = (uint64_t)L.n_active * NODE_GROUP;
  for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < lanes; t += (uint64_t)gridDim.x * 256) {
    const uint32_t k = (uint32_t)(t / NODE_GROUP), q = (uint32_t)(t % NODE_GROUP);
    const uint32_t i = L.active[k];
    const v3 p = ld3(L.npos[i]), n = ld3(L.nnorm[i]);
    // the group's four lanes take the same branch, so the butterfly below stays within live lanes
    if (INST && L.moved && !near_moved(L, p)) continue;
    const uint32_t m = L.nmat[i];
    const uint32_t hs = pcg(i * 2246822519u + 3u);
    const float s0 = u01(pcg(hs)), s1 = u01(pcg(hs + 1u)), s2 = u01(pcg(hs + 2u));
    v3 B = mk(0.f, 0.f, 0.f);
    uint32_t nlit = 0, nvis = 0;
    const v3 o = p + n * EPS_OFF;
    const EmitPose E = emitter_pose(L.S, 0.5f);      // mid-exposure, as the shadow rays
    for (uint32_t s = q; s < L.node_samples && L.S.emit_count > 0; s += NODE_GROUP) {
      v3 pl, nle;
      const uint32_t e = sample_emitter(L.S, E, kr(s, KL0, s0), kr(s, KL1, s1), kr(s, KL2, s2), pl, nle);
      v3 wi = pl - p;
      const float d2 = dot(wi, wi), dist = sqrtf(d2);
      wi = wi * (1.0f / dist);
      const float cosE = -dot(wi, nle);
      const float cosN = dot(n, wi);
      if (cosE <= 0.0f || cosN <= 0.0f) continue;
      const uint32_t lid = __float_as_uint(L.S.emit[4 * e + 2].w);
      const v3 Le = ld3(L.S.lights[lid]) * light_profile(L.S, lid, emit_local(E, -wi));
      if (maxc(Le) <= 0.0f) continue;
      nlit++;
      Hit sh;
      if (trace_scene<true, INST>(L.S, o, wi, 0.0f, dist - 2e-5f, 0.5f, L.S.shadow_skip, sh, st)) continue;
      nvis++;
      const v3 ev = Le * (cosE * cosN * L.S.emit_area / d2);
      B = B + ev * albedo_at(L.alb, ALB_BINS, m, cosN);
    }
#pragma unroll
    for (int off = NODE_GROUP / 2; off >= 1; off >>= 1) {
      B.x += __shfl_xor(B.x, off, 64);
      B.y += __shfl_xor(B.y, off, 64);
      B.z += __shfl_xor(B.z, off, 64);
      nlit += __shfl_xor(nlit, off, 64);
      nvis += __shfl_xor(nvis, off, 64);
    }
    if (q) continue;
    const float inv = 1.0f / (float)max(L.node_samples, 1u);
    B = B * inv;
    const float vis = nlit == 0 ? -1.0f : (nvis == nlit ? 1.0f : (nvis == 0 ? 0.0f : 0.5f));
    L.state[4 * (size_t)i + 3] = make_float4(B.x, B.y, B.z, vis);
    const float4 old = L.state[4 * (size_t)i];
    L.state[4 * (size_t)i] = make_float4(old.x, old.y, old.z, INST ? capsule_vis(L.S, p, n) : 1.0f);
  }
}

// nodeSummary's eAvg (gpu/renderer.js:824-852) on the device: area-weighted E_free V + B_dir / max(A_hh, 1e-3)
// over the active nodes in FP64. Stage 1: block b sums the b-th contiguous share of the active list (a fixed
// partition, so the result does not depend on scheduling); stage 2: one block adds the EAVG_BLOCKS partials in
// order and writes radu[0] = eAvg + bits(slot).
constexpr uint32_t EAVG_BLOCKS = 256;

__device__ __forceinline__ void block_sum4(double v[4], double (*sh)[256]) {
  for (int q = 0; q < 4; q++) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (uint32_t st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 4; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_eavg_part(NodeLaunch L, const float *area, double *part) {
  __shared__ double sh[4][256];
  const uint32_t per = (L.n_active + EAVG_BLOCKS - 1) / EAVG_BLOCKS;
  const uint32_t k0 = blockIdx.x * per, k1 = min(L.n_active, k0 + per);
  double v[4] = {0, 0, 0, 0};
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += 256) {
    const uint32_t i = L.active[k];
    const double w = area[i];
    const float4 s0 = L.state[4 * (size_t)i], s3 = L.state[4 * (size_t)i + 3];
    const float4 h = L.hh[L.nmat[i]];
    v[0] += w;
    v[1] += w * ((double)s0.x * s0.w + (double)s3.x / fmax((double)h.x, 1e-3));
    v[2] += w * ((double)s0.y * s0.w + (double)s3.y / fmax((double)h.y, 1e-3));
    v[3] += w * ((double)s0.z * s0.w + (double)s3.z / fmax((double)h.z, 1e-3));
  }
  block_sum4(v, sh);
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_eavg(const double *part, float4 *radu, uint32_t slot) {
  __shared__ double sh[4][256];
  double v[4];
  for (int q = 0; q < 4; q++) v[q] = threadIdx.x < EAVG_BLOCKS ? part[4 * threadIdx.x + q] : 0.0;
  block_sum4(v, sh);
  if (threadIdx.x == 0) {
    const double A = sh[0][0];
    radu[0] = A > 0 ? make_float4((float)(sh[1][0] / A), (float)(sh[2][0] / A), (float)(sh[3][0] / A), __uint_as_float(slot))
                    : make_float4(0.f, 0.f, 0.f, __uint_as_float(slot));
  }
}

__global__ void k_fill_f4(float4 *p, uint32_t n, float4 v) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

uint32_t grid_of(const NodeLaunch &L, uint64_t lanes) {
  const uint64_t g = (lanes + 255) / 256;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, L.max_blocks));
}

}  // namespace

hipError_t launch_eavg(const NodeLaunch &L, const float *area, float4 *radu, uint32_t slot, double *part,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_eavg_part, dim3(EAVG_BLOCKS), dim3(256), 0, s, L, area, part);
  hipLaunchKernelGGL(k_eavg, dim3(1), dim3(256), 0, s, (const double *)part, radu, slot);
  return hipGetLastError();
}
hipError_t launch_fill_f4(float4 *p, uint32_t n, float4 v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f4, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}
hipError_t launch_albedo_tables(const NodeLaunch &L, hipStream_t s) {
  if (L.n_mat == 0) return hipSuccess;
  hipLaunchKernelGGL(k_albedo, dim3(L.n_mat), dim3(64), 0, s, L);
  return hipGetLastError();
}
hipError_t launch_gather_build(const NodeLaunch &L, hipStream_t s) {
  if (L.n_nodes == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_build, dim3(grid_of(L, (uint64_t)L.n_nodes * L.K)), dim3(256), 0, s, L);
  return hipGetLastError();
}
hipError_t launch_gather_count(const NodeLaunch &L, hipStream_t s) {
  if (L.n_nodes == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_count, dim3((L.n_nodes + 255) / 256), dim3(256), 0, s, L);
  return hipGetLastError();
}
hipError_t launch_gather_remap(const NodeLaunch &L, hipStream_t s) {
  if (L.n_nodes == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_remap, dim3(grid_of(L, (uint64_t)L.n_nodes * L.K)), dim3(256), 0, s, L);
  return hipGetLastError();
}
hipError_t launch_node_direct(const NodeLaunch &L, hipStream_t s) {
  if (L.n_active == 0) return hipSuccess;
  if (L.S.inst_count)
    hipLaunchKernelGGL(k_node_direct<true>, dim3(grid_of(L, (uint64_t)L.n_active * NODE_GROUP)), dim3(256), 0, s, L);
  else
    hipLaunchKernelGGL(k_node_direct<false>, dim3(grid_of(L, (uint64_t)L.n_active * NODE_GROUP)), dim3(256), 0, s, L);
  return hipGetLastError();
}

}  // namespace vt
