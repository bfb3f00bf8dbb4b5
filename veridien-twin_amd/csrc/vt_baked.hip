// The lite renderer's radiosity on its baked node table (lite/shaders.js:220-283; lite/renderer.js:214-222,
// 486-517) on gfx950: K cosine gather rays per node, stored as 16-bit hit node ids (0xffff: the ray escapes).
//  * k_baked_occ (gatherOcc): one lane per word of four gather rays. A ray's transmission byte is 255 unless a
//    posed body's capsule comes within its radius of the segment from the node (lifted 1 mm along its normal) to
//    the hit node; then it is the smallest round(255 rho) of the capsules that cut it: the light the ray would
//    have brought is replaced by what the body re-emits, about its mean albedo times the same radiance.
//  * k_baked_jacobi (jacobi): E_ind = pi / K sum_k t_k B(hit_k), B' = (E_dir + hh E_ind) / pi. One wave per node,
//    lane l taking rays l, l + 64, ...: the node's hit ids (2 B each) and bytes come in as contiguous 128-B and
//    64-B stretches, the K gathers of B are in flight at once, and the lanes' sums meet in a fixed butterfly, so a
//    step is deterministic. The table is 3 B per ray and B a 16-B record per node: a step at K = 256 streams
//    768 B of table per node plus the hits' B records, which the L2 holds for the 12,976-node table.
#include "vt_internal.h"

namespace vt {

namespace {

// the closest points of the segments p0-p1 and q0-q1 (parameters s, t) and their distance (segSeg,
// lite/shaders.js:86-104)
__device__ __forceinline__ float seg_seg_dist(v3 p0, v3 p1, v3 q0, v3 q1) {
  const v3 d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
  const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  float s = 0.0f, t = 0.0f;
  if (a <= 1e-14f) {
    t = fminf(fmaxf(f / fmaxf(e, 1e-14f), 0.0f), 1.0f);
  } else {
    const float c = dot(d1, r);
    if (e <= 1e-14f) {
      s = fminf(fmaxf(-c / a, 0.0f), 1.0f);
    } else {
      const float b = dot(d1, d2);
      const float den = a * e - b * b;
      if (den > 1e-14f) s = fminf(fmaxf((b * f - c * e) / den, 0.0f), 1.0f);
      t = (b * s + f) / e;
      if (t < 0.0f) {
        t = 0.0f;
        s = fminf(fmaxf(-c / a, 0.0f), 1.0f);
      } else if (t > 1.0f) {
        t = 1.0f;
        s = fminf(fmaxf((b - c) / a, 0.0f), 1.0f);
      }
    }
  }
  const v3 w = (p0 + d1 * s) - (q0 + d2 * t);
  return sqrtf(dot(w, w));
}

__global__ __launch_bounds__(256) void k_baked_occ(BakedDev B) {
  const uint64_t words = (uint64_t)B.nodes * B.K / 4;
  for (uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x; w < words; w += (uint64_t)gridDim.x * 256) {
    const uint32_t node = (uint32_t)(w * 4 / B.K);
    const v3 p = ld3(B.node_a[node]) + ld3(B.node_n[node]) * 1e-3f;
    // four consecutive 16-bit ids: one 8-B load (K is a multiple of 4, so a word never straddles two nodes)
    const uint2 hw = reinterpret_cast<const uint2 *>(B.gather)[w];
    uint32_t word = 0u;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t hit = ((q < 2 ? hw.x : hw.y) >> (16 * (q & 1))) & 0xffffu;
      uint32_t t = 255u;
      if (hit != 0xffffu) {
        const v3 e = ld3(B.node_a[hit]);
        for (uint32_t i = 0; i < B.n_caps; i++) {
          const float4 a = B.caps[2 * i], b = B.caps[2 * i + 1];
          if (seg_seg_dist(p, e, ld3(a), ld3(b)) < a.w)
            t = min(t, (uint32_t)rintf(255.0f * fminf(fmaxf(b.w, 0.0f), 1.0f)));
        }
      }
      word |= t << (8 * q);
    }
    B.occ[w] = word;
  }
}

__global__ __launch_bounds__(256) void k_baked_jacobi(BakedDev B, const float4 *b_in, float4 *b_out) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t i = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (i >= B.nodes) return;     // uniform over the wave
  const uint16_t *row = B.gather + (size_t)i * B.K;
  const uint8_t *orow = reinterpret_cast<const uint8_t *>(B.occ) + (size_t)i * B.K;
  v3 acc = mk(0.f, 0.f, 0.f);
  for (uint32_t k = lane; k < B.K; k += 64u) {
    const uint32_t hit = row[k];
    if (hit == 0xffffu) continue;
    const float t = (float)orow[k] * (1.0f / 255.0f);
    acc = acc + ld3(b_in[hit]) * t;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    acc.x += __shfl_xor(acc.x, off, 64);
    acc.y += __shfl_xor(acc.y, off, 64);
    acc.z += __shfl_xor(acc.z, off, 64);
  }
  if (lane) return;
  const v3 E = acc * (PI_F / (float)B.K);
  const uint32_t mesh = __float_as_uint(B.node_n[i].w);
  const v3 hh = ld3(B.hh[mesh]);
  B.e_ind[i] = make_float4(E.x, E.y, E.z, 0.0f);
  const v3 Bo = (ld3(B.e_dir[i]) + hh * E) * INV_PI_F;
  b_out[i] = make_float4(Bo.x, Bo.y, Bo.z, 0.0f);
}

}  // namespace

hipError_t launch_baked_occ(const BakedDev &B, hipStream_t s) {
  const uint64_t words = (uint64_t)B.nodes * B.K / 4;
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((words + 255) / 256, 65536));
  hipLaunchKernelGGL(k_baked_occ, dim3(grid), dim3(256), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_baked_jacobi(const BakedDev &B, const float4 *b_in, float4 *b_out, hipStream_t s) {
  hipLaunchKernelGGL(k_baked_jacobi, dim3((B.nodes + 3) / 4), dim3(256), 0, s, B, b_in, b_out);
  return hipGetLastError();
}

}  // namespace vt
