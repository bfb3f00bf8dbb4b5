// Device helpers of the node radiosity cache and of the reference camera estimator that reads it
// (gpu/wgsl/shade.wgsl.js:57-243, gpu/wgsl/kernels.wgsl.js:271-281, gpu/wgsl/bsdf.wgsl.js:157-173).
#pragma once
#include "vt_device.h"

namespace vt {

constexpr uint32_t F_STATIC = 1u;
constexpr uint32_t OPT_NO_INDIRECT = 4u, OPT_DIRECT_ONLY = 8u, OPT_NO_CAPSULE = 32u, OPT_NORMALS = 64u,
                   OPT_NO_MIS = 128u, OPT_REFINE_SHADOWS = 1024u, OPT_SHADOW_CACHE = 8192u,
                   OPT_REFERENCE = 1u << 16,
                   OPT_CULL_SHADOWS = 1u << 17;

// gather-table ray j of node i (normal n): K cosine-distributed rays, a Fibonacci spiral on the disk rotated per
// node (gatherDir, kernels.wgsl.js:271-278)
__device__ __forceinline__ v3 gather_dir(uint32_t i, uint32_t j, uint32_t K, v3 n) {
  v3 t, b;
  basis(n, t, b);
  const float rot = u01(pcg(i * 2654435761u + 17u));
  const float rr = sqrtf(((float)j + 0.5f) / (float)K);
  const float ph = 2.0f * PI_F * kr(j, 0.6180339887498949f, rot);
  float sn, cs;
  sincosf(ph, &sn, &cs);
  return normalize(t * (rr * cs) + b * (rr * sn) + n * sqrtf(fmaxf(0.0f, 1.0f - rr * rr)));
}

// directional albedo of material m at view cosine mu: specular + diffuse parts, linear in the 32 bins
// (albedoSpec + albedoDiff, bsdf.wgsl.js:157-173); the table holds per bin [spec.rgb, -][diff.rgb, -]
__device__ __forceinline__ v3 albedo_at(const float4 *alb, uint32_t bins, uint32_t m, float mu) {
  const float x = fminf(fmaxf(mu * (float)bins - 0.5f, 0.0f), (float)(bins - 1));
  const uint32_t i = (uint32_t)floorf(x);
  const uint32_t j = min(bins - 1, i + 1);
  const float f = x - (float)i;
  const float4 *b = alb + (size_t)m * bins * 2;
  const v3 s0 = ld3(b[2 * i]), s1 = ld3(b[2 * j]), d0 = ld3(b[2 * i + 1]), d1 = ld3(b[2 * j + 1]);
  return (s0 * (1.0f - f) + s1 * f) + (d0 * (1.0f - f) + d1 * f);
}

// cosine-weighted fraction of the hemisphere above (p, n) hidden by a thin cylinder of radius r on a-b, by the
// closed-form integral over the axis angle, clipped to the horizon (capsuleOcc, shade.wgsl.js:57-81)
__device__ __forceinline__ float capsule_occ(v3 p, v3 n, v3 a, v3 bb, float r) {
  const v3 ab = bb - a;
  const float L = sqrtf(dot(ab, ab));
  if (L < 1e-9f || r <= 0.0f) return 0.0f;
  const v3 e = ab * (1.0f / L);
  const v3 ap = p - a;
  const float s0 = dot(ap, e);
  const v3 perp = ap - e * s0;
  const float lp = sqrtf(dot(perp, perp));
  const float h = sqrtf(lp * lp + r * r);
  v3 ep = perp * (-1.0f / fmaxf(lp, 1e-12f));
  if (lp < 1e-12f) {
    v3 t, b;
    basis(e, t, b);
    ep = normalize(cross(e, t));
  }
  const float np = dot(n, ep), na = dot(n, e);
  float f1 = atanf((0.0f - s0) / h);
  float f2 = atanf((L - s0) / h);
  const float f0 = atan2f(na, np);
  f1 = fmaxf(f1, f0 - 0.5f * PI_F);
  f2 = fminf(f2, f0 + 0.5f * PI_F);
  if (f2 <= f1) return 0.0f;
  const float s1 = sinf(f1), s2 = sinf(f2);
  const float I = np * (0.5f * (f2 - f1) + 0.25f * (sinf(2.0f * f2) - sinf(2.0f * f1))) + na * 0.5f * (s2 * s2 - s1 * s1);
  return fminf(fmaxf(INV_PI_F * (2.0f * r / h) * I, 0.0f), 1.0f);
}

// 1 - sum over the visible bodies of (1 - rho_eff) O(occlusion capsule): the indirect-light visibility of a
// static point (capsuleVis, shade.wgsl.js:83-94)
__device__ __forceinline__ float capsule_vis(const SceneDev &S, v3 p, v3 n) {
  float v = 1.0f;
  for (uint32_t i = 0; i < S.inst_count; i++) {
    const float4 *I = S.inst + INST_WORDS * (size_t)i;
    if (__float_as_uint(I[3].w) & INST_HIDDEN) continue;
    const float4 a = I[8], b = I[9];
    v -= capsule_occ(p, n, ld3(a), ld3(b), a.w) * (1.0f - b.w);
  }
  return fminf(fmaxf(v, 0.0f), 1.0f);
}

// barycentric weights renormalised over the triangle's solved nodes (nodeWeights, shade.wgsl.js:214-221)
__device__ __forceinline__ v3 node_weights(const SceneDev &S, uint4 vi, v3 b, uint32_t slot) {
  const v3 w = mk(b.x * S.nstate[4 * (size_t)vi.x + slot].w, b.y * S.nstate[4 * (size_t)vi.y + slot].w,
                  b.z * S.nstate[4 * (size_t)vi.z + slot].w);
  const float sw = w.x + w.y + w.z;
  if (sw <= 1e-6f) return b;
  return w * (1.0f / sw);
}
// state word k of the three nodes, interpolated with the renormalised weights (nodeEfree k = 0, nodeLout k = slot)
__device__ __forceinline__ v3 node_interp(const SceneDev &S, uint4 vi, v3 b, uint32_t slot, uint32_t k) {
  const v3 w = node_weights(S, vi, b, slot);
  return ld3(S.nstate[4 * (size_t)vi.x + k]) * w.x + ld3(S.nstate[4 * (size_t)vi.y + k]) * w.y +
         ld3(S.nstate[4 * (size_t)vi.z + k]) * w.z;
}

}  // namespace vt
