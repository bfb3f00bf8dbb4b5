// Device-side building blocks of the camera path for gfx950: vector helpers, the hash streams and sample
// pattern of gpu/wgsl/common.wgsl.js:94-113 and kernels.wgsl.js:34-53, the FP32 KB4 inverse of
// gpu/wgsl/lens.wgsl.js:9-29, the BSDF of gpu/wgsl/bsdf.wgsl.js (plus the Lambertian kind), emitter sampling
// of gpu/wgsl/shade.wgsl.js:31-47, and the four-wide BVH traversal with a watertight triangle test.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vt {

constexpr float PI_F = 3.14159265358979f;
constexpr float INV_PI_F = 0.318309886183791f;
constexpr uint32_t MISS = 0xffffffffu;
constexpr float EPS_OFF = 2e-6f;      // ray origin offset along the geometric normal (m), common.wgsl.js:33
constexpr uint32_t F_EMITTER = 2u;
constexpr uint32_t F_TEXTURED = 128u;   // the mesh samples the print mask (gpu/wgsl/common.wgsl.js:42)
constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t EMPTY_REF = 0xffffffffu;

// ------------------------------------------------------------------ small vector type
struct v3 {
  float x, y, z;
};
__device__ __forceinline__ v3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ v3 operator*(float s, v3 a) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ v3 normalize(v3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ float maxc(v3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
__device__ __forceinline__ v3 ld3(const float4 &f) { return {f.x, f.y, f.z}; }

// ------------------------------------------------------------------ hashes and the sample pattern
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
  uint32_t s = v * 747796405u + 2891336453u;
  uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (w >> 22u) ^ w;
}
__device__ __forceinline__ uint32_t hash3(uint32_t a, uint32_t b, uint32_t c) { return pcg(a ^ pcg(b ^ pcg(c))); }
__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8u) * (1.0f / 16777216.0f); }
// fract(shift + f32(i) * a), product and sum each rounded: the WGSL sequence, never fused
__device__ __forceinline__ float kr(uint32_t i, float a, float shift) {
#pragma clang fp contract(off)
  float p = (float)i * a;
  float s = shift + p;
  return s - floorf(s);
}

constexpr float R2A = 0.7548776662466927f, R2B = 0.5698402909980532f;
constexpr float KL0 = 0.41421356237f, KL1 = 0.73205080757f, KL2 = 0.23606797750f;
constexpr float KB0 = 0.64575131106f, KB1 = 0.31662479036f, KB2 = 0.60555127546f;

// ------------------------------------------------------------------ cameras
struct DevCam {
  float R[9];
  float pos[3];
  uint32_t mode;
  float lens[9];                 // fx fy cx cy k1..k4 alpha
  uint32_t width, height, spp, seed;
  uint32_t cam_index, nl, depth, tiles_x;
  uint32_t tiles_y, nspp, spp_begin, accumulate;
  uint64_t item_begin, item_count;   // in items of this launch's queue (multiples of 64)
  float4 *acc;                       // per-pixel radiance sum (rgb), sample count (a)
  const float4 *adj;                 // adjoint image dLoss/dI (rgb) for the backward pass, else null
  float scale;                       // output scale (the image is acc * scale / spp)
  float shutter0, shutter1;          // exposure interval on the pose timeline (CamU.expo.xy)
  uint32_t skip;                     // bodies the camera's own rays ignore (bit per body)
  uint32_t opts;                     // option bits (VT_OPT_*)
  // the reference estimator's refinement pass (camRefine): extra sub-samples, their light samples, and the
  // per-pixel keys of the first pass {mesh of sub-sample 0, lowest and highest mesh over the sub-samples,
  // 0x10000 posed body | 0x20000 shadow ray traced}; null keys: no refinement this frame
  uint32_t refine, nl_refine;
  uint4 *keys;
  // the static first-hit cache of a fixed camera (staticFirstHit, kernels.wgsl.js:58-67): per sub-sample layer i
  // and pixel, (t, triangle, u, v) of the static scene's first hit; t > 0 marks an entry written (a miss keeps
  // t = 1e30, triangle MISS). Null: no cache.
  float4 *hcache;
  // the static occlusion cache of a still light (shadowCache): per sub-sample layer and pixel, bit j = light sample
  // j hidden by the static scene, bit 31 = written. Null: no cache.
  uint32_t *scache;
};

// KB4 inverse, FP32 with a fixed 8 Newton steps (lens.wgsl.js:9-29)
__device__ __forceinline__ v3 kb4_unproject(const float *L, float u, float v) {
  const float fx = L[0], fy = L[1], cx = L[2], cy = L[3];
  const float k1 = L[4], k2 = L[5], k3 = L[6], k4 = L[7], al = L[8];
  float yd = (v - cy) / fy;
  float xd = (u - cx) / fx - al * yd;
  float thd = sqrtf(xd * xd + yd * yd);
  if (thd < 1e-20f) return mk(0.f, 0.f, 1.f);
  float th = thd;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    float t2 = th * th;
    float f = th * (1.0f + t2 * (k1 + t2 * (k2 + t2 * (k3 + t2 * k4)))) - thd;
    float fp = 1.0f + t2 * (3.0f * k1 + t2 * (5.0f * k2 + t2 * (7.0f * k3 + t2 * 9.0f * k4)));
    th = th - f / fp;
  }
  float sn, cs;
  sincosf(th, &sn, &cs);
  float s = sn / thd;
  return mk(xd * s, yd * s, cs);
}

struct PixSeeds {
  uint32_t hs;
  float s[9];
};
__device__ __forceinline__ void pix_seeds(const DevCam &C, uint32_t x, uint32_t y, PixSeeds &P) {
  P.hs = hash3(x, y, C.seed * 64u + C.cam_index);
#pragma unroll
  for (int k = 0; k < 9; k++) P.s[k] = u01(pcg(P.hs + (uint32_t)k));
}

constexpr float GOLD_F = 0.6180339887498949f;
// the sub-sample's time in the exposure (kernels.wgsl.js:44): every ray of its path sees the bodies at that time
__device__ __forceinline__ float sample_time(const DevCam &C, const PixSeeds &P, uint32_t i) {
  const float k = kr(i, GOLD_F, P.s[2]);
  return C.shutter0 + (C.shutter1 - C.shutter0) * k;
}

__device__ __forceinline__ v3 camera_dir(const DevCam &C, const PixSeeds &P, uint32_t x, uint32_t y, uint32_t i) {
  const float u = ((float)x + kr(i, R2A, P.s[0])) - 0.5f;
  const float v = ((float)y + kr(i, R2B, P.s[1])) - 0.5f;
  v3 dc;
  if (C.mode == 0) dc = kb4_unproject(C.lens, u, v);
  else dc = normalize(mk((u - C.lens[2]) / C.lens[0], (v - C.lens[3]) / C.lens[1], 1.0f));
  v3 d = mk(C.R[0], C.R[1], C.R[2]) * dc.x + mk(C.R[3], C.R[4], C.R[5]) * dc.y + mk(C.R[6], C.R[7], C.R[8]) * dc.z;
  return normalize(d);
}

// ------------------------------------------------------------------ BSDF (bsdf.wgsl.js; kind 3 = Lambert)
struct Mat {
  v3 c;
  float alpha, ior, fbar;
  int kind;
  v3 tau;
  float alpha_c;
};
__device__ __forceinline__ Mat load_mat(const float4 *mats, uint32_t m) {
  float4 a = mats[4 * m], b = mats[4 * m + 1], c = mats[4 * m + 2];
  Mat M;
  M.c = mk(a.x, a.y, a.z);
  M.alpha = a.w;
  M.ior = b.x;
  M.kind = (int)b.y;
  M.fbar = b.z;
  M.tau = mk(c.x, c.y, c.z);
  M.alpha_c = c.w;
  return M;
}
__device__ __forceinline__ float fresnel_dielectric(float cosi, float eta) {
  float c = fminf(fmaxf(cosi, 0.0f), 1.0f);
  float s2 = (1.0f - c * c) / (eta * eta);
  if (s2 >= 1.0f) return 1.0f;
  float ct = sqrtf(1.0f - s2);
  float rs = (c - eta * ct) / (c + eta * ct);
  float rp = (eta * c - ct) / (eta * c + ct);
  return 0.5f * (rs * rs + rp * rp);
}
__device__ __forceinline__ float pow5(float x) { float x2 = x * x; return x2 * x2 * x; }
__device__ __forceinline__ v3 schlick3(v3 f0, float c) {
  float m = pow5(1.0f - fminf(fmaxf(c, 0.0f), 1.0f));
  return f0 + (mk(1.f, 1.f, 1.f) - f0) * m;
}
// GGX D = a^2 / (pi (cos^2 (a^2 - 1) + 1)^2) written as a^2 / (pi (sin^2 + cos^2 a^2)^2) with sin^2 = |n x h|^2:
// the same function for unit n and h, without the cancellation of 1 - cos^2 at the lobe's peak, where the
// smoothest materials (a^2 ~ 6e-6) would lose two digits in binary32
__device__ __forceinline__ float ggx_d(v3 n, v3 h, float nh, float a) {
  if (nh <= 0.0f) return 0.0f;
  const v3 x = cross(n, h);
  const float s2 = dot(x, x), a2 = a * a;
  const float d = s2 + nh * nh * a2;
  return a2 / (PI_F * d * d);
}
__device__ __forceinline__ float ggx_lambda(float c, float a) {
  float c2 = fmaxf(c * c, 1e-12f);
  float t2 = (1.0f - c2) / c2;
  return 0.5f * (-1.0f + sqrtf(1.0f + a * a * t2));
}
__device__ __forceinline__ float ggx_g1(float c, float a) { return 1.0f / (1.0f + ggx_lambda(c, a)); }
__device__ __forceinline__ float ggx_g2(float nl, float nv, float a) {
  return 1.0f / (1.0f + ggx_lambda(nl, a) + ggx_lambda(nv, a));
}
__device__ __forceinline__ float luminance(v3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

__device__ __forceinline__ v3 eval_bsdf(const Mat &M, v3 n, v3 l, v3 v) {
  float nl = dot(n, l), nv = dot(n, v);
  if (nl <= 0.0f || nv <= 0.0f) return mk(0.f, 0.f, 0.f);
  if (M.kind == 3) return M.c * INV_PI_F;
  v3 h = normalize(l + v);
  float nh = dot(n, h), vh = fmaxf(0.0f, dot(v, h));
  if (M.kind == 2) {
    float coat = ggx_d(n, h, nh, M.alpha_c) * ggx_g2(nl, nv, M.alpha_c) * fresnel_dielectric(vh, M.ior) / (4.0f * nl * nv);
    float t = (1.0f - fresnel_dielectric(nl, M.ior)) * (1.0f - fresnel_dielectric(nv, M.ior));
    float metal = ggx_d(n, h, nh, M.alpha) * ggx_g2(nl, nv, M.alpha) / (4.0f * nl * nv);
    float e = 1.0f / nl + 1.0f / nv;
    v3 tp = mk(powf(M.tau.x, e), powf(M.tau.y, e), powf(M.tau.z, e));
    return mk(coat, coat, coat) + tp * schlick3(M.c, vh) * (t * metal);
  }
  float spec = ggx_d(n, h, nh, M.alpha) * ggx_g2(nl, nv, M.alpha) / (4.0f * nl * nv);
  v3 F = M.kind == 1 ? schlick3(M.c, vh) : [&] { float f = fresnel_dielectric(vh, M.ior); return mk(f, f, f); }();
  v3 out = F * spec;
  if (M.kind == 0) {
    float kd = INV_PI_F * (1.0f - fresnel_dielectric(nl, M.ior)) * (1.0f - fresnel_dielectric(nv, M.ior)) / (1.0f - M.fbar);
    out = out + M.c * kd;
  }
  return out;
}
// d f / d c per channel (c = rho_d of dielectric and Lambert, F0 of conductor and coated conductor), the
// derivative the adjoint needs; 0 outside the upper hemisphere like eval_bsdf
__device__ __forceinline__ v3 eval_bsdf_dc(const Mat &M, v3 n, v3 l, v3 v) {
  float nl = dot(n, l), nv = dot(n, v);
  if (nl <= 0.0f || nv <= 0.0f) return mk(0.f, 0.f, 0.f);
  if (M.kind == 3) return mk(INV_PI_F, INV_PI_F, INV_PI_F);
  if (M.kind == 0) {
    float kd = INV_PI_F * (1.0f - fresnel_dielectric(nl, M.ior)) * (1.0f - fresnel_dielectric(nv, M.ior)) / (1.0f - M.fbar);
    return mk(kd, kd, kd);
  }
  v3 h = normalize(l + v);
  float nh = dot(n, h), vh = fmaxf(0.0f, dot(v, h));
  float one_m = 1.0f - pow5(1.0f - fminf(fmaxf(vh, 0.0f), 1.0f));
  if (M.kind == 1) {
    float spec = ggx_d(n, h, nh, M.alpha) * ggx_g2(nl, nv, M.alpha) / (4.0f * nl * nv);
    return mk(spec * one_m, spec * one_m, spec * one_m);
  }
  float t = (1.0f - fresnel_dielectric(nl, M.ior)) * (1.0f - fresnel_dielectric(nv, M.ior));
  float metal = ggx_d(n, h, nh, M.alpha) * ggx_g2(nl, nv, M.alpha) / (4.0f * nl * nv);
  float e = 1.0f / nl + 1.0f / nv;
  return mk(powf(M.tau.x, e), powf(M.tau.y, e), powf(M.tau.z, e)) * (t * metal * one_m);
}
__device__ __forceinline__ float spec_prob(const Mat &M, float nv) {
  if (M.kind == 3) return 0.0f;
  if (M.kind == 2) {
    float Fc = fresnel_dielectric(nv, M.ior);
    return Fc / (Fc + (1.0f - Fc) * (M.c.x + M.c.y + M.c.z) / 3.0f);
  }
  if (M.kind == 1) return 1.0f;
  float F = fresnel_dielectric(nv, M.ior);
  float d = luminance(M.c) * (1.0f - F);
  float s = fmaxf(F, 0.04f);
  return fminf(fmaxf(s / (s + d), 0.05f), 0.95f);
}
__device__ __forceinline__ float pdf_bsdf(const Mat &M, v3 n, v3 v, v3 l) {
  float nl = dot(n, l), nv = dot(n, v);
  if (nl <= 0.0f || nv <= 0.0f) return 0.0f;
  if (M.kind == 3) return nl * INV_PI_F;
  v3 h = normalize(l + v);
  float nh = dot(n, h), vh = dot(v, h);
  float ps = spec_prob(M, nv);
  if (M.kind == 2) {
    if (!(vh > 0.0f)) return 0.0f;
    return ps * ggx_g1(nv, M.alpha_c) * ggx_d(n, h, nh, M.alpha_c) / (4.0f * nv) +
           (1.0f - ps) * ggx_g1(nv, M.alpha) * ggx_d(n, h, nh, M.alpha) / (4.0f * nv);
  }
  float pdf_spec = vh > 0.0f ? ggx_g1(nv, M.alpha) * ggx_d(n, h, nh, M.alpha) / (4.0f * nv) : 0.0f;
  return ps * pdf_spec + (1.0f - ps) * nl * INV_PI_F;
}
// Duff et al. 2017 basis (common.wgsl.js:146-157)
__device__ __forceinline__ void basis(v3 n, v3 &t, v3 &b) {
  float s = n.z >= 0.0f ? 1.0f : -1.0f;
  float a = -1.0f / (s + n.z);
  float bb = n.x * n.y * a;
  t = mk(1.0f + s * n.x * n.x * a, s * bb, -s * n.x);
  b = mk(bb, s + n.y * n.y * a, -n.y);
}
// Heitz 2018 visible-normal sampling (bsdf.wgsl.js:86-101)
__device__ __forceinline__ v3 sample_vndf(v3 ve, float a, float u1, float u2) {
  v3 vh = normalize(mk(a * ve.x, a * ve.y, ve.z));
  float lensq = vh.x * vh.x + vh.y * vh.y;
  v3 T1 = mk(1.f, 0.f, 0.f);
  if (lensq > 0.0f) T1 = mk(-vh.y, vh.x, 0.0f) * (1.0f / sqrtf(lensq));
  v3 T2 = cross(vh, T1);
  float r = sqrtf(u1), phi = 2.0f * PI_F * u2;
  float sn, cs;
  sincosf(phi, &sn, &cs);
  float t1 = r * cs, t2 = r * sn;
  float s = 0.5f * (1.0f + vh.z);
  t2 = (1.0f - s) * sqrtf(fmaxf(0.0f, 1.0f - t1 * t1)) + s * t2;
  float t3 = sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2));
  v3 nh = T1 * t1 + T2 * t2 + vh * t3;
  return normalize(mk(a * nh.x, a * nh.y, fmaxf(1e-7f, nh.z)));
}
// returns pdf (0: no sample), writes l (bsdf.wgsl.js:133-155)
__device__ __forceinline__ float sample_bsdf(const Mat &M, v3 n, v3 v, float u0, float u1, float u2, v3 &l) {
  float nv = dot(n, v);
  if (nv <= 0.0f) return 0.0f;
  v3 t, b;
  basis(n, t, b);
  float ps = spec_prob(M, nv);
  if (M.kind == 2 || (M.kind != 3 && u0 < ps)) {
    float a = M.kind == 2 ? (u0 < ps ? M.alpha_c : M.alpha) : M.alpha;
    v3 hl = sample_vndf(mk(dot(t, v), dot(b, v), nv), a, u1, u2);
    v3 h = t * hl.x + b * hl.y + n * hl.z;
    l = h * (2.0f * dot(v, h)) - v;
  } else {
    float r = sqrtf(u1), phi = 2.0f * PI_F * u2;
    float sn, cs;
    sincosf(phi, &sn, &cs);
    l = t * (r * cs) + b * (r * sn) + n * sqrtf(fmaxf(0.0f, 1.0f - u1));
  }
  if (dot(n, l) <= 0.0f) return 0.0f;
  return pdf_bsdf(M, n, v, l);
}

// ------------------------------------------------------------------ scene on the device
struct __attribute__((aligned(128))) Node4 {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  uint32_t ref[4];      // interior: node index; leaf: LEAF_BIT | (count - 1) << 26 | first slot; EMPTY_REF
  uint32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 is one 128-B cache line");

struct SceneDev {
  const Node4 *nodes;
  const float4 *tri;        // 3 float4 per BVH slot: v0, v1, v2 (w unused)
  const float4 *tshade;     // 4 per slot, one 64-B line: vertex normals n0 (w: mesh id bits), n1, n2, and the
                            // geometric normal normalize((v1 - v0) x (v2 - v0))
  const uint4 *meshes;      // material, flags, light, -
  const float4 *mats;       // 4 float4 per material (16 floats)
  const float4 *lights;     // radiance rgb
  const float4 *emit;       // 4 float4 per emitter triangle
  uint32_t emit_count, root;
  float emit_area;
  const float4 *inst;       // posed bodies, INST_WORDS float4 each (layout at inst_pose below)
  uint32_t inst_count;
  uint32_t shadow_skip;     // bodies shadow rays ignore (bit per body)
  uint32_t emit_body;       // the body whose frame holds the emitter records and light axes (NO_INST: the world)
  const float4 *lprof;      // per light: emission axis xyz, w = bits(knot count); knots at light * MAX_KNOTS
  const float2 *knots;      // (theta, p) knots of the angular emission profiles
  const uint4 *tvid;        // per slot: vertex ids v0 v1 v2, mesh
  // node radiosity cache (null until built): 4 float4 per vertex (E_free + V, slot-1 source, slot-2 source,
  // B_dir + visibility code); albedo tables (per material ALB_BINS x [spec, diff]) and hemispherical albedo;
  // radu[0] = eAvg.xyz + bits(current source slot), radu[1] = emitter centroid xyz + radius
  const float4 *nstate;
  const float4 *alb;
  const float4 *hh;
  const float4 *radu;
  // the printed board: texture coordinates per vertex (null without textured meshes) and the mask, 8 bits per
  // texel, 255 = paper
  const float2 *uv;
  const uint8_t *tex;
  uint32_t tex_w, tex_h;
};
constexpr uint32_t ALB_BINS = 32;

// The print mask at texture coordinates (u, v): bilinear between the four nearest texel centres, texel (i, j)
// centred at ((i + 0.5) / w, (j + 0.5) / h), clamped to the edge texels (textureSampleLevel with a linear,
// clamp-to-edge sampler, gpu/renderer.js:398 and gpu/wgsl/shade.wgsl.js:262), in binary32 with every product
// rounded on its own; 1 = paper.
__device__ __forceinline__ float board_mask(const SceneDev &S, float u, float v) {
#pragma clang fp contract(off)
  const float fw = (float)S.tex_w, fh = (float)S.tex_h;
  const float x = fminf(fmaxf(u * fw - 0.5f, -1.0f), fw), y = fminf(fmaxf(v * fh - 0.5f, -1.0f), fh);
  const float x0 = floorf(x), y0 = floorf(y);
  const float ax = x - x0, ay = y - y0;
  const int last_x = (int)S.tex_w - 1, last_y = (int)S.tex_h - 1;
  const int i0 = min(max((int)x0, 0), last_x), i1 = min(max((int)x0 + 1, 0), last_x);
  const int j0 = min(max((int)y0, 0), last_y), j1 = min(max((int)y0 + 1, 0), last_y);
  const float t00 = (float)S.tex[(size_t)j0 * S.tex_w + i0] * (1.0f / 255.0f);
  const float t10 = (float)S.tex[(size_t)j0 * S.tex_w + i1] * (1.0f / 255.0f);
  const float t01 = (float)S.tex[(size_t)j1 * S.tex_w + i0] * (1.0f / 255.0f);
  const float t11 = (float)S.tex[(size_t)j1 * S.tex_w + i1] * (1.0f / 255.0f);
  const float top = t00 * (1.0f - ax) + t10 * ax, bot = t01 * (1.0f - ax) + t11 * ax;
  return top * (1.0f - ay) + bot * ay;
}
// The material at a hit of slot `tri` (barycentrics bu, bv): the mesh's record, and on a textured mesh
// (charuco_print, gpu/wgsl/shade.wgsl.js:258-265) rho_d and alpha mixed between the ink (the record's words 8-11)
// and the paper by the mask at the interpolated texture coordinates: c = ink (1 - m) + paper m.
__device__ __forceinline__ Mat hit_mat(const SceneDev &S, uint4 ms, uint32_t tri, float bu, float bv) {
#pragma clang fp contract(off)
  Mat M = load_mat(S.mats, ms.x);
  if (ms.y & F_TEXTURED) {
    const uint4 vi = S.tvid[tri];
    const float2 a = S.uv[vi.x], b = S.uv[vi.y], c = S.uv[vi.z];
    const float w0 = 1.0f - bu - bv;
    const float u = a.x * w0 + b.x * bu + c.x * bv, v = a.y * w0 + b.y * bu + c.y * bv;
    const float m = board_mask(S, u, v), k = 1.0f - m;
    M.c = mk(M.tau.x * k + M.c.x * m, M.tau.y * k + M.c.y * m, M.tau.z * k + M.c.z * m);
    M.alpha = M.alpha_c * k + M.alpha * m;
  }
  return M;
}

constexpr uint32_t MAX_KNOTS = 64;
// p(theta) of light `light` toward the unit direction w leaving the emitter, theta from the light's axis: linear
// between knots, 0 past the last one (shade.wgsl.js:12-27); 1 for a light without a profile
__device__ __forceinline__ float light_profile(const SceneDev &S, uint32_t light, v3 w) {
  const float4 a = S.lprof[light];
  const uint32_t cnt = __float_as_uint(a.w);
  if (cnt == 0) return 1.0f;
  const float2 *K = S.knots + (size_t)light * MAX_KNOTS;
  const float th = acosf(fminf(fmaxf(w.x * a.x + w.y * a.y + w.z * a.z, -1.0f), 1.0f));
  const float2 last = K[cnt - 1];
  if (th >= last.x) return th == last.x ? last.y : 0.0f;
  uint32_t lo = 0, hi = cnt - 1;
  for (int it = 0; it < 8; it++) {
    if (hi - lo <= 1) break;
    const uint32_t m = (lo + hi) >> 1;
    if (K[m].x <= th) lo = m; else hi = m;
  }
  const float2 ka = K[lo], kb = K[hi];
  const float f = fminf(fmaxf((th - ka.x) / fmaxf(kb.x - ka.x, 1e-12f), 0.0f), 1.0f);
  return ka.y + (kb.y - ka.y) * f;
}

constexpr uint32_t NO_INST = 0xffffffffu;
struct Hit {
  float t, u, v;
  uint32_t tri;
  uint32_t inst;     // posed body of the hit, NO_INST for the static scene
};

// per-lane traversal stack: SLDS entries in LDS, the rest spilled to a per-lane global area. The LDS arrays are
// held as LDS-address-space pointers: as generic pointers, the branch between an LDS and a spill read is merged
// into one generic (flat) load, which waits on every outstanding memory access of the wave.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) float lds_f32;
struct Stack {
  lds_u32 *lds_node;    // base of this block's node-index array, [slot * BLOCK + tid]
  lds_f32 *lds_t;
  uint2 *spill;         // [(slot - SLDS) * spill_stride + gtid]
  uint32_t tid, block, slds, spill_depth;
  uint64_t spill_stride, gtid;
  uint32_t *overflow;   // global counter of entries that found no slot
};
__device__ __forceinline__ void push(Stack &S, uint32_t &sp, uint32_t node, float t) {
  if (sp < S.slds) {
    S.lds_node[sp * S.block + S.tid] = node;
    S.lds_t[sp * S.block + S.tid] = t;
  } else if (sp < S.slds + S.spill_depth) {
    S.spill[(uint64_t)(sp - S.slds) * S.spill_stride + S.gtid] = make_uint2(node, __float_as_uint(t));
  } else {
    atomicAdd(S.overflow, 1u);
    return;
  }
  sp++;
}
__device__ __forceinline__ void pop(const Stack &S, uint32_t sp, uint32_t &node, float &t) {
  if (sp < S.slds) {
    node = S.lds_node[sp * S.block + S.tid];
    t = S.lds_t[sp * S.block + S.tid];
  } else {
    uint2 e = S.spill[(uint64_t)(sp - S.slds) * S.spill_stride + S.gtid];
    node = e.x;
    t = __uint_as_float(e.y);
  }
}

// Watertight ray-triangle test (Woop, Benthin, Wald, JCGT 2013), two-sided, with the FP64 fallback on an
// exact edge. Edge functions are evaluated unfused so a shared edge gives exactly negated values.
struct RayW {
  v3 o;
  int kx, ky, kz;
  float Sx, Sy, Sz;
};
__device__ __forceinline__ float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__device__ __forceinline__ RayW ray_w(v3 o, v3 d) {
  RayW R;
  R.o = o;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  int kx = kz == 2 ? 0 : kz + 1;
  int ky = kx == 2 ? 0 : kx + 1;
  if (comp(d, kz) < 0.0f) { int s = kx; kx = ky; ky = s; }
  R.kx = kx; R.ky = ky; R.kz = kz;
  float dz = comp(d, kz);
  R.Sx = comp(d, kx) / dz;
  R.Sy = comp(d, ky) / dz;
  R.Sz = 1.0f / dz;
  return R;
}
__device__ __forceinline__ bool tri_hit(const RayW &R, v3 p0, v3 p1, v3 p2, float tmin, float tmax, float &t,
                                        float &u, float &v) {
#pragma clang fp contract(off)
  v3 A = p0 - R.o, B = p1 - R.o, C = p2 - R.o;
  float Az = comp(A, R.kz), Bz = comp(B, R.kz), Cz = comp(C, R.kz);
  float Ax = comp(A, R.kx) - R.Sx * Az, Ay = comp(A, R.ky) - R.Sy * Az;
  float Bx = comp(B, R.kx) - R.Sx * Bz, By = comp(B, R.ky) - R.Sy * Bz;
  float Cx = comp(C, R.kx) - R.Sx * Cz, Cy = comp(C, R.ky) - R.Sy * Cz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float T = U * (R.Sz * Az) + V * (R.Sz * Bz) + W * (R.Sz * Cz);
  float rd = 1.0f / det;
  float tt = T * rd;
  if (!(tt > tmin && tt < tmax)) return false;
  t = tt;
  u = V * rd;
  v = W * rd;
  return true;
}

// The triangles of one leaf in slot order. Returns true when ANY stops at a hit. (Loading two triangles before
// either test measured the same on the full-scale frame and spilled registers, so the loop stays one at a time.)
template <bool ANY>
__device__ __forceinline__ bool leaf_test(const float4 *tri, const RayW &R, uint32_t leaf, float tmin, float &best,
                                          Hit &h, bool &found) {
  const uint32_t first = leaf & 0x03ffffffu, end = first + ((leaf >> 26) & 31u) + 1u;
  for (uint32_t t = first; t < end; t++) {
    const float4 *tp = tri + 3 * (size_t)t;
    const float4 a = tp[0], b = tp[1], c = tp[2];
    float tt, u, v;
    if (tri_hit(R, ld3(a), ld3(b), ld3(c), tmin, best, tt, u, v)) {
      best = tt;
      h.t = tt; h.u = u; h.v = v; h.tri = t;
      found = true;
      if (ANY) return true;
    }
  }
  return false;
}

__device__ __forceinline__ v3 safe_inv(v3 d) {
  return mk(1.0f / copysignf(fmaxf(fabsf(d.x), 1e-20f), d.x), 1.0f / copysignf(fmaxf(fabsf(d.y), 1e-20f), d.y),
            1.0f / copysignf(fmaxf(fabsf(d.z), 1e-20f), d.z));
}

// Per-ray constants of the four-wide slab test. Along each axis the near plane is the lo plane where the
// direction is >= 0 (inv > 0) and the hi plane otherwise; fma(x, inv, oi) is monotone in x for a fixed inv, so
// the near plane's distance is exactly min(t_lo, t_hi) and the far plane's max(t_lo, t_hi): each child's entry
// and exit are a max3 and a min3 of plane distances, with no min/max per axis. An empty child (lo = +inf,
// hi = -inf) gives entry +inf > exit -inf; its ref is tested all the same, which keeps the ref load issued with
// the plane loads rather than after the test (a second round trip to memory per node). `near` packs the byte offsets of
// the near plane arrays in a Node4, one per byte (x: 0 or 16, y: 32 or 48, z: 64 or 80); a far plane array is at
// its near one ^ 16. One register instead of three: the trace kernels are bound by memory latency, so resident
// waves count for more than the three bit-field extracts per node.
struct Slab {
  v3 inv, oi;
  uint32_t near;
};
__device__ __forceinline__ Slab make_slab(v3 o, v3 d) {
  Slab R;
  R.inv = safe_inv(d);
  R.oi = mk(-o.x * R.inv.x, -o.y * R.inv.y, -o.z * R.inv.z);
  R.near = (R.inv.x > 0.0f ? 0u : 16u) | (R.inv.y > 0.0f ? 32u : 48u) << 8 | (R.inv.z > 0.0f ? 64u : 80u) << 16;
  return R;
}
// The node array as a buffer resource: a node's seven 16-B loads then take one 32-bit lane offset each
// (ref * 128 + the plane's offset) instead of a 64-bit address.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t node_rsrc(const Node4 *nodes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<Node4 *>(nodes), (short)0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ float4 ld_node16(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
// entry distances of the four children of node `ref` in ascending order (INFINITY: missed, or entered past
// tfar) and their refs
__device__ __forceinline__ void node4_test(__amdgpu_buffer_rsrc_t nodes, uint32_t ref, const Slab &R, float tmin,
                                           float tfar, float tn[4], uint32_t rr[4]) {
  const uint32_t b = ref << 7;
  const uint32_t ox = R.near & 0xffu, oy = (R.near >> 8) & 0xffu, oz = R.near >> 16;
  const float4 nx = ld_node16(nodes, b + ox), fx = ld_node16(nodes, b + (ox ^ 16u));
  const float4 ny = ld_node16(nodes, b + oy), fy = ld_node16(nodes, b + (oy ^ 16u));
  const float4 nz = ld_node16(nodes, b + oz), fz = ld_node16(nodes, b + (oz ^ 16u));
  const uint4 rf = __builtin_bit_cast(uint4, ld_node16(nodes, b + 96u));
  uint32_t r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
  const uint32_t rref[4] = {r0, r1, r2, r3};
  const float pnx[4] = {nx.x, nx.y, nx.z, nx.w}, pfx[4] = {fx.x, fx.y, fx.z, fx.w};
  const float pny[4] = {ny.x, ny.y, ny.z, ny.w}, pfy[4] = {fy.x, fy.y, fy.z, fy.w};
  const float pnz[4] = {nz.x, nz.y, nz.z, nz.w}, pfz[4] = {fz.x, fz.y, fz.z, fz.w};
  float t[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const float a = fmaxf(fmaxf(fmaf(pnx[c], R.inv.x, R.oi.x), fmaf(pny[c], R.inv.y, R.oi.y)),
                          fmaxf(fmaf(pnz[c], R.inv.z, R.oi.z), tmin));
    const float e = fminf(fminf(fmaf(pfx[c], R.inv.x, R.oi.x), fmaf(pfy[c], R.inv.y, R.oi.y)),
                          fminf(fmaf(pfz[c], R.inv.z, R.oi.z), tfar));
    t[c] = a <= e && rref[c] != EMPTY_REF ? a : INFINITY;
  }
  float t0 = t[0], t1 = t[1], t2 = t[2], t3 = t[3];
  // ascending (t, ref): a five-comparator network of selects (a stable order: equal entries keep theirs)
#define VT_CSWAP(ti, tj, ri, rj)                                          \
  {                                                                       \
    const bool s = tj < ti;                                               \
    const float lo = s ? tj : ti, hi = s ? ti : tj;                       \
    const uint32_t rlo = s ? rj : ri, rhi = s ? ri : rj;                  \
    ti = lo; tj = hi; ri = rlo; rj = rhi;                                 \
  }
  VT_CSWAP(t0, t1, r0, r1) VT_CSWAP(t2, t3, r2, r3) VT_CSWAP(t0, t2, r0, r2) VT_CSWAP(t1, t3, r1, r3)
  VT_CSWAP(t1, t2, r1, r2)
#undef VT_CSWAP
  tn[0] = t0; tn[1] = t1; tn[2] = t2; tn[3] = t3;
  rr[0] = r0; rr[1] = r1; rr[2] = r2; rr[3] = r3;
}

// Closest (any_hit = false) or any hit in (tmin, tmax), "while-while" with postponed leaves (Aila and Laine,
// HPG 2009): a lane walks interior nodes and sets aside the first leaf it reaches, and the wave tests leaves only
// when every active lane holds one, so the node phase and the triangle phase each run with full lanes instead of
// alternating inside one loop.
constexpr uint32_t TRAV_DONE = 0xfffffffeu;
__device__ __forceinline__ bool is_leaf_ref(uint32_t r) { return (r & LEAF_BIT) && r != TRAV_DONE; }

// next entry nearer than the best hit, or TRAV_DONE
__device__ __forceinline__ uint32_t pop_next(const Stack &st, uint32_t &sp, float best) {
  while (sp > 0) {
    sp--;
    uint32_t r;
    float t;
    pop(st, sp, r, t);
    if (t < best) return r;
  }
  return TRAV_DONE;
}

template <bool ANY>
__device__ __forceinline__ bool traverse(const SceneDev &S, uint32_t root, v3 o, v3 d, float tmin, float tmax, Hit &h,
                                         Stack &st) {
  const Slab B = make_slab(o, d);
  const __amdgpu_buffer_rsrc_t NR = node_rsrc(S.nodes);
  const RayW R = ray_w(o, d);
  float best = tmax;
  bool found = false;
  uint32_t sp = 0;
  uint32_t ref = root;
  uint32_t leaf = TRAV_DONE;      // postponed leaf
  while (ref != TRAV_DONE || leaf != TRAV_DONE) {
    // node phase
    while (ref != TRAV_DONE && !(ref & LEAF_BIT)) {
      float tn[4];
      uint32_t rr[4];
      node4_test(NR, ref, B, tmin, best * 1.00000024f, tn, rr);
      if (tn[0] < INFINITY) {
        if (tn[3] < INFINITY) push(st, sp, rr[3], tn[3]);
        if (tn[2] < INFINITY) push(st, sp, rr[2], tn[2]);
        if (tn[1] < INFINITY) push(st, sp, rr[1], tn[1]);
        ref = rr[0];
      } else {
        ref = pop_next(st, sp, best);
      }
      // set the first leaf aside and keep walking nodes while other lanes of the wave still need a leaf
      if (is_leaf_ref(ref) && leaf == TRAV_DONE) {
        leaf = ref;
        ref = pop_next(st, sp, best);
      }
      if (!__any(leaf == TRAV_DONE)) break;
    }
    // triangle phase: the postponed leaf, then any leaf the node phase stopped on
    while (leaf != TRAV_DONE) {
      leaf_test<ANY>(S.tri, R, leaf, tmin, best, h, found);
      if (ANY && found) {
        leaf = TRAV_DONE;
        ref = TRAV_DONE;
        break;
      }
      leaf = TRAV_DONE;
      if (is_leaf_ref(ref)) {
        leaf = ref;
        ref = pop_next(st, sp, best);
      }
    }
  }
  return found;
}

// ------------------------------------------------------------------ posed bodies (trace.wgsl.js:122-164)
// Instance record, INST_WORDS float4: [0] q0 and [1] q1 (unit quaternions x y z w at shutter open and close),
// [2] p0.xyz + BLAS root (bits), [3] p1.xyz + flags (bits: 1 moving, 2 hidden), [4] world box min, [5] world box
// max (swept over the exposure), [6] bounding capsule a.xyz + radius, [7] capsule b.xyz, [8] occlusion capsule
// a.xyz + radius and [9] b.xyz + rho_eff (world frame at mid-exposure; radius 0 = none).
constexpr int INST_WORDS = 10;
constexpr uint32_t INST_MOVING = 1u, INST_HIDDEN = 2u;

struct Pose {
  v3 c0, c1, c2;    // rotation columns
  v3 p;
};
// quaternion (x, y, z, w) -> rotation columns (common.wgsl.js:117-123 quatToMat)
__device__ __forceinline__ void quat_cols(float4 q, v3 &c0, v3 &c1, v3 &c2) {
  const float x = q.x, y = q.y, z = q.z, w = q.w;
  c0 = mk(1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + z * w), 2.0f * (x * z - y * w));
  c1 = mk(2.0f * (x * y - z * w), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + x * w));
  c2 = mk(2.0f * (x * z + y * w), 2.0f * (y * z - x * w), 1.0f - 2.0f * (x * x + y * y));
}
// the shorter-arc spherical interpolation, normalised lerp when the ends nearly agree (common.wgsl.js:125-133)
__device__ __forceinline__ float4 slerp4(float4 a, float4 b, float t) {
  float c = a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
  if (c < 0.0f) {
    b = make_float4(-b.x, -b.y, -b.z, -b.w);
    c = -c;
  }
  if (c > 0.9995f) {
    float4 m = make_float4(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, a.w + (b.w - a.w) * t);
    const float il = 1.0f / sqrtf(m.x * m.x + m.y * m.y + m.z * m.z + m.w * m.w);
    return make_float4(m.x * il, m.y * il, m.z * il, m.w * il);
  }
  const float th = acosf(c), s = sinf(th);
  const float wa = sinf((1.0f - t) * th) / s, wb = sinf(t * th) / s;
  return make_float4(a.x * wa + b.x * wb, a.y * wa + b.y * wb, a.z * wa + b.z * wb, a.w * wa + b.w * wb);
}
// pose of instance i at exposure time t (common.wgsl.js:137-143 instPose)
__device__ __forceinline__ Pose inst_pose(const SceneDev &S, uint32_t i, float t) {
  const float4 *I = S.inst + INST_WORDS * (size_t)i;
  const float4 p0 = I[2], p1 = I[3];
  Pose P;
  if (__float_as_uint(p1.w) & INST_MOVING) {
    quat_cols(slerp4(I[0], I[1], t), P.c0, P.c1, P.c2);
    P.p = mk(p0.x + (p1.x - p0.x) * t, p0.y + (p1.y - p0.y) * t, p0.z + (p1.z - p0.z) * t);
  } else {
    quat_cols(I[0], P.c0, P.c1, P.c2);
    P.p = mk(p0.x, p0.y, p0.z);
  }
  return P;
}
__device__ __forceinline__ v3 rot(const Pose &P, v3 v) { return P.c0 * v.x + P.c1 * v.y + P.c2 * v.z; }
__device__ __forceinline__ v3 rot_t(const Pose &P, v3 v) { return mk(dot(P.c0, v), dot(P.c1, v), dot(P.c2, v)); }

// does the segment o + t d, t in [t0, t1], pass within r of the segment a-b (trace.wgsl.js:122-136)
__device__ __forceinline__ bool seg_near_capsule(v3 o, v3 d, float t0, float t1, v3 a, v3 b, float r) {
  const v3 p0 = o + d * t0, u = d * (t1 - t0), v = b - a, w = p0 - a;
  const float A = dot(u, u), B = dot(u, v), Cc = dot(v, v), D = dot(u, w), E = dot(v, w);
  const float den = A * Cc - B * B;
  float s = 0.0f;
  if (den > 1e-20f * A * Cc) s = fminf(fmaxf((B * E - Cc * D) / den, 0.0f), 1.0f);
  float t = Cc > 1e-30f ? (B * s + E) / Cc : 0.0f;
  if (t < 0.0f) {
    t = 0.0f;
    s = fminf(fmaxf(-D / fmaxf(A, 1e-30f), 0.0f), 1.0f);
  } else if (t > 1.0f) {
    t = 1.0f;
    s = fminf(fmaxf((B - D) / fmaxf(A, 1e-30f), 0.0f), 1.0f);
  }
  const v3 q = (p0 + u * s) - (a + v * t);
  return dot(q, q) <= r * r;
}

__device__ __forceinline__ float box_entry(float4 lo, float4 hi, v3 o, v3 inv, float tmin, float tmax) {
  const float x0 = (lo.x - o.x) * inv.x, x1 = (hi.x - o.x) * inv.x;
  const float y0 = (lo.y - o.y) * inv.y, y1 = (hi.y - o.y) * inv.y;
  const float z0 = (lo.z - o.z) * inv.z, z1 = (hi.z - o.z) * inv.z;
  const float a = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
  const float b = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
  return a <= b ? a : INFINITY;
}

// The static scene, then every visible posed body (world box, bounding capsule, then its BLAS in its own frame at
// the ray's exposure time), as traceSceneSkip (trace.wgsl.js:144-164) does. h.inst names the body of the hit.
// the posed bodies after the static scene: `found` and h carry the static result in and the final one out
// skip: bit i set = ignore body i (a camera's own lens front for its camera rays, a light's housing for shadow rays)
template <bool ANY>
__device__ __forceinline__ void trace_instances(const SceneDev &S, v3 o, v3 d, float tmin, float tmax, float time,
                                                uint32_t skip, Hit &h, bool &found, Stack &st) {
  float best = found ? h.t : tmax;
  const v3 inv = safe_inv(d);
  for (uint32_t i = 0; i < S.inst_count; i++) {
    const float4 *I = S.inst + INST_WORDS * (size_t)i;
    const float4 p0 = I[2], p1 = I[3];
    if ((__float_as_uint(p1.w) & INST_HIDDEN) || ((skip >> i) & 1u)) continue;
    if (box_entry(I[4], I[5], o, inv, tmin, best) >= best) continue;
    const float4 ca = I[6], cb = I[7];
    if (!seg_near_capsule(o, d, tmin, fminf(best, 1e3f), ld3(ca), ld3(cb), ca.w)) continue;
    const Pose P = inst_pose(S, i, time);
    const v3 ol = rot_t(P, o - P.p), dl = rot_t(P, d);
    Hit hl;
    if (traverse<ANY>(S, __float_as_uint(p0.w), ol, dl, tmin, best, hl, st)) {
      h = hl;
      h.inst = i;
      best = hl.t;
      found = true;
      if (ANY) return;
    }
  }
}

// INST = false: kernels built for scenes without posed bodies (the instance walk and its registers compiled out)
template <bool ANY, bool INST>
__device__ __forceinline__ bool trace_scene(const SceneDev &S, v3 o, v3 d, float tmin, float tmax, float time,
                                            uint32_t skip, Hit &h, Stack &st) {
  bool found = traverse<ANY>(S, S.root, o, d, tmin, tmax, h, st);
  if (found) {
    h.inst = NO_INST;
    if (ANY) return true;
  }
  if (INST && S.inst_count) trace_instances<ANY>(S, o, d, tmin, tmax, time, skip, h, found, st);
  return found;
}

// The emitter's frame at exposure time t: the pose of the body it rides on (the reference's light rides on the
// scope: shade.wgsl.js:45, 55, 120), or the world for an emitter in the static scene.
struct EmitPose {
  Pose P;
  bool posed;
};
__device__ __forceinline__ EmitPose emitter_pose(const SceneDev &S, float t) {
  EmitPose E;
  E.posed = S.emit_body != NO_INST;
  if (E.posed) E.P = inst_pose(S, S.emit_body, t);
  return E;
}
// a world direction in the emitter's frame, where the light axes of the emission profiles are
__device__ __forceinline__ v3 emit_local(const EmitPose &E, v3 w) { return E.posed ? rot_t(E.P, w) : w; }
// the emitter's centre (radu[1].xyz, emitter frame) in the world at the pose E
__device__ __forceinline__ v3 emit_centre(const SceneDev &S, const EmitPose &E) {
  const v3 c = ld3(S.radu[1]);
  return E.posed ? rot(E.P, c) + E.P.p : c;
}

// emitter sample (shade.wgsl.js:31-47): area-CDF pick and a uniform point on the triangle, posed by E; p and the
// emitter's normal n in the world; returns the record index
__device__ __forceinline__ uint32_t sample_emitter(const SceneDev &S, const EmitPose &E, float ux, float uy, float uz,
                                                   v3 &p, v3 &n) {
  uint32_t lo = 0, hi = S.emit_count - 1;
  for (int it = 0; it < 40; it++) {
    if (lo >= hi) break;
    uint32_t m = (lo + hi) / 2;
    if (S.emit[4 * m].w < ux) lo = m + 1; else hi = m;
  }
  float4 v0 = S.emit[4 * lo], e1 = S.emit[4 * lo + 1], e2 = S.emit[4 * lo + 2];
  float su = sqrtf(uy);
  p = ld3(v0) + ld3(e1) * (su * (1.0f - uz)) + ld3(e2) * (su * uz);
  n = ld3(S.emit[4 * lo + 3]);
  if (E.posed) {
    p = rot(E.P, p) + E.P.p;
    n = rot(E.P, n);
  }
  return lo;
}

}  // namespace vt
