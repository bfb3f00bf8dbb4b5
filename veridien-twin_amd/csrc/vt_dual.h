// Forward-mode derivatives for the camera adjoint: a value with ND tangents, and the BSDF of vt_device.h
// restated on such values. The tangents are four directions of the camera ray: the origin and the direction
// each moved along two unit vectors perpendicular to the ray (a move along the ray leaves the first hit, and
// so the path, unchanged, and the camera's unit direction only ever moves perpendicular to itself). Branches
// are taken on the value, exactly as the float code takes them, so every function below is the derivative of
// its float counterpart wherever that is differentiable (the interior term: no visibility or branch changes).
#pragma once
#include "vt_device.h"

namespace vt {

constexpr int ND = 4;

struct D {
  float v;
  float g[ND];
};
__device__ __forceinline__ D dk(float v) {
  D r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < ND; k++) r.g[k] = 0.0f;
  return r;
}
__device__ __forceinline__ D operator+(const D &a, const D &b) {
  D r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < ND; k++) r.g[k] = a.g[k] + b.g[k];
  return r;
}
__device__ __forceinline__ D operator-(const D &a, const D &b) {
  D r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < ND; k++) r.g[k] = a.g[k] - b.g[k];
  return r;
}
__device__ __forceinline__ D operator-(const D &a) {
  D r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < ND; k++) r.g[k] = -a.g[k];
  return r;
}
__device__ __forceinline__ D operator*(const D &a, const D &b) {
  D r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < ND; k++) r.g[k] = a.g[k] * b.v + a.v * b.g[k];
  return r;
}
__device__ __forceinline__ D operator*(const D &a, float s) {
  D r;
  r.v = a.v * s;
#pragma unroll
  for (int k = 0; k < ND; k++) r.g[k] = a.g[k] * s;
  return r;
}
__device__ __forceinline__ D operator*(float s, const D &a) { return a * s; }
__device__ __forceinline__ D operator+(const D &a, float s) {
  D r = a;
  r.v = a.v + s;
  return r;
}
__device__ __forceinline__ D operator+(float s, const D &a) { return a + s; }
__device__ __forceinline__ D operator-(const D &a, float s) { return a + (-s); }
__device__ __forceinline__ D operator-(float s, const D &a) { return (-a) + s; }
__device__ __forceinline__ D rcp(const D &a) {
  D r;
  r.v = 1.0f / a.v;
  const float d = -r.v * r.v;
#pragma unroll
  for (int k = 0; k < ND; k++) r.g[k] = a.g[k] * d;
  return r;
}
__device__ __forceinline__ D operator/(const D &a, const D &b) {
  D r;
  r.v = a.v / b.v;
  const float ib = 1.0f / b.v;
#pragma unroll
  for (int k = 0; k < ND; k++) r.g[k] = (a.g[k] - r.v * b.g[k]) * ib;
  return r;
}
__device__ __forceinline__ D operator/(const D &a, float s) { return a * (1.0f / s); }
__device__ __forceinline__ D operator/(float s, const D &b) { return rcp(b) * s; }
__device__ __forceinline__ D dsqrt(const D &a) {
  D r;
  r.v = sqrtf(a.v);
  const float d = r.v > 0.0f ? 0.5f / r.v : 0.0f;
#pragma unroll
  for (int k = 0; k < ND; k++) r.g[k] = a.g[k] * d;
  return r;
}
__device__ __forceinline__ D dmax(const D &a, float b) { return a.v >= b ? a : dk(b); }
__device__ __forceinline__ D dmin(const D &a, float b) { return a.v <= b ? a : dk(b); }
__device__ __forceinline__ D dclamp01(const D &a) { return dmin(dmax(a, 0.0f), 1.0f); }
// t^e for a constant base t > 0 and a variable exponent
__device__ __forceinline__ D dpow_const_base(float t, const D &e) {
  D r;
  r.v = powf(t, e.v);
  const float d = t > 0.0f ? r.v * logf(t) : 0.0f;
#pragma unroll
  for (int k = 0; k < ND; k++) r.g[k] = e.g[k] * d;
  return r;
}

struct DV {
  D x, y, z;
};
__device__ __forceinline__ DV dv(v3 a) { return {dk(a.x), dk(a.y), dk(a.z)}; }
__device__ __forceinline__ v3 val(const DV &a) { return mk(a.x.v, a.y.v, a.z.v); }
__device__ __forceinline__ DV operator+(const DV &a, const DV &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ DV operator-(const DV &a, const DV &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ DV operator-(const DV &a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ DV operator+(const DV &a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ DV operator-(v3 a, const DV &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ DV operator-(const DV &a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ DV operator*(const DV &a, const D &s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ DV operator*(const DV &a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ DV operator*(v3 a, const D &s) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ DV operator*(const DV &a, const DV &b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ DV operator*(const DV &a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ D ddot(const DV &a, const DV &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D ddot(const DV &a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ DV dcross(const DV &a, const DV &b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ DV dnormalize(const DV &a) { return a * rcp(dsqrt(ddot(a, a))); }

// ------------------------------------------------------------------ BSDF on tangent values
__device__ __forceinline__ D d_fresnel_dielectric(const D &cosi, float eta) {
  const D c = dclamp01(cosi);
  const D s2 = (1.0f - c * c) * (1.0f / (eta * eta));
  if (s2.v >= 1.0f) return dk(1.0f);
  const D ct = dsqrt(1.0f - s2);
  const D rs = (c - eta * ct) / (c + eta * ct);
  const D rp = (eta * c - ct) / (eta * c + ct);
  return 0.5f * (rs * rs + rp * rp);
}
__device__ __forceinline__ D d_pow5(const D &x) {
  const D x2 = x * x;
  return x2 * x2 * x;
}
__device__ __forceinline__ DV d_schlick3(v3 f0, const D &c) {
  const D m = d_pow5(1.0f - dclamp01(c));
  return {f0.x + (1.0f - f0.x) * m, f0.y + (1.0f - f0.y) * m, f0.z + (1.0f - f0.z) * m};
}
// ggx_d of vt_device.h: sin^2 of the half vector's angle from |n x h|^2
__device__ __forceinline__ D d_ggx_d(const DV &n, const DV &h, const D &nh, float a) {
  if (nh.v <= 0.0f) return dk(0.0f);
  const DV x = dcross(n, h);
  const float a2 = a * a;
  const D d = ddot(x, x) + (nh * nh) * a2;
  return a2 * rcp(PI_F * (d * d));
}
__device__ __forceinline__ D d_ggx_lambda(const D &c, float a) {
  const D c2 = dmax(c * c, 1e-12f);
  const D t2 = (1.0f - c2) / c2;
  return 0.5f * (dsqrt(1.0f + (a * a) * t2) - 1.0f);
}
__device__ __forceinline__ D d_ggx_g1(const D &c, float a) { return rcp(1.0f + d_ggx_lambda(c, a)); }
__device__ __forceinline__ D d_ggx_g2(const D &nl, const D &nv, float a) {
  return rcp(1.0f + d_ggx_lambda(nl, a) + d_ggx_lambda(nv, a));
}

__device__ __forceinline__ DV d_eval_bsdf(const Mat &M, const DV &n, const DV &l, const DV &v) {
  const D nl = ddot(n, l), nv = ddot(n, v);
  if (nl.v <= 0.0f || nv.v <= 0.0f) return dv(mk(0.f, 0.f, 0.f));
  if (M.kind == 3) return dv(M.c * INV_PI_F);
  const DV h = dnormalize(l + v);
  const D nh = ddot(n, h), vh = dmax(ddot(v, h), 0.0f);
  const D inv4 = rcp(4.0f * nl * nv);
  if (M.kind == 2) {
    const D coat = d_ggx_d(n, h, nh, M.alpha_c) * d_ggx_g2(nl, nv, M.alpha_c) * d_fresnel_dielectric(vh, M.ior) * inv4;
    const D t = (1.0f - d_fresnel_dielectric(nl, M.ior)) * (1.0f - d_fresnel_dielectric(nv, M.ior));
    const D metal = d_ggx_d(n, h, nh, M.alpha) * d_ggx_g2(nl, nv, M.alpha) * inv4;
    const D e = rcp(nl) + rcp(nv);
    const DV tp = {dpow_const_base(M.tau.x, e), dpow_const_base(M.tau.y, e), dpow_const_base(M.tau.z, e)};
    const DV sf = d_schlick3(M.c, vh);
    const D tm = t * metal;
    return {coat + tp.x * sf.x * tm, coat + tp.y * sf.y * tm, coat + tp.z * sf.z * tm};
  }
  const D spec = d_ggx_d(n, h, nh, M.alpha) * d_ggx_g2(nl, nv, M.alpha) * inv4;
  DV out;
  if (M.kind == 1) {
    out = d_schlick3(M.c, vh) * spec;
  } else {
    const D f = d_fresnel_dielectric(vh, M.ior) * spec;
    out = {f, f, f};
  }
  if (M.kind == 0) {
    const D kd = (1.0f - d_fresnel_dielectric(nl, M.ior)) * (1.0f - d_fresnel_dielectric(nv, M.ior)) *
                 (INV_PI_F / (1.0f - M.fbar));
    out = out + M.c * kd;
  }
  return out;
}
__device__ __forceinline__ D d_spec_prob(const Mat &M, const D &nv) {
  if (M.kind == 3) return dk(0.0f);
  if (M.kind == 2) {
    const D Fc = d_fresnel_dielectric(nv, M.ior);
    return Fc / (Fc + (1.0f - Fc) * ((M.c.x + M.c.y + M.c.z) / 3.0f));
  }
  if (M.kind == 1) return dk(1.0f);
  const D F = d_fresnel_dielectric(nv, M.ior);
  const D d = luminance(M.c) * (1.0f - F);
  const D s = dmax(F, 0.04f);
  return dmin(dmax(s / (s + d), 0.05f), 0.95f);
}
__device__ __forceinline__ D d_pdf_bsdf(const Mat &M, const DV &n, const DV &v, const DV &l) {
  const D nl = ddot(n, l), nv = ddot(n, v);
  if (nl.v <= 0.0f || nv.v <= 0.0f) return dk(0.0f);
  if (M.kind == 3) return nl * INV_PI_F;
  const DV h = dnormalize(l + v);
  const D nh = ddot(n, h), vh = ddot(v, h);
  const D ps = d_spec_prob(M, nv);
  const D inv4 = rcp(4.0f * nv);
  if (M.kind == 2) {
    if (!(vh.v > 0.0f)) return dk(0.0f);
    return ps * d_ggx_g1(nv, M.alpha_c) * d_ggx_d(n, h, nh, M.alpha_c) * inv4 +
           (1.0f - ps) * d_ggx_g1(nv, M.alpha) * d_ggx_d(n, h, nh, M.alpha) * inv4;
  }
  const D pdf_spec = vh.v > 0.0f ? d_ggx_g1(nv, M.alpha) * d_ggx_d(n, h, nh, M.alpha) * inv4 : dk(0.0f);
  return ps * pdf_spec + (1.0f - ps) * nl * INV_PI_F;
}
__device__ __forceinline__ void d_basis(const DV &n, DV &t, DV &b) {
  const float s = n.z.v >= 0.0f ? 1.0f : -1.0f;
  const D a = -rcp(s + n.z);
  const D bb = n.x * n.y * a;
  t = {1.0f + s * (n.x * n.x * a), s * bb, -s * n.x};
  b = {bb, s + n.y * n.y * a, -n.y};
}
__device__ __forceinline__ DV d_sample_vndf(const DV &ve, float a, float u1, float u2) {
  const DV vh = dnormalize(DV{a * ve.x, a * ve.y, ve.z});
  const D lensq = vh.x * vh.x + vh.y * vh.y;
  DV T1 = dv(mk(1.f, 0.f, 0.f));
  if (lensq.v > 0.0f) {
    const D il = rcp(dsqrt(lensq));
    T1 = {-vh.y * il, vh.x * il, dk(0.0f)};
  }
  const DV T2 = dcross(vh, T1);
  const float r = sqrtf(u1), phi = 2.0f * PI_F * u2;
  float sn, cs;
  sincosf(phi, &sn, &cs);
  const float t1 = r * cs, t2r = r * sn;
  const D s = 0.5f * (1.0f + vh.z);
  const D t2 = (1.0f - s) * sqrtf(fmaxf(0.0f, 1.0f - t1 * t1)) + s * t2r;
  const D t3 = dsqrt(dmax(1.0f - t1 * t1 - t2 * t2, 0.0f));
  const DV nh = T1 * dk(t1) + T2 * t2 + vh * t3;
  D nz = nh.z.v >= 1e-7f ? nh.z : dk(1e-7f);
  return dnormalize(DV{a * nh.x, a * nh.y, nz});
}
// sample_bsdf on tangent values: the same lobe choice (made on the value), l and its pdf
__device__ __forceinline__ D d_sample_bsdf(const Mat &M, const DV &n, const DV &v, float u0, float u1, float u2,
                                           DV &l) {
  const D nv = ddot(n, v);
  if (nv.v <= 0.0f) return dk(0.0f);
  DV t, b;
  d_basis(n, t, b);
  const D ps = d_spec_prob(M, nv);
  if (M.kind == 2 || (M.kind != 3 && u0 < ps.v)) {
    const float a = M.kind == 2 ? (u0 < ps.v ? M.alpha_c : M.alpha) : M.alpha;
    const DV hl = d_sample_vndf(DV{ddot(t, v), ddot(b, v), nv}, a, u1, u2);
    const DV h = t * hl.x + b * hl.y + n * hl.z;
    l = h * (2.0f * ddot(v, h)) - v;
  } else {
    const float r = sqrtf(u1), phi = 2.0f * PI_F * u2;
    float sn, cs;
    sincosf(phi, &sn, &cs);
    l = t * dk(r * cs) + b * dk(r * sn) + n * dk(sqrtf(fmaxf(0.0f, 1.0f - u1)));
  }
  if (ddot(n, l).v <= 0.0f) return dk(0.0f);
  return d_pdf_bsdf(M, n, v, l);
}

}  // namespace vt
