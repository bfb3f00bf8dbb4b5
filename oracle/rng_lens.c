/* Oracle (test infrastructure only): seeded streams and the KB4 lens.
 *
 * mulberry32 follows control/rng.js:4-13 word for word in uint32 arithmetic (Math.imul == mod-2^32 product);
 * pinned by golden vectors made by running control/rng.js under Node (tests/golden/make_golden.mjs).
 * Rng.normal (control/rng.js:19-27) uses host libm here; V8's Math.log/sin/cos may differ in the last ulp, so
 * normals are pinned to 1e-12 relative, not bit for bit.
 * pcg/hash3/u01/kr follow gpu/wgsl/common.wgsl.js:94-113 (u32 exact; kr in binary32, no contraction).
 * The lens follows gpu/lens.js:19-83 in FP64 and gpu/wgsl/lens.wgsl.js:9-29 in binary32; pinned by the
 * closed-form project/unproject round trip, the equidistant k = 0 case and maxMonotoneTheta's definition.
 */
#include "vt_oracle.h"
#include <math.h>

uint32_t or_mulberry32_u32(uint32_t *state) {
  uint32_t a = *state + 0x6d2b79f5u;
  *state = a;
  uint32_t t = a;
  t = (t ^ (t >> 15)) * (t | 1u);
  t ^= t + (t ^ (t >> 7)) * (t | 61u);
  return t ^ (t >> 14);
}
double or_mulberry32_uniform(uint32_t *state) { return (double)or_mulberry32_u32(state) / 4294967296.0; }
void or_mulberry32_fill(uint32_t seed, double *out, size_t n) {
  uint32_t s = seed;
  for (size_t i = 0; i < n; i++) out[i] = or_mulberry32_uniform(&s);
}
void or_rng_init(or_rng *r, uint32_t seed) { r->a = seed; r->has_spare = 0; r->spare = 0; }
double or_rng_uniform(or_rng *r, double lo, double hi) { return lo + (hi - lo) * or_mulberry32_uniform(&r->a); }
int64_t or_rng_int(or_rng *r, int64_t lo, int64_t hi) {
  return lo + (int64_t)floor(or_mulberry32_uniform(&r->a) * (double)(hi - lo + 1));
}
double or_rng_normal(or_rng *r, double mu, double sigma) {
  if (r->has_spare) { r->has_spare = 0; return mu + sigma * r->spare; }
  double u = 0, v;
  while (u <= 1e-300) u = or_mulberry32_uniform(&r->a);
  v = or_mulberry32_uniform(&r->a);
  double rr = sqrt(-2 * log(u));
  r->spare = rr * sin(2 * M_PI * v);
  r->has_spare = 1;
  return mu + sigma * rr * cos(2 * M_PI * v);
}

uint32_t or_pcg(uint32_t v) {
  uint32_t s = v * 747796405u + 2891336453u;
  uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (w >> 22u) ^ w;
}
uint32_t or_hash3(uint32_t a, uint32_t b, uint32_t c) { return or_pcg(a ^ or_pcg(b ^ or_pcg(c))); }
float or_u01(uint32_t x) { return (float)(x >> 8u) * (1.0f / 16777216.0f); }
float or_kr(uint32_t i, float a, float shift) {
  volatile float p = (float)i * a;   /* binary32 product, then sum: the WGSL order, never fused */
  float s = shift + p;
  return s - floorf(s);
}

/* ------------------------------------------------------------------ KB4, FP64 */
double or_distort(const or_lens *L, double th) {
  double t2 = th * th;
  return th * (1 + t2 * (L->k[0] + t2 * (L->k[1] + t2 * (L->k[2] + t2 * L->k[3]))));
}
double or_distort_deriv(const or_lens *L, double th) {
  double t2 = th * th;
  return 1 + t2 * (3 * L->k[0] + t2 * (5 * L->k[1] + t2 * (7 * L->k[2] + t2 * 9 * L->k[3])));
}
double or_undistort_theta(const or_lens *L, double thd, int iters) {
  double th = thd;
  for (int i = 0; i < iters; i++) {
    double f = or_distort(L, th) - thd;
    double step = f / or_distort_deriv(L, th);
    th -= step;
    if (fabs(step) < 1e-15) break;
  }
  return th;
}
void or_project(const or_lens *L, const double X[3], double uv[2]) {
  double x = X[0], y = X[1], z = X[2];
  double rho = hypot(x, y);
  if (rho < 1e-300) { uv[0] = L->cx; uv[1] = L->cy; return; }
  double th = atan2(rho, z);
  double s = or_distort(L, th) / rho;
  double xd = x * s, yd = y * s;
  uv[0] = L->fx * (xd + L->alpha * yd) + L->cx;
  uv[1] = L->fy * yd + L->cy;
}
void or_unproject(const or_lens *L, const double uv[2], double ray[3]) {
  double yd = (uv[1] - L->cy) / L->fy;
  double xd = (uv[0] - L->cx) / L->fx - L->alpha * yd;
  double thd = hypot(xd, yd);
  if (thd < 1e-300) { ray[0] = 0; ray[1] = 0; ray[2] = 1; return; }
  double th = or_undistort_theta(L, thd, 20);
  double s = sin(th) / thd;
  ray[0] = xd * s; ray[1] = yd * s; ray[2] = cos(th);
}
double or_max_monotone_theta(const or_lens *L, double limit) {
  const int n = 4096;
  for (int i = 1; i <= n; i++) {
    double th = ((double)i / n) * limit;
    if (or_distort_deriv(L, th) <= 0) return ((double)(i - 1) / n) * limit;
  }
  return limit;
}
double or_max_image_theta(const or_lens *L, int w, int h) {
  double us[3] = {-0.5, (w - 1) / 2.0, w - 0.5}, vs[3] = {-0.5, (h - 1) / 2.0, h - 0.5};
  double m = 0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double uv[2] = {us[a], vs[b]}, r[3];
      or_unproject(L, uv, r);
      double c = r[2] < -1 ? -1 : r[2] > 1 ? 1 : r[2];
      double th = acos(c);
      if (th > m) m = th;
    }
  return m;
}

/* binary32 KB4 inverse as gpu/wgsl/lens.wgsl.js:9-29 (a = fx fy cx cy, b = k1..k4, c = alpha) */
void or_unproject_f32(const float lens[9], const float uv[2], float ray[3]) {
  const float fx = lens[0], fy = lens[1], cx = lens[2], cy = lens[3];
  const float k1 = lens[4], k2 = lens[5], k3 = lens[6], k4 = lens[7], al = lens[8];
  float yd = (uv[1] - cy) / fy;
  volatile float ay = al * yd;
  float xd = (uv[0] - cx) / fx - ay;
  volatile float xx = xd * xd, yy = yd * yd;
  float thd = sqrtf(xx + yy);
  if (thd < 1e-20f) { ray[0] = 0; ray[1] = 0; ray[2] = 1; return; }
  float th = thd;
  for (int i = 0; i < 8; i++) {
    volatile float t2 = th * th;
    volatile float p = k3 + t2 * k4; p = k2 + t2 * p; p = k1 + t2 * p; p = 1.0f + t2 * p;
    volatile float f = th * p; f = f - thd;
    volatile float t9 = t2 * 9.0f; volatile float q = 7.0f * k3 + t9 * k4; q = 5.0f * k2 + t2 * q; q = 3.0f * k1 + t2 * q; q = 1.0f + t2 * q;
    th = th - f / q;
  }
  float s = (float)sin((double)th) / thd;
  ray[0] = xd * s; ray[1] = yd * s; ray[2] = (float)cos((double)th);
}
