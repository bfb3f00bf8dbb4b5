/* Oracle (test infrastructure only): the sensor model of gpu/wgsl/sensor.wgsl.js:46-208 and gpu/sensor.js.
 *
 * Evaluated in binary32 with the WGSL's operation order and no contraction, so the GPU's fused epilogue can
 * be compared bit for bit. The reference uses the WGSL built-ins exp/log/cos, whose precision the WebGPU
 * implementation defines; this build states its own binary32 routines from published methods:
 *   exp: Cody & Waite (1980) argument reduction by ln 2 (two-part constant), degree-7 Taylor polynomial;
 *   log: exponent/mantissa split, mantissa in [sqrt(1/2), sqrt(2)), the atanh series
 *        log m = 2 (s + s^3/3 + ... + s^11/11), s = (m - 1)/(m + 1);
 *   cos: Cody & Waite reduction by pi/2 (three-part constant), Taylor polynomials of sin and cos on [-pi/4, pi/4].
 * Pins (tests/test_oracle_pins.py): the routines against libm (max error in ulp over dense grids); with noise
 * off, DN equals the closed form clamp(floor(e cg gain + black + 0.5), 0, 2^bits - 1); with noise on, the
 * per-level mean and variance match gpu/sensor.js:180-182 (rawNoiseDN); Poisson inversion against scipy's pmf.
 */
#include "vt_oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* 2^n for n in [-126, 127] */
static inline float pow2i(int n) { return u2f((uint32_t)(n + 127) << 23); }

float or_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return INFINITY;
  if (x < -103.0f) return 0.0f;
  float fn = floorf(x * 1.44269504088896341f + 0.5f);
  int n = (int)fn;
  float r = x - fn * 0.693359375f;
  r = r - fn * -2.12194440e-4f;
  float p = 1.98412698e-4f;             /* 1/5040 */
  p = p * r + 1.38888889e-3f;           /* 1/720 */
  p = p * r + 8.33333333e-3f;           /* 1/120 */
  p = p * r + 4.16666667e-2f;           /* 1/24 */
  p = p * r + 1.66666667e-1f;           /* 1/6 */
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  int n1 = n / 2, n2 = n - n1;
  return (p * pow2i(n1)) * pow2i(n2);
}

float or_logf(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  int e = 0;
  if (x < 1.17549435e-38f) { x = x * 16777216.0f; e = -24; }
  uint32_t u = f2u(x);
  e += (int)((u >> 23) & 0xff) - 127;
  float m = u2f((u & 0x007fffffu) | 0x3f800000u);       /* [1, 2) */
  if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
  float s = (m - 1.0f) / (m + 1.0f);
  float z = s * s;
  float q = 9.09090909e-2f;              /* 1/11 */
  q = q * z + 1.11111111e-1f;            /* 1/9 */
  q = q * z + 1.42857143e-1f;            /* 1/7 */
  q = q * z + 2.0e-1f;                   /* 1/5 */
  q = q * z + 3.33333333e-1f;            /* 1/3 */
  q = q * z;
  float lm = s + s;
  lm = lm + lm * q;
  float fe = (float)e;
  return fe * 0.693359375f + (fe * -2.12194440e-4f + lm);
}

float or_cosf(float x) {
  float ax = fabsf(x);
  float fj = floorf(ax * 0.636619772367581343f + 0.5f);
  int j = (int)fj;
  float r = ax - fj * 1.5703125f;
  r = r - fj * 4.837512969970703125e-4f;
  r = r - fj * 7.54978995489188216e-8f;
  float z = r * r;
  float c = -2.75573192e-7f;             /* -1/10! */
  c = c * z + 2.48015873e-5f;            /* 1/8! */
  c = c * z + -1.38888889e-3f;           /* -1/6! */
  c = c * z + 4.16666667e-2f;            /* 1/4! */
  c = c * z + -0.5f;
  c = c * z + 1.0f;
  float s = 2.75573192e-6f;              /* 1/9! */
  s = s * z + -1.98412698e-4f;           /* -1/7! */
  s = s * z + 8.33333333e-3f;            /* 1/5! */
  s = s * z + -1.66666667e-1f;           /* -1/3! */
  s = s * z;
  s = r + r * s;
  switch (j & 3) {
    case 0: return c;
    case 1: return -s;
    case 2: return -c;
    default: return s;
  }
}

float or_erfa(float z) {
  float az = fabsf(z);
  float t = 1.0f / (1.0f + 0.3275911f * az);
  float p = 1.061405429f * t - 1.453152027f;
  p = p * t;
  p = p + 1.421413741f;
  p = p * t - 0.284496736f;
  p = p * t + 0.254829592f;
  p = p * t;
  float y = 1.0f - p * or_expf(-z * z);
  return z >= 0.0f ? y : -y;
}

static int gauss_taps(float sg) {
  float c = ceilf(3.0f * sg + 0.5f);
  int r = (int)c;
  return r < 0 ? 0 : r > 12 ? 12 : r;
}
static float tap_w(int k, float sg) {
  if (sg <= 0.0f) return k == 0 ? 1.0f : 0.0f;
  float s2 = sg * 1.41421356f;
  float a = or_erfa(((float)k + 0.5f) / s2), b = or_erfa(((float)k - 0.5f) / s2);
  return 0.5f * (a - b);
}
static inline float us01(uint32_t x) { return ((float)(x >> 8u) + 0.5f) * (1.0f / 16777216.0f); }
static float gauss_h(uint32_t h) {
  float u1 = us01(or_pcg(h)), u2 = us01(or_pcg(h ^ 0x68bc21ebu));
  return sqrtf(-2.0f * or_logf(u1)) * or_cosf(6.283185307f * u2);
}
static float poisson_h(float lam, uint32_t h) {
  if (lam <= 0.0f) return 0.0f;
  if (lam > 64.0f) {
    float v = floorf((lam + sqrtf(lam) * gauss_h(h)) + 0.5f);
    return v > 0.0f ? v : 0.0f;
  }
  float u = us01(or_pcg(h ^ 0x1b873593u));
  float p = or_expf(-lam), F = p, k = 0.0f;
  for (int i = 0; i < 200; i++) {
    if (u <= F) break;
    k = k + 1.0f;
    p = (p * lam) / k;
    F = F + p;
  }
  return k;
}
static inline int cfa_channel(uint32_t x, uint32_t y, uint32_t code) {
  static const int pat[4][4] = {{0, 1, 1, 2}, {1, 0, 2, 1}, {1, 2, 0, 1}, {2, 1, 1, 0}};
  return pat[code & 3][(y & 1u) * 2u + (x & 1u)];
}

/* veiling glare (sensor.wgsl.js:87-127): block means on the 1/F level (at most 32 x 32 samples per block), then a
 * Gaussian along rows and then columns of the level, zero outside it, divided by sigma sqrt(2 pi) */
void or_glare_level(const or_sensor *s, const float *L, const float *ri, float *G) {
  const int W = (int)s->width, H = (int)s->height, F = (int)s->glare_f;
  const int gw = (W + F - 1) / F, gh = (H + F - 1) / F, fl = F < 32 ? F : 32;
  float *T = malloc(sizeof(float) * 3 * (size_t)gw * gh);
  for (int y = 0; y < gh; y++)
    for (int x = 0; x < gw; x++) {
      float a[3] = {0, 0, 0}, n = 0.0f;
      for (int j = 0; j < fl; j++)
        for (int i = 0; i < fl; i++) {
          const int xx = x * F + i, yy = y * F + j;
          if (xx >= W || yy >= H) continue;
          const float *p = L + 4 * ((size_t)yy * W + xx);
          const float q = ri[(size_t)yy * W + xx];
          for (int c = 0; c < 3; c++) a[c] = a[c] + p[c] * q;
          n = n + 1.0f;
        }
      const float sc = s->optics_k / (n > 1.0f ? n : 1.0f);
      for (int c = 0; c < 3; c++) G[3 * ((size_t)y * gw + x) + c] = a[c] * sc;
    }
  const float sg = s->glare_sigma > 1e-3f ? s->glare_sigma : 1e-3f;
  const float nrm = sg * 2.5066282746f;
  for (int pass = 0; pass < 2; pass++) {      /* rows (G -> T), then columns (T -> G) */
    const float *in = pass == 0 ? G : T;
    float *out = pass == 0 ? T : G;
    const int n = pass == 0 ? gw : gh, lim = n < 256 ? n : 256;
    for (int y = 0; y < gh; y++)
      for (int x = 0; x < gw; x++) {
        float a[3] = {0, 0, 0};
        for (int k = 0; k < lim; k++) {
          const float d = (float)(k - (pass == 0 ? x : y));
          const float w = or_expf(-0.5f * d * d / (sg * sg));
          const float *c = in + 3 * (pass == 0 ? (size_t)y * gw + k : (size_t)k * gw + x);
          for (int ch = 0; ch < 3; ch++) a[ch] = a[ch] + w * c[ch];
        }
        for (int ch = 0; ch < 3; ch++) out[3 * ((size_t)y * gw + x) + ch] = a[ch] / nrm;
      }
  }
  free(T);
}
static inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

void or_sensor_expose(const or_sensor *s, const float *L, const float *ri, uint16_t *dn) {
  const int W = (int)s->width, H = (int)s->height;
  const float sg = s->esf_sigma;
  const int r = gauss_taps(sg);
  float wt[25];
  for (int k = -r; k <= r; k++) wt[k + r] = tap_w(k, sg);
  float *A = malloc(sizeof(float) * 3 * (size_t)W * H);
  /* sensorOptics: horizontal pass of the radiance times RI (sensor.wgsl.js:68-84) */
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      float a0 = 0, a1 = 0, a2 = 0, ws = 0;
      for (int k = -r; k <= r; k++) {
        int xx = x + k; xx = xx < 0 ? 0 : xx > W - 1 ? W - 1 : xx;
        float w = wt[k + r];
        const float *p = L + 4 * ((size_t)y * W + xx);
        float q = ri[(size_t)y * W + xx];
        a0 = a0 + (w * p[0]) * q; a1 = a1 + (w * p[1]) * q; a2 = a2 + (w * p[2]) * q;
        ws = ws + w;
      }
      float sc = s->optics_k / ws;
      float *o = A + 3 * ((size_t)y * W + x);
      o[0] = a0 * sc; o[1] = a1 * sc; o[2] = a2 * sc;
    }
  const uint32_t hs = or_pcg((s->frame * 0x85ebca6bu) ^ or_pcg(s->seed ^ (s->cam_index * 0xc2b2ae35u)));
  float *G = NULL;
  int gw = 0, gh = 0;
  if (s->glare_k > 0.0f) {
    gw = (W + (int)s->glare_f - 1) / (int)s->glare_f;
    gh = (H + (int)s->glare_f - 1) / (int)s->glare_f;
    G = malloc(sizeof(float) * 3 * (size_t)gw * gh);
    or_glare_level(s, L, ri, G);
  }
  /* sensorExpose (sensor.wgsl.js:177-208); glare read bilinearly from the level (glareAt :129-141) */
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      float e0 = 0, e1 = 0, e2 = 0, ws = 0;
      for (int k = -r; k <= r; k++) {
        int yy = y + k; yy = yy < 0 ? 0 : yy > H - 1 ? H - 1 : yy;
        float w = wt[k + r];
        const float *p = A + 3 * ((size_t)yy * W + x);
        e0 = e0 + w * p[0]; e1 = e1 + w * p[1]; e2 = e2 + w * p[2];
        ws = ws + w;
      }
      e0 = e0 / ws; e1 = e1 / ws; e2 = e2 / ws;
      if (G) {
        const float F = (float)s->glare_f;
        const float px = ((float)x + 0.5f) / F - 0.5f, py = ((float)y + 0.5f) / F - 0.5f;
        const float fx0 = floorf(px), fy0 = floorf(py), fx = px - fx0, fy = py - fy0;
        int xs[2] = {(int)fx0, (int)fx0 + 1}, ys[2] = {(int)fy0, (int)fy0 + 1};
        for (int k = 0; k < 2; k++) {
          xs[k] = xs[k] < 0 ? 0 : xs[k] > gw - 1 ? gw - 1 : xs[k];
          ys[k] = ys[k] < 0 ? 0 : ys[k] > gh - 1 ? gh - 1 : ys[k];
        }
        float g[3];
        for (int c = 0; c < 3; c++) {
          const float a = G[3 * ((size_t)ys[0] * gw + xs[0]) + c], b = G[3 * ((size_t)ys[0] * gw + xs[1]) + c];
          const float cc = G[3 * ((size_t)ys[1] * gw + xs[0]) + c], d = G[3 * ((size_t)ys[1] * gw + xs[1]) + c];
          g[c] = mixf(mixf(a, b, fx), mixf(cc, d, fx), fy);
        }
        const float k = s->glare_k;
        e0 = (1.0f - k) * e0 + k * g[0];
        e1 = (1.0f - k) * e1 + k * g[1];
        e2 = (1.0f - k) * e2 + k * g[2];
      }
      int ch = cfa_channel((uint32_t)x, (uint32_t)y, s->cfa);
      const float *row = s->m + 3 * ch;
      float dt = (row[0] * e0 + row[1] * e1) + row[2] * e2;
      float lam = (dt > 0.0f ? dt : 0.0f) * s->exposure + s->dark * s->exposure;
      uint32_t h = or_pcg(((uint32_t)(y * W + x) * 0x9E3779B9u) ^ hs);
      float e = lam;
      if (s->noise_on > 0.5f) e = poisson_h(lam, h) + s->read_noise * gauss_h(h ^ 0x27d4eb2fu);
      e = e < s->full_well ? e : s->full_well;
      float v = floorf(((e * s->conv_gain) * s->gain + s->black) + 0.5f);
      v = v < 0.0f ? 0.0f : v > s->max_dn ? s->max_dn : v;
      dn[(size_t)y * W + x] = (uint16_t)v;
    }
  free(G);
  free(A);
}

static double vignetting(int vn, const double *vt, const double *vv, double th) {
  if (vn <= 0) return 1;
  if (th <= vt[0]) return vv[0];
  for (int k = 1; k < vn; k++)
    if (th <= vt[k]) return vv[k - 1] + (vv[k] - vv[k - 1]) * (th - vt[k - 1]) / (vt[k] - vt[k - 1]);
  return vv[vn - 1];
}

void or_ri_map(const or_lens *L, int w, int h, double q, int vn, const double *vt, const double *vv, float *out) {
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      double uv[2] = {x, y}, r[3];
      or_unproject(L, uv, r);
      double c = r[2] < -1 ? -1 : r[2] > 1 ? 1 : r[2];
      double th = acos(c);
      double ri = 1;
      if (th > 1e-6) ri = pow(cos(th), q) * sin(th) / (or_distort(L, th) * or_distort_deriv(L, th));
      out[(size_t)y * w + x] = (float)((ri > 0 ? ri : 0) * vignetting(vn, vt, vv, th));
    }
}

double or_raw_noise_dn(double e, double read_noise, double conv_gain, double gain) {
  return sqrt((e > 0 ? e : 0) + read_noise * read_noise) * conv_gain * gain;
}
