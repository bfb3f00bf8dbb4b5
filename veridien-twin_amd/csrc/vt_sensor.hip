// The sensor epilogue of gpu/wgsl/sensor.wgsl.js:46-208 as ONE kernel: per 32x8 output tile the horizontal
// optics pass (radiance x RI, pixel-integrated Gaussian taps) is computed once into LDS for the rows the
// vertical taps need, then the vertical pass, CFA response, Poisson shot + Gaussian read noise, full-well clip
// and the 12-bit ADC write a uint16 DN. HBM traffic: the radiance sums and RI map in, 2 B/pixel out.
//
// Integer DNs match the host sensor model bit for bit: every binary32 operation is unfused (this file is built
// with -ffp-contract=off), in the WGSL's order, and exp/log/cos are this library's own routines (published
// methods, stated below) instead of device intrinsics whose last-ulp behaviour differs from the host's.
#include <mutex>

#include "vt_internal.h"

namespace vt {
namespace {

__device__ __forceinline__ float pow2i(int n) { return __uint_as_float((uint32_t)(n + 127) << 23); }

// exp: Cody & Waite (1980) reduction by ln 2 (two-part constant), degree-7 Taylor polynomial, scale by 2^n
__device__ float vt_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return INFINITY;
  if (x < -103.0f) return 0.0f;
  float fn = floorf(x * 1.44269504088896341f + 0.5f);
  int n = (int)fn;
  float r = x - fn * 0.693359375f;
  r = r - fn * -2.12194440e-4f;
  float p = 1.98412698e-4f;
  p = p * r + 1.38888889e-3f;
  p = p * r + 8.33333333e-3f;
  p = p * r + 4.16666667e-2f;
  p = p * r + 1.66666667e-1f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  int n1 = n / 2, n2 = n - n1;
  return (p * pow2i(n1)) * pow2i(n2);
}

// log: exponent split, mantissa in [sqrt(1/2), sqrt(2)), atanh series 2(s + s^3/3 + ... + s^11/11)
__device__ float vt_logf(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  int e = 0;
  if (x < 1.17549435e-38f) { x = x * 16777216.0f; e = -24; }
  uint32_t u = __float_as_uint(x);
  e += (int)((u >> 23) & 0xffu) - 127;
  float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
  float s = (m - 1.0f) / (m + 1.0f);
  float z = s * s;
  float q = 9.09090909e-2f;
  q = q * z + 1.11111111e-1f;
  q = q * z + 1.42857143e-1f;
  q = q * z + 2.0e-1f;
  q = q * z + 3.33333333e-1f;
  q = q * z;
  float lm = s + s;
  lm = lm + lm * q;
  float fe = (float)e;
  return fe * 0.693359375f + (fe * -2.12194440e-4f + lm);
}

// cos: Cody & Waite reduction by pi/2 (three-part constant), Taylor polynomials on [-pi/4, pi/4]
__device__ float vt_cosf(float x) {
  float ax = fabsf(x);
  float fj = floorf(ax * 0.636619772367581343f + 0.5f);
  int j = (int)fj;
  float r = ax - fj * 1.5703125f;
  r = r - fj * 4.837512969970703125e-4f;
  r = r - fj * 7.54978995489188216e-8f;
  float z = r * r;
  float c = -2.75573192e-7f;
  c = c * z + 2.48015873e-5f;
  c = c * z + -1.38888889e-3f;
  c = c * z + 4.16666667e-2f;
  c = c * z + -0.5f;
  c = c * z + 1.0f;
  float s = 2.75573192e-6f;
  s = s * z + -1.98412698e-4f;
  s = s * z + 8.33333333e-3f;
  s = s * z + -1.66666667e-1f;
  s = s * z;
  s = r + r * s;
  switch (j & 3) {
    case 0: return c;
    case 1: return -s;
    case 2: return -c;
    default: return s;
  }
}

// Abramowitz & Stegun 7.1.26 (sensor.wgsl.js:57-61)
__device__ float vt_erfa(float z) {
  float az = fabsf(z);
  float t = 1.0f / (1.0f + 0.3275911f * az);
  float p = 1.061405429f * t - 1.453152027f;
  p = p * t;
  p = p + 1.421413741f;
  p = p * t - 0.284496736f;
  p = p * t + 0.254829592f;
  p = p * t;
  float y = 1.0f - p * vt_expf(-z * z);
  return z >= 0.0f ? y : -y;
}

__device__ __forceinline__ int gauss_taps(float sg) {
  int r = (int)ceilf(3.0f * sg + 0.5f);
  return r < 0 ? 0 : (r > 12 ? 12 : r);
}
__device__ float tap_w(int k, float sg) {
  if (sg <= 0.0f) return k == 0 ? 1.0f : 0.0f;
  float s2 = sg * 1.41421356f;
  float a = vt_erfa(((float)k + 0.5f) / s2), b = vt_erfa(((float)k - 0.5f) / s2);
  return 0.5f * (a - b);
}
__device__ __forceinline__ float us01(uint32_t x) { return ((float)(x >> 8u) + 0.5f) * (1.0f / 16777216.0f); }
__device__ float gauss_h(uint32_t h) {
  float u1 = us01(pcg(h)), u2 = us01(pcg(h ^ 0x68bc21ebu));
  return sqrtf(-2.0f * vt_logf(u1)) * vt_cosf(6.283185307f * u2);
}
// Poisson by sequential-search inversion up to 64 e-, rounded normal above (sensor.wgsl.js:160-175)
__device__ float poisson_h(float lam, uint32_t h) {
  if (lam <= 0.0f) return 0.0f;
  if (lam > 64.0f) {
    float v = floorf((lam + sqrtf(lam) * gauss_h(h)) + 0.5f);
    return v > 0.0f ? v : 0.0f;
  }
  float u = us01(pcg(h ^ 0x1b873593u));
  float p = vt_expf(-lam), F = p, k = 0.0f;
  for (int i = 0; i < 200; i++) {
    if (u <= F) break;
    k = k + 1.0f;
    p = (p * lam) / k;
    F = F + p;
  }
  return k;
}
__device__ __forceinline__ int cfa_channel(uint32_t x, uint32_t y, uint32_t code) {
  const uint32_t k = (y & 1u) * 2u + (x & 1u);
  // RGGB, GRBG, GBRG, BGGR, row-major 2x2, 0 R 1 G 2 B
  const uint32_t pat = code == 0 ? 0x2110u : code == 1 ? 0x1201u : code == 2 ? 0x1021u : 0x0112u;
  return (int)((pat >> (4u * k)) & 0xfu);
}

constexpr int TW = 32, TH = 8, RMAX = 12;

// a pixel's mean radiance factor: the output scale over its sample count (the refinement pass gives some pixels
// more samples than others)
__device__ __forceinline__ float pix_scale(const SensorDev &P, float count) {
  return count > 0.0f ? P.scale / count : 0.0f;
}

// ------------------------------------------------------------------ veiling glare on the 1/F level
// (sensor.wgsl.js:87-141): F x F block means of the irradiance, then a Gaussian along each row and each column of the
// level (zero outside the sensor, normalised on the infinite line), read back bilinearly in the exposure
__global__ void k_glare_down(SensorDev P, const float4 *acc, const float *ri, float4 *out) {
  const int W = (int)P.w, H = (int)P.h, F = (int)P.glare_f;
  const int gw = (W + F - 1) / F, gh = (H + F - 1) / F;
  const int x = blockIdx.x * 8 + threadIdx.x, y = blockIdx.y * 8 + threadIdx.y;
  if (x >= gw || y >= gh) return;
  const int fl = F < 32 ? F : 32;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, n = 0.0f;
  for (int j = 0; j < fl; j++)
    for (int i = 0; i < fl; i++) {
      const int xx = x * F + i, yy = y * F + j;
      if (xx >= W || yy >= H) continue;
      const float4 s = acc[(size_t)yy * W + xx];
      const float q = ri[(size_t)yy * W + xx];
      const float inv = pix_scale(P, s.w);
      a0 = a0 + (s.x * inv) * q;
      a1 = a1 + (s.y * inv) * q;
      a2 = a2 + (s.z * inv) * q;
      n = n + 1.0f;
    }
  const float sc = P.optics_k / fmaxf(n, 1.0f);
  out[(size_t)y * gw + x] = make_float4(a0 * sc, a1 * sc, a2 * sc, 1.0f);
}
template <int DIR>
__global__ void k_glare_blur(SensorDev P, const float4 *in, float4 *out) {
  const int W = (int)P.w, H = (int)P.h, F = (int)P.glare_f;
  const int gw = (W + F - 1) / F, gh = (H + F - 1) / F;
  const int x = blockIdx.x * 8 + threadIdx.x, y = blockIdx.y * 8 + threadIdx.y;
  if (x >= gw || y >= gh) return;
  const float sg = fmaxf(P.glare_sigma, 1e-3f);
  const int n = DIR == 0 ? gw : gh, lim = n < 256 ? n : 256, at = DIR == 0 ? x : y;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
  for (int k = 0; k < lim; k++) {
    const float d = (float)(k - at);
    const float w = vt_expf(-0.5f * d * d / (sg * sg));
    const float4 c = in[DIR == 0 ? (size_t)y * gw + k : (size_t)k * gw + x];
    a0 = a0 + w * c.x;
    a1 = a1 + w * c.y;
    a2 = a2 + w * c.z;
  }
  const float nrm = sg * 2.5066282746f;
  out[(size_t)y * gw + x] = make_float4(a0 / nrm, a1 / nrm, a2 / nrm, 1.0f);
}
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ void glare_at(const SensorDev &P, const float4 *G, int x, int y, float &g0, float &g1, float &g2) {
  const int W = (int)P.w, H = (int)P.h, Fi = (int)P.glare_f;
  const int gw = (W + Fi - 1) / Fi, gh = (H + Fi - 1) / Fi;
  const float F = (float)P.glare_f;
  const float px = ((float)x + 0.5f) / F - 0.5f, py = ((float)y + 0.5f) / F - 0.5f;
  const float fx0 = floorf(px), fy0 = floorf(py);
  const int ix = (int)fx0, iy = (int)fy0;
  const float fx = px - fx0, fy = py - fy0;
  const int x0 = min(max(ix, 0), gw - 1), x1 = min(max(ix + 1, 0), gw - 1);
  const int y0 = min(max(iy, 0), gh - 1), y1 = min(max(iy + 1, 0), gh - 1);
  const float4 a = G[(size_t)y0 * gw + x0], b = G[(size_t)y0 * gw + x1], c = G[(size_t)y1 * gw + x0],
               d = G[(size_t)y1 * gw + x1];
  g0 = mixf(mixf(a.x, b.x, fx), mixf(c.x, d.x, fx), fy);
  g1 = mixf(mixf(a.y, b.y, fx), mixf(c.y, d.y, fx), fy);
  g2 = mixf(mixf(a.z, b.z, fx), mixf(c.z, d.z, fx), fy);
}

__global__ __launch_bounds__(TW *TH) void k_sensor(SensorDev P, const float4 *acc, const float *ri,
                                                  const float4 *glare, uint16_t *dn) {
  __shared__ float sA[3][(TH + 2 * RMAX) * TW];
  __shared__ float sw[2 * RMAX + 1];
  const int W = (int)P.w, H = (int)P.h;
  const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
  const int tid = threadIdx.y * TW + threadIdx.x;
  const float sg = P.esf_sigma;
  const int r = gauss_taps(sg);
  if (tid <= 2 * r) sw[tid] = tap_w(tid - r, sg);
  __syncthreads();
  float wsum = 0.0f;
  for (int k = 0; k <= 2 * r; k++) wsum = wsum + sw[k];
  // horizontal pass (sensorOptics) for rows y0 - r .. y0 + TH - 1 + r, each stored at its clamped row
  const int rows = TH + 2 * r;
  for (int idx = tid; idx < rows * TW; idx += TW * TH) {
    const int j = idx / TW, c = idx % TW;
    const int x = x0 + c;
    int yy = y0 - r + j;
    yy = yy < 0 ? 0 : (yy > H - 1 ? H - 1 : yy);
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    if (x < W) {
      for (int k = -r; k <= r; k++) {
        int xx = x + k;
        xx = xx < 0 ? 0 : (xx > W - 1 ? W - 1 : xx);
        const float w = sw[k + r];
        const float4 s = acc[(size_t)yy * W + xx];
        const float q = ri[(size_t)yy * W + xx];
        const float inv = pix_scale(P, s.w);
        const float l0 = s.x * inv, l1 = s.y * inv, l2 = s.z * inv;
        a0 = a0 + (w * l0) * q;
        a1 = a1 + (w * l1) * q;
        a2 = a2 + (w * l2) * q;
      }
      const float sc = P.optics_k / wsum;
      a0 = a0 * sc; a1 = a1 * sc; a2 = a2 * sc;
    }
    sA[0][idx] = a0; sA[1][idx] = a1; sA[2][idx] = a2;
  }
  __syncthreads();
  const int x = x0 + (int)threadIdx.x, y = y0 + (int)threadIdx.y;
  if (x >= W || y >= H) return;
  // vertical pass + exposure (sensorExpose)
  float e0 = 0.0f, e1 = 0.0f, e2 = 0.0f;
  for (int k = -r; k <= r; k++) {
    const int j = (int)threadIdx.y + k + r;
    const float w = sw[k + r];
    e0 = e0 + w * sA[0][j * TW + threadIdx.x];
    e1 = e1 + w * sA[1][j * TW + threadIdx.x];
    e2 = e2 + w * sA[2][j * TW + threadIdx.x];
  }
  e0 = e0 / wsum; e1 = e1 / wsum; e2 = e2 / wsum;
  if (P.glare_k > 0.0f) {
    float g0, g1, g2;
    glare_at(P, glare, x, y, g0, g1, g2);
    const float k = P.glare_k;
    e0 = (1.0f - k) * e0 + k * g0;
    e1 = (1.0f - k) * e1 + k * g1;
    e2 = (1.0f - k) * e2 + k * g2;
  }
  const int ch = cfa_channel((uint32_t)x, (uint32_t)y, P.cfa);
  const float *row = P.m + 3 * ch;
  const float dt = (row[0] * e0 + row[1] * e1) + row[2] * e2;
  const float lam = (dt > 0.0f ? dt : 0.0f) * P.exposure + P.dark * P.exposure;
  const uint32_t hs = pcg((P.frame * 0x85ebca6bu) ^ pcg(P.seed ^ (P.cam_index * 0xc2b2ae35u)));
  const uint32_t h = pcg(((uint32_t)(y * W + x) * 0x9E3779B9u) ^ hs);
  float e = lam;
  if (P.noise_on) e = poisson_h(lam, h) + P.read_noise * gauss_h(h ^ 0x27d4eb2fu);
  e = e < P.full_well ? e : P.full_well;
  float v = floorf(((e * P.conv_gain) * P.gain + P.black) + 0.5f);
  v = v < 0.0f ? 0.0f : (v > P.max_dn ? P.max_dn : v);
  dn[(size_t)y * W + x] = (uint16_t)v;
}

// ------------------------------------------------------------------ the ISP after the raw DN
// (sensor.wgsl.js:210-427). 8-bit outputs equal the host model's bit for bit: the same unfused binary32 order, the
// library's own exp / log (pow = exp(y log x)).
__device__ __forceinline__ float isp_pow(float x, float y) { return x > 0.0f ? vt_expf(y * vt_logf(x)) : 0.0f; }

// raw DN mirrored at the borders with the CFA phase kept (reflect by 2) (rawAt :216-223)
__device__ __forceinline__ float raw_at(const IspDev &P, const uint16_t *dn, int x, int y) {
  const int W = (int)P.w, H = (int)P.h;
  int xx = x, yy = y;
  if (xx < 0) xx = -xx;
  if (xx >= W) xx = 2 * (W - 1) - xx;
  if (yy < 0) yy = -yy;
  if (yy >= H) yy = 2 * (H - 1) - yy;
  xx = xx < 0 ? 0 : (xx > W - 1 ? W - 1 : xx);
  yy = yy < 0 ? 0 : (yy > H - 1 ? H - 1 : yy);
  return (float)dn[(size_t)yy * W + xx];
}

// One block demosaics a 16 x 16 tile (demosaicG :229-256 and demosaicRB :263-319 fused): the raw tile with a
// 3-pixel mirrored halo and then the green plane with a 1-pixel halo are staged in LDS, so every DN is read from
// HBM once and the green plane never leaves the CU. Green: the sample at G sites; elsewhere the 4-neighbour mean
// (bilinear), the mean along the smaller green gradient (edge-aware), or Hamilton-Adams with the site's own second
// derivative. Red and blue: colour differences over the nearest sites (Hamilton-Adams) or the sites' plain means,
// then normalisation, the grey-world sums (x 1023, integers; one partial per block, no atomics), white balance and
// CCM -> linear camera RGB.
constexpr int DT = 16, DR = DT + 6, DG = DT + 2;

__device__ __forceinline__ unsigned int wave_sum(unsigned int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_isp_demosaic(IspDev P, const uint16_t *dn, float4 *rgb, uint4 *part) {
  __shared__ float sR[DR * DR];
  __shared__ float sG[DG * DG];
  __shared__ unsigned int sred[3][4];
  const int W = (int)P.w, H = (int)P.h;
  const int x0 = blockIdx.x * DT, y0 = blockIdx.y * DT;
  const int tid = threadIdx.y * DT + threadIdx.x;
  for (int i = tid; i < DR * DR; i += 256) sR[i] = raw_at(P, dn, x0 - 3 + i % DR, y0 - 3 + i / DR);
  __syncthreads();
  // raw_at of any coordinate within 3 of the tile
#define RAW(X, Y) sR[((Y) - (y0 - 3)) * DR + ((X) - (x0 - 3))]
  for (int i = tid; i < DG * DG; i += 256) {
    const int cx = min(max(x0 - 1 + i % DG, 0), W - 1), cy = min(max(y0 - 1 + i / DG, 0), H - 1);
    float v = RAW(cx, cy);
    if (cfa_channel((uint32_t)cx, (uint32_t)cy, P.cfa) != 1) {
      const float gl = RAW(cx - 1, cy), gr = RAW(cx + 1, cy), gu = RAW(cx, cy - 1), gd = RAW(cx, cy + 1);
      if (P.demosaic == 0) {
        v = 0.25f * (((gl + gr) + gu) + gd);
      } else if (P.demosaic == 1) {
        const float dh = fabsf(gl - gr), dv = fabsf(gu - gd);
        v = dh < dv ? 0.5f * (gl + gr) : (dv < dh ? 0.5f * (gu + gd) : 0.25f * (((gl + gr) + gu) + gd));
      } else {
        const float c = v;
        const float lh = (2.0f * c - RAW(cx - 2, cy)) - RAW(cx + 2, cy);
        const float lv = (2.0f * c - RAW(cx, cy - 2)) - RAW(cx, cy + 2);
        const float dh = fabsf(gl - gr) + fabsf(lh), dv = fabsf(gu - gd) + fabsf(lv);
        const float gh = 0.5f * (gl + gr) + 0.25f * lh, gv = 0.5f * (gu + gd) + 0.25f * lv;
        v = dh < dv ? gh : (dv < dh ? gv : 0.5f * (gh + gv));
        v = v < 0.0f ? 0.0f : (v > P.max_dn ? P.max_dn : v);
      }
    }
    sG[i] = v;
  }
  __syncthreads();
  // the green plane at an in-image coordinate within 1 of the tile
#define GRN(X, Y) sG[((Y) - (y0 - 1)) * DG + ((X) - (x0 - 1))]
  const int x = x0 + (int)threadIdx.x, y = y0 + (int)threadIdx.y;
  unsigned int q0 = 0u, q1 = 0u, q2 = 0u;
  if (x < W && y < H) {
    const int own = cfa_channel((uint32_t)x, (uint32_t)y, P.cfa);
    const float gv = GRN(x, y);
    float r, gg, b;
    if (P.demosaic == 2) {
      float sx = 0.0f, sy = 0.0f, cx = 0.0f, cy = 0.0f;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int xx = min(max(x + dx, 0), W - 1), yy = min(max(y + dy, 0), H - 1);
          const int ch = cfa_channel((uint32_t)xx, (uint32_t)yy, P.cfa);
          if (ch == 1) continue;
          if (ch == own && !(dx == 0 && dy == 0)) continue;
          const float d = RAW(xx, yy) - GRN(xx, yy);
          if (ch == 0) { sx += d; cx += 1.0f; } else { sy += d; cy += 1.0f; }
        }
      r = gv + sx / fmaxf(cx, 1.0f);
      gg = gv;
      b = gv + sy / fmaxf(cy, 1.0f);
    } else {
      float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int xx = min(max(x + dx, 0), W - 1), yy = min(max(y + dy, 0), H - 1);
          const int ch = cfa_channel((uint32_t)xx, (uint32_t)yy, P.cfa);
          if (ch == own && !(dx == 0 && dy == 0)) continue;
          const float v = RAW(xx, yy);
          if (ch == 0) { s0 += v; c0 += 1.0f; } else if (ch == 1) { s1 += v; c1 += 1.0f; } else { s2 += v; c2 += 1.0f; }
        }
      r = s0 / fmaxf(c0, 1.0f);
      gg = P.demosaic == 1 ? gv : s1 / fmaxf(c1, 1.0f);
      b = s2 / fmaxf(c2, 1.0f);
    }
#undef RAW
#undef GRN
    const float full = fmaxf(P.max_dn - P.black, 1.0f);
    float cr = (r - P.black) / full, cg = (gg - P.black) / full, cb = (b - P.black) / full;
    q0 = (unsigned int)rintf(fminf(fmaxf(cr, 0.0f), 1.0f) * 1023.0f);
    q1 = (unsigned int)rintf(fminf(fmaxf(cg, 0.0f), 1.0f) * 1023.0f);
    q2 = (unsigned int)rintf(fminf(fmaxf(cb, 0.0f), 1.0f) * 1023.0f);
    cr = cr * P.wb[0];
    cg = cg * P.wb[1];
    cb = cb * P.wb[2];
    const float o0 = (P.ccm[0] * cr + P.ccm[1] * cg) + P.ccm[2] * cb;
    const float o1 = (P.ccm[3] * cr + P.ccm[4] * cg) + P.ccm[5] * cb;
    const float o2 = (P.ccm[6] * cr + P.ccm[7] * cg) + P.ccm[8] * cb;
    rgb[(size_t)y * W + x] = make_float4(o0, o1, o2, 1.0f);
  }
  q0 = wave_sum(q0);
  q1 = wave_sum(q1);
  q2 = wave_sum(q2);
  if ((tid & 63) == 0) {
    sred[0][tid >> 6] = q0;
    sred[1][tid >> 6] = q1;
    sred[2][tid >> 6] = q2;
  }
  __syncthreads();
  if (tid == 0)
    part[blockIdx.y * gridDim.x + blockIdx.x] = make_uint4(0u, (sred[0][0] + sred[0][1]) + (sred[0][2] + sred[0][3]),
                                                           (sred[1][0] + sred[1][1]) + (sred[1][2] + sred[1][3]),
                                                           (sred[2][0] + sred[2][1]) + (sred[2][2] + sred[2][3]));
}

// tone curves (toneC :350-366): sRGB OETF, gamma, linear, or sRGB + a contrast S-curve with a black offset
__device__ float tone_c(const IspDev &P, float x) {
  if (P.tone == 3) {
    const float b = P.black_off;
    float v = (x - b) / (1.0f - b);
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    const float u = v <= 0.0031308f ? 12.92f * v : 1.055f * isp_pow(v, 1.0f / 2.4f) - 0.055f;
    const float a = isp_pow(u, P.s_c);
    return a / fmaxf(a + isp_pow(1.0f - u, P.s_c), 1e-12f);
  }
  const float c = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
  if (P.tone == 0) return c <= 0.0031308f ? 12.92f * c : 1.055f * isp_pow(c, 1.0f / 2.4f) - 0.055f;
  if (P.tone == 1) return isp_pow(c, 1.0f / fmaxf(P.gamma, 1e-3f));
  return c;
}

// 3x3 bilateral on linear RGB at an in-image pixel of the LDS tile: range on luma with sigma = scale x the modelled
// noise at the pixel (sensorNR :321-348); C(X, Y) is the RGB at any coordinate within 1 of it, clamped to the image
template <typename Fetch>
__device__ __forceinline__ float3 bilateral(const IspDev &P, int x, int y, Fetch C) {
  const int W = (int)P.w, H = (int)P.h;
  const float4 c0 = C(x, y);
  const float Y0 = fmaxf((c0.x * 0.2126f + c0.y * 0.7152f) + c0.z * 0.0722f, 0.0f);
  const float fullE = fmaxf(P.max_dn - P.black, 1.0f) / fmaxf(P.conv_gain * P.gain, 1e-9f);
  const float sn = sqrtf(Y0 * fullE + P.read_noise * P.read_noise) / fullE;
  const float rs = fmaxf(P.nr_scale * sn, 1e-6f);
  const float ss = P.nr_sigma;
  float ax = 0.0f, ay = 0.0f, az = 0.0f, ws = 0.0f;
  for (int dy = -1; dy <= 1; dy++)
    for (int dx = -1; dx <= 1; dx++) {
      const float4 c = C(min(max(x + dx, 0), W - 1), min(max(y + dy, 0), H - 1));
      const float dY = ((c.x - c0.x) * 0.2126f + (c.y - c0.y) * 0.7152f) + (c.z - c0.z) * 0.0722f;
      const float w = vt_expf(-0.5f * (float)(dx * dx + dy * dy) / (ss * ss) - 0.5f * dY * dY / (rs * rs));
      ax += w * c.x;
      ay += w * c.y;
      az += w * c.z;
      ws += w;
    }
  return make_float3(ax / ws, ay / ws, az / ws);
}

// One block finishes a 32 x 32 tile (sensorNR, toneC and sensorOutput :321-427 fused): the linear RGB with the
// halo the noise reduction and the unsharp box need is staged in LDS, each position is denoised and toned once,
// then one lane per 2x2 block applies the unsharp mask on luma, BT.601 full-range YCbCr, 4:2:0 chroma (2x2 mean)
// and the RGB rebuilt from it (or 4:4:4 RGB), writes RGBA and the Y / Cb / Cr planes and adds its luma to the
// block's partial sum.
constexpr int FT = 32, FSH = 4, FC = FT + 2 * FSH, FR = FC + 2;

__global__ __launch_bounds__(256) void k_isp_finish(IspDev P, const float4 *src, uchar4 *rgba, uint8_t *yuv,
                                                    uint4 *part) {
  __shared__ float4 sC[FR * FR];
  __shared__ float sT[3][FC * FC];
  __shared__ unsigned int sred[4];
  const int W = (int)P.w, H = (int)P.h;
  const int x0 = blockIdx.x * FT, y0 = blockIdx.y * FT;
  const int tid = threadIdx.y * 16 + threadIdx.x;
  const float amt = P.sharpen_amount;
  const int rad = P.sharpen_rad;
  const int R = amt > 0.0f ? rad : 0;       // toned halo
  const int hc = R + (P.nr ? 1 : 0);        // RGB halo
  // RGB at coordinates x0 - hc .. x0 + FT - 1 + hc (clamped to the image), stored at the tile's fixed origin
  const int nc = FT + 2 * hc;
  for (int i = tid; i < nc * nc; i += 256) {
    const int X = x0 - hc + i % nc, Y = y0 - hc + i / nc;
    sC[(Y - (y0 - FSH - 1)) * FR + (X - (x0 - FSH - 1))] =
        src[(size_t)min(max(Y, 0), H - 1) * W + min(max(X, 0), W - 1)];
  }
  __syncthreads();
  auto C = [&](int X, int Y) { return sC[(Y - (y0 - FSH - 1)) * FR + (X - (x0 - FSH - 1))]; };
  const int nt = FT + 2 * R;
  for (int i = tid; i < nt * nt; i += 256) {
    const int X = x0 - R + i % nt, Y = y0 - R + i / nt;
    const int cx = min(max(X, 0), W - 1), cy = min(max(Y, 0), H - 1);
    float3 c;
    if (P.nr) {
      c = bilateral(P, cx, cy, C);
    } else {
      const float4 v = C(cx, cy);
      c = make_float3(v.x, v.y, v.z);
    }
    const int k = (Y - (y0 - FSH)) * FC + (X - (x0 - FSH));
    sT[0][k] = tone_c(P, c.x);
    sT[1][k] = tone_c(P, c.y);
    sT[2][k] = tone_c(P, c.z);
  }
  __syncthreads();
#define TONED(X, Y, ch) sT[ch][((Y) - (y0 - FSH)) * FC + ((X) - (x0 - FSH))]
  const int bx = x0 + 2 * (int)threadIdx.x, by = y0 + 2 * (int)threadIdx.y;
  unsigned int lsum = 0;
  if (bx < W && by < H) {
    float Ys[4];
    float cb = 0.0f, cr = 0.0f, nb = 0.0f;
    for (int q = 0; q < 4; q++) {
      const int x = bx + (q & 1), y = by + (q >> 1);
      if (x >= W || y >= H) { Ys[q] = 0.0f; continue; }
      float r = TONED(x, y, 0), g = TONED(x, y, 1), b = TONED(x, y, 2);
      if (amt > 0.0f) {
        float m = 0.0f, n = 0.0f;
        for (int dy = -rad; dy <= rad; dy++)
          for (int dx = -rad; dx <= rad; dx++) {
            m += (TONED(x + dx, y + dy, 0) * 0.299f + TONED(x + dx, y + dy, 1) * 0.587f) +
                 TONED(x + dx, y + dy, 2) * 0.114f;
            n += 1.0f;
          }
        const float Y0 = (r * 0.299f + g * 0.587f) + b * 0.114f;
        const float d = amt * (Y0 - m / n);
        r += d; g += d; b += d;
      }
      const float r8 = fminf(fmaxf(r, 0.0f), 1.0f) * 255.0f, g8 = fminf(fmaxf(g, 0.0f), 1.0f) * 255.0f,
                  b8 = fminf(fmaxf(b, 0.0f), 1.0f) * 255.0f;
      Ys[q] = (0.299f * r8 + 0.587f * g8) + 0.114f * b8;
      cb += ((128.0f - 0.168736f * r8) - 0.331264f * g8) + 0.5f * b8;
      cr += ((128.0f + 0.5f * r8) - 0.418688f * g8) - 0.081312f * b8;
      nb += 1.0f;
      if (!P.chroma420) {
        const float y8 = fminf(fmaxf(rintf(Ys[q]), 0.0f), 255.0f);
        rgba[(size_t)y * W + x] = make_uchar4((uint8_t)rintf(r8), (uint8_t)rintf(g8), (uint8_t)rintf(b8), 255);
        lsum += (unsigned int)y8;
      }
    }
#undef TONED
    if (P.chroma420) {
      const float Cb = fminf(fmaxf(rintf(cb / nb), 0.0f), 255.0f), Cr = fminf(fmaxf(rintf(cr / nb), 0.0f), 255.0f);
      const int cw = (W + 1) / 2, ch = (H + 1) / 2;
      const size_t bi = (size_t)(by / 2) * cw + (bx / 2);
      for (int q = 0; q < 4; q++) {
        const int x = bx + (q & 1), y = by + (q >> 1);
        if (x >= W || y >= H) continue;
        const float y8 = fminf(fmaxf(rintf(Ys[q]), 0.0f), 255.0f);
        const float rr = y8 + 1.402f * (Cr - 128.0f);
        const float gg = (y8 - 0.344136f * (Cb - 128.0f)) - 0.714136f * (Cr - 128.0f);
        const float bb = y8 + 1.772f * (Cb - 128.0f);
        rgba[(size_t)y * W + x] = make_uchar4((uint8_t)fminf(fmaxf(rintf(rr), 0.0f), 255.0f),
                                              (uint8_t)fminf(fmaxf(rintf(gg), 0.0f), 255.0f),
                                              (uint8_t)fminf(fmaxf(rintf(bb), 0.0f), 255.0f), 255);
        yuv[(size_t)y * W + x] = (uint8_t)y8;
        lsum += (unsigned int)y8;
      }
      yuv[(size_t)W * H + bi] = (uint8_t)Cb;
      yuv[(size_t)W * H + (size_t)cw * ch + bi] = (uint8_t)Cr;
    }
  }
  lsum = wave_sum(lsum);
  if ((tid & 63) == 0) sred[tid >> 6] = lsum;
  __syncthreads();
  if (tid == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = make_uint4((sred[0] + sred[1]) + (sred[2] + sred[3]), 0u, 0u, 0u);
}

// the frame's statistics from the blocks' partial sums (integer, so any order gives the same words)
__global__ __launch_bounds__(256) void k_isp_stats(const uint4 *pa, uint32_t na, const uint4 *pb, uint32_t nb,
                                                   unsigned int *stats) {
  __shared__ unsigned int sred[4][4];
  unsigned int a = 0u, r = 0u, g = 0u, b = 0u;
  for (uint32_t i = threadIdx.x; i < na; i += 256) {
    const uint4 v = pa[i];
    r += v.y;
    g += v.z;
    b += v.w;
  }
  for (uint32_t i = threadIdx.x; i < nb; i += 256) a += pb[i].x;
  a = wave_sum(a);
  r = wave_sum(r);
  g = wave_sum(g);
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0) {
    sred[0][threadIdx.x >> 6] = a;
    sred[1][threadIdx.x >> 6] = r;
    sred[2][threadIdx.x >> 6] = g;
    sred[3][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x < 4)
    stats[threadIdx.x] = (sred[threadIdx.x][0] + sred[threadIdx.x][1]) + (sred[threadIdx.x][2] + sred[threadIdx.x][3]);
}

__global__ void k_finalize(const float4 *acc, float scale, uint64_t n, float4 *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 a = acc[i];
  const float s = a.w > 0.0f ? scale / a.w : 0.0f;
  out[i] = make_float4(a.x * s, a.y * s, a.z * s, 1.0f);
}

}  // namespace

hipError_t launch_sensor(const SensorDev &P, const float4 *acc, const float *ri, float4 *glare_a, float4 *glare_b,
                         uint16_t *dn, hipStream_t s) {
  if (P.glare_k > 0.0f) {
    const uint32_t F = P.glare_f, gw = (P.w + F - 1) / F, gh = (P.h + F - 1) / F;
    const dim3 g((gw + 7) / 8, (gh + 7) / 8), b(8, 8);
    hipLaunchKernelGGL(k_glare_down, g, b, 0, s, P, acc, ri, glare_a);
    hipLaunchKernelGGL(k_glare_blur<0>, g, b, 0, s, P, (const float4 *)glare_a, glare_b);
    hipLaunchKernelGGL(k_glare_blur<1>, g, b, 0, s, P, (const float4 *)glare_b, glare_a);
  }
  dim3 grid((P.w + TW - 1) / TW, (P.h + TH - 1) / TH);
  hipLaunchKernelGGL(k_sensor, grid, dim3(TW, TH), 0, s, P, acc, ri, (const float4 *)glare_a, dn);
  return hipGetLastError();
}
__global__ void k_pack_rgb8(const uchar4 *rgba, uint64_t n, uint8_t *rgb) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uchar4 c = rgba[i];
  rgb[3 * i] = c.x;
  rgb[3 * i + 1] = c.y;
  rgb[3 * i + 2] = c.z;
}
hipError_t launch_pack_rgb8(const uchar4 *rgba, uint64_t n, uint8_t *rgb, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_rgb8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rgba, n, rgb);
  return hipGetLastError();
}
uint32_t isp_partials(uint32_t w, uint32_t h) {
  return ((w + DT - 1) / DT) * ((h + DT - 1) / DT) + ((w + FT - 1) / FT) * ((h + FT - 1) / FT);
}
hipError_t launch_isp(const IspDev &P, const uint16_t *dn, float4 *rgb, uint4 *part, uchar4 *rgba, uint8_t *yuv,
                      unsigned int *stats, hipStream_t s) {
  if (P.w == 0 || P.h == 0) return hipSuccess;
  if (P.sharpen_rad < 1 || P.sharpen_rad > FSH) return hipErrorInvalidValue;
  const dim3 blk(16, 16), gd((P.w + DT - 1) / DT, (P.h + DT - 1) / DT), gf((P.w + FT - 1) / FT, (P.h + FT - 1) / FT);
  hipLaunchKernelGGL(k_isp_demosaic, gd, blk, 0, s, P, dn, rgb, part);
  uint4 *pf = part + gd.x * gd.y;
  hipLaunchKernelGGL(k_isp_finish, gf, blk, 0, s, P, (const float4 *)rgb, rgba, yuv, pf);
  hipLaunchKernelGGL(k_isp_stats, dim3(1), dim3(256), 0, s, (const uint4 *)part, gd.x * gd.y, (const uint4 *)pf,
                     gf.x * gf.y, stats);
  return hipGetLastError();
}

hipError_t launch_finalize_linear(const float4 *acc, float scale, uint64_t n, float4 *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, acc, scale, n, out);
  return hipGetLastError();
}

}  // namespace vt

// ------------------------------------------------------------------ the raw stage without allocation
// The two glare levels of vt_sensor_raw_device live in this file-scope device array, GLARE_LEVEL_MAX float4 each
// (8 MiB; every GPU the library runs on holds its own copy until the process ends). Host code never names the array
// in a launch or copy: it uses the address hipGetSymbolAddress returned for the device, kept per device ordinal.
// The glare kernels write every element of a level before reading it, so the array needs no clearing between calls;
// calls that overlap on one GPU share it and race.
static __device__ __align__(256) float4 scratch[2 * vt::GLARE_LEVEL_MAX];

namespace {
std::mutex g_scratch_mu;
float4 *g_scratch[vt::MAX_DEVICES] = {};
}  // namespace

// with `device` current (vt_create): the array's address on that device, once
hipError_t vt::sensor_scratch_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_scratch[device] = static_cast<float4 *>(p);
  return hipSuccess;
}

hipError_t vt::launch_sensor_noalloc(int device, const SensorDev &P, const float4 *acc, const float *ri, uint16_t *dn,
                                     hipStream_t s) {
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  float4 *base;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch[device];
  }
  if (!base) return hipErrorNotInitialized;
  if (P.glare_k > 0.0f) {
    const uint64_t F = P.glare_f, level = ((P.w + F - 1) / F) * ((P.h + F - 1) / F);
    if (F < 1 || level > (uint64_t)GLARE_LEVEL_MAX) return hipErrorInvalidValue;
  }
  return launch_sensor(P, acc, ri, base, base + GLARE_LEVEL_MAX, dn, s);
}
