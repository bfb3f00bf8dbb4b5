// The sensor epilogue of gpu/wgsl/sensor.wgsl.js:46-208 as ONE kernel: per 32x8 output tile the horizontal
// optics pass (radiance x RI, pixel-integrated Gaussian taps) is computed once into LDS for the rows the
// vertical taps need, then the vertical pass, CFA response, Poisson shot + Gaussian read noise, full-well clip
// and the 12-bit ADC write a uint16 DN. HBM traffic: the radiance sums and RI map in, 2 B/pixel out.
//
// Integer DNs match the host sensor model bit for bit: every binary32 operation is unfused (this file is built
// with -ffp-contract=off), in the WGSL's order, and exp/log/cos are this library's own routines (published
// methods, stated below) instead of device intrinsics whose last-ulp behaviour differs from the host's.
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

__global__ __launch_bounds__(TW *TH) void k_sensor(SensorDev P, const float4 *acc, const float *ri, uint16_t *dn) {
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
        const float l0 = s.x * P.inv_spp, l1 = s.y * P.inv_spp, l2 = s.z * P.inv_spp;
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
  // vertical pass + exposure (sensorExpose, glare off)
  float e0 = 0.0f, e1 = 0.0f, e2 = 0.0f;
  for (int k = -r; k <= r; k++) {
    const int j = (int)threadIdx.y + k + r;
    const float w = sw[k + r];
    e0 = e0 + w * sA[0][j * TW + threadIdx.x];
    e1 = e1 + w * sA[1][j * TW + threadIdx.x];
    e2 = e2 + w * sA[2][j * TW + threadIdx.x];
  }
  e0 = e0 / wsum; e1 = e1 / wsum; e2 = e2 / wsum;
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

__global__ void k_finalize(const float4 *acc, float s, uint64_t n, float4 *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 a = acc[i];
  out[i] = make_float4(a.x * s, a.y * s, a.z * s, 1.0f);
}

}  // namespace

hipError_t launch_sensor(const SensorDev &P, const float4 *acc, const float *ri, uint16_t *dn, hipStream_t s) {
  dim3 grid((P.w + TW - 1) / TW, (P.h + TH - 1) / TH);
  hipLaunchKernelGGL(k_sensor, grid, dim3(TW, TH), 0, s, P, acc, ri, dn);
  return hipGetLastError();
}
hipError_t launch_finalize_linear(const float4 *acc, float scale_over_spp, uint64_t n, float4 *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, acc, scale_over_spp, n, out);
  return hipGetLastError();
}

}  // namespace vt
