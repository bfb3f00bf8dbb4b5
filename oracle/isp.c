/* Oracle (test infrastructure only): the ISP after the raw DN, gpu/wgsl/sensor.wgsl.js:210-427 (demosaicG,
 * demosaicRB, sensorNR, toneC, sensorOutput) with its parameters from gpu/sensor.js:169-176.
 *
 * Evaluated in binary32 with the WGSL's operation order (dot products left to right, mix = a (1 - t) + b t) and no
 * contraction, so the 8-bit outputs can be compared with the GPU's bit for bit; pow(x, y) = exp(y log x) with the
 * library's own or_expf / or_logf (sensor.c). Pins (tests/test_isp.py): a flat raw field gives the flat value in
 * every demosaic mode and at every border; a linear ramp is reproduced exactly away from the borders; bilinear mode
 * equals the textbook Bayer convolution kernels; grey gives Cb = Cr = 128; the BT.601 primaries give the published
 * YCbCr values; tone curves against their closed forms in FP64; the bilateral filter against a plain Gaussian when
 * the range term vanishes, and identity on flat input; the unsharp mask leaves a ramp alone and overshoots a step
 * by the closed form; the AE / AWB sums against a direct count.
 */
#include "vt_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
static inline float clampf(float v, float lo, float hi) { return v < lo ? lo : v > hi ? hi : v; }
static inline int bayer(uint32_t x, uint32_t y, uint32_t code) {
  static const int pat[4][4] = {{0, 1, 1, 2}, {1, 0, 2, 1}, {1, 2, 0, 1}, {2, 1, 1, 0}};
  return pat[code & 3][(y & 1u) * 2u + (x & 1u)];
}

/* raw sample, mirrored at the borders (reflect by 2 keeps the CFA phase), rawAt :216-223 */
static float raw_px(const or_isp_p *P, const uint16_t *dn, int x, int y) {
  const int W = (int)P->w, H = (int)P->h;
  if (x < 0) x = -x;
  if (x >= W) x = 2 * (W - 1) - x;
  if (y < 0) y = -y;
  if (y >= H) y = 2 * (H - 1) - y;
  return (float)dn[(size_t)clampi(y, 0, H - 1) * W + clampi(x, 0, W - 1)];
}
static float g_px(const or_isp_p *P, const float *g, int x, int y) {
  return g[(size_t)clampi(y, 0, (int)P->h - 1) * P->w + clampi(x, 0, (int)P->w - 1)];
}

/* demosaicG :229-256 */
void or_isp_green(const or_isp_p *P, const uint16_t *dn, float *g) {
  for (int y = 0; y < (int)P->h; y++)
    for (int x = 0; x < (int)P->w; x++) {
      float v = raw_px(P, dn, x, y);
      if (bayer((uint32_t)x, (uint32_t)y, P->cfa) != 1) {
        const float l = raw_px(P, dn, x - 1, y), r = raw_px(P, dn, x + 1, y);
        const float u = raw_px(P, dn, x, y - 1), d = raw_px(P, dn, x, y + 1);
        const float mean4 = 0.25f * (((l + r) + u) + d);
        if (P->demosaic == 0) {
          v = mean4;
        } else if (P->demosaic == 1) {
          const float dh = fabsf(l - r), dv = fabsf(u - d);
          v = dh < dv ? 0.5f * (l + r) : dv < dh ? 0.5f * (u + d) : mean4;
        } else {
          const float lh = (2.0f * v - raw_px(P, dn, x - 2, y)) - raw_px(P, dn, x + 2, y);
          const float lv = (2.0f * v - raw_px(P, dn, x, y - 2)) - raw_px(P, dn, x, y + 2);
          const float dh = fabsf(l - r) + fabsf(lh), dv = fabsf(u - d) + fabsf(lv);
          const float gh = 0.5f * (l + r) + 0.25f * lh, gv = 0.5f * (u + d) + 0.25f * lv;
          v = dh < dv ? gh : dv < dh ? gv : 0.5f * (gh + gv);
          v = clampf(v, 0.0f, P->max_dn);
        }
      }
      g[(size_t)y * P->w + x] = v;
    }
}

/* demosaicRB :263-319; rgb: w*h*4 (linear camera RGB, a = 1); stats[1..3] += grey-world sums */
void or_isp_rb(const or_isp_p *P, const uint16_t *dn, const float *g, float *rgb, uint32_t *stats) {
  const int W = (int)P->w, H = (int)P->h;
  const float full = P->max_dn - P->black > 1.0f ? P->max_dn - P->black : 1.0f;
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const int own = bayer((uint32_t)x, (uint32_t)y, P->cfa);
      const float gc = g_px(P, g, x, y);
      float c[3];
      if (P->demosaic == 2) {
        /* colour differences (c - G) over the nearest R and B sites */
        float sd[2] = {0, 0}, nd[2] = {0, 0};
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            const int xx = clampi(x + dx, 0, W - 1), yy = clampi(y + dy, 0, H - 1);
            const int ch = bayer((uint32_t)xx, (uint32_t)yy, P->cfa);
            if (ch == 1 || (ch == own && (dx || dy))) continue;
            const int k = ch == 0 ? 0 : 1;
            sd[k] = sd[k] + (raw_px(P, dn, xx, yy) - g_px(P, g, xx, yy));
            nd[k] = nd[k] + 1.0f;
          }
        c[0] = gc + sd[0] / (nd[0] > 1.0f ? nd[0] : 1.0f);
        c[1] = gc;
        c[2] = gc + sd[1] / (nd[1] > 1.0f ? nd[1] : 1.0f);
      } else {
        /* plain means of each channel's sites in the 3x3 window (own channel: the centre only) */
        float s[3] = {0, 0, 0}, n[3] = {0, 0, 0};
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            const int xx = clampi(x + dx, 0, W - 1), yy = clampi(y + dy, 0, H - 1);
            const int ch = bayer((uint32_t)xx, (uint32_t)yy, P->cfa);
            if (ch == own && (dx || dy)) continue;
            s[ch] = s[ch] + raw_px(P, dn, xx, yy);
            n[ch] = n[ch] + 1.0f;
          }
        for (int k = 0; k < 3; k++) c[k] = s[k] / (n[k] > 1.0f ? n[k] : 1.0f);
        if (P->demosaic == 1) c[1] = gc;
      }
      for (int k = 0; k < 3; k++) {
        c[k] = (c[k] - P->black) / full;
        stats[1 + k] += (uint32_t)rintf(clampf(c[k], 0.0f, 1.0f) * 1023.0f);
        c[k] = c[k] * P->wb[k];
      }
      float *o = rgb + 4 * ((size_t)y * W + x);
      for (int r = 0; r < 3; r++) o[r] = (P->ccm[3 * r] * c[0] + P->ccm[3 * r + 1] * c[1]) + P->ccm[3 * r + 2] * c[2];
      o[3] = 1.0f;
    }
}

/* sensorNR :321-348 (the caller skips it when off) */
void or_isp_nr(const or_isp_p *P, const float *src, float *dst) {
  const int W = (int)P->w, H = (int)P->h;
  const float fullE = (P->max_dn - P->black > 1.0f ? P->max_dn - P->black : 1.0f) /
                      (P->conv_gain * P->gain > 1e-9f ? P->conv_gain * P->gain : 1e-9f);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const float *c0 = src + 4 * ((size_t)y * W + x);
      float Y0 = (c0[0] * 0.2126f + c0[1] * 0.7152f) + c0[2] * 0.0722f;
      Y0 = Y0 > 0.0f ? Y0 : 0.0f;
      const float sn = sqrtf(Y0 * fullE + P->read_noise * P->read_noise) / fullE;
      float rs = P->nr_scale * sn;
      rs = rs > 1e-6f ? rs : 1e-6f;
      const float ss = P->nr_sigma;
      float acc[3] = {0, 0, 0}, ws = 0.0f;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const float *c = src + 4 * ((size_t)clampi(y + dy, 0, H - 1) * W + clampi(x + dx, 0, W - 1));
          const float dY = ((c[0] - c0[0]) * 0.2126f + (c[1] - c0[1]) * 0.7152f) + (c[2] - c0[2]) * 0.0722f;
          const float w = or_expf(-0.5f * (float)(dx * dx + dy * dy) / (ss * ss) - 0.5f * dY * dY / (rs * rs));
          for (int k = 0; k < 3; k++) acc[k] = acc[k] + w * c[k];
          ws = ws + w;
        }
      float *o = dst + 4 * ((size_t)y * W + x);
      for (int k = 0; k < 3; k++) o[k] = acc[k] / ws;
      o[3] = 1.0f;
    }
}

static float powr(float x, float y) { return x > 0.0f ? or_expf(y * or_logf(x)) : 0.0f; }
static float srgb_oetf(float v) { return v <= 0.0031308f ? 12.92f * v : 1.055f * powr(v, 1.0f / 2.4f) - 0.055f; }

/* toneC :350-366 */
float or_isp_tone(const or_isp_p *P, float x) {
  if (P->tone == 3) {
    const float u = srgb_oetf(clampf((x - P->black_off) / (1.0f - P->black_off), 0.0f, 1.0f));
    const float a = powr(u, P->s_c);
    const float den = a + powr(1.0f - u, P->s_c);
    return a / (den > 1e-12f ? den : 1e-12f);
  }
  const float c = clampf(x, 0.0f, 1.0f);
  if (P->tone == 0) return srgb_oetf(c);
  if (P->tone == 1) return powr(c, 1.0f / (P->gamma > 1e-3f ? P->gamma : 1e-3f));
  return c;
}
static void toned_px(const or_isp_p *P, const float *src, int x, int y, float t[3]) {
  const float *c = src + 4 * ((size_t)clampi(y, 0, (int)P->h - 1) * P->w + clampi(x, 0, (int)P->w - 1));
  for (int k = 0; k < 3; k++) t[k] = or_isp_tone(P, c[k]);
}

/* sensorOutput :374-427, one 2x2 block at a time: rgba w*h*4 bytes; yuv (4:2:0 only) Y w*h, Cb, Cr
 * ceil(w/2) x ceil(h/2); stats[0] += the luma sum */
void or_isp_out(const or_isp_p *P, const float *src, uint8_t *rgba, uint8_t *yuv, uint32_t *stats) {
  const int W = (int)P->w, H = (int)P->h, cw = (W + 1) / 2, chh = (H + 1) / 2;
  const int rad = P->sharpen_rad;
  for (int by = 0; by < H; by += 2)
    for (int bx = 0; bx < W; bx += 2) {
      float Ys[4] = {0, 0, 0, 0}, cb = 0.0f, cr = 0.0f, nb = 0.0f;
      for (int q = 0; q < 4; q++) {
        const int x = bx + (q & 1), y = by + (q >> 1);
        if (x >= W || y >= H) continue;
        float c[3];
        toned_px(P, src, x, y, c);
        if (P->sharpen_amount > 0.0f) {
          float m = 0.0f, n = 0.0f;
          for (int dy = -rad; dy <= rad; dy++)
            for (int dx = -rad; dx <= rad; dx++) {
              float t[3];
              toned_px(P, src, x + dx, y + dy, t);
              m += (t[0] * 0.299f + t[1] * 0.587f) + t[2] * 0.114f;
              n += 1.0f;
            }
          const float Y0 = (c[0] * 0.299f + c[1] * 0.587f) + c[2] * 0.114f;
          const float d = P->sharpen_amount * (Y0 - m / n);
          for (int k = 0; k < 3; k++) c[k] += d;
        }
        float c8[3];
        for (int k = 0; k < 3; k++) c8[k] = clampf(c[k], 0.0f, 1.0f) * 255.0f;
        Ys[q] = (0.299f * c8[0] + 0.587f * c8[1]) + 0.114f * c8[2];
        cb += ((128.0f - 0.168736f * c8[0]) - 0.331264f * c8[1]) + 0.5f * c8[2];
        cr += ((128.0f + 0.5f * c8[0]) - 0.418688f * c8[1]) - 0.081312f * c8[2];
        nb += 1.0f;
        if (!P->chroma420) {
          uint8_t *o = rgba + 4 * ((size_t)y * W + x);
          for (int k = 0; k < 3; k++) o[k] = (uint8_t)rintf(c8[k]);
          o[3] = 255;
          stats[0] += (uint32_t)clampf(rintf(Ys[q]), 0.0f, 255.0f);
        }
      }
      if (!P->chroma420) continue;
      const float Cb = clampf(rintf(cb / nb), 0.0f, 255.0f), Cr = clampf(rintf(cr / nb), 0.0f, 255.0f);
      const size_t bi = (size_t)(by / 2) * cw + (size_t)(bx / 2);
      for (int q = 0; q < 4; q++) {
        const int x = bx + (q & 1), y = by + (q >> 1);
        if (x >= W || y >= H) continue;
        const float y8 = clampf(rintf(Ys[q]), 0.0f, 255.0f);
        const float rgb[3] = {y8 + 1.402f * (Cr - 128.0f), (y8 - 0.344136f * (Cb - 128.0f)) - 0.714136f * (Cr - 128.0f),
                              y8 + 1.772f * (Cb - 128.0f)};
        uint8_t *o = rgba + 4 * ((size_t)y * W + x);
        for (int k = 0; k < 3; k++) o[k] = (uint8_t)clampf(rintf(rgb[k]), 0.0f, 255.0f);
        o[3] = 255;
        yuv[(size_t)y * W + x] = (uint8_t)y8;
        stats[0] += (uint32_t)y8;
      }
      yuv[(size_t)W * H + bi] = (uint8_t)Cb;
      yuv[(size_t)W * H + (size_t)cw * chh + bi] = (uint8_t)Cr;
    }
}

/* the four passes in order (the sensor pass list, gpu/renderer.js:1034-1037); lin (w*h*4, optional) receives the
 * linear RGB that the output pass reads */
void or_isp(const or_isp_p *P, const uint16_t *dn, uint8_t *rgba, uint8_t *yuv, uint32_t *stats, float *lin) {
  const size_t px = (size_t)P->w * P->h;
  float *g = malloc(sizeof(float) * px), *a = malloc(sizeof(float) * 4 * px), *b = malloc(sizeof(float) * 4 * px);
  memset(stats, 0, 4 * sizeof(uint32_t));
  or_isp_green(P, dn, g);
  or_isp_rb(P, dn, g, a, stats);
  const float *img = a;
  if (P->nr) {
    or_isp_nr(P, a, b);
    img = b;
  }
  or_isp_out(P, img, rgba, yuv, stats);
  if (lin) memcpy(lin, img, sizeof(float) * 4 * px);
  free(g);
  free(a);
  free(b);
}
