/* Oracle (test infrastructure only): the twin's BSDF and the shading helpers of gpu/shade_ref.js, FP64.
 *
 * Follows gpu/bsdf.js:25-280 and gpu/shade_ref.js:5-78. OR_LAMBERT (f = rho/pi, cosine sampling) is the
 * textbook Lambertian surface that the synthetic scene configurations name; it is not in the reference.
 * Pins (tests/test_oracle_pins.py): reciprocity f(l, v) = f(v, l); the analytic diffuse albedo of
 * gpu/bsdf.js:240-241 against quadrature; the pdf integrating to 1; sample histograms against the pdf;
 * Fbar by Simpson against scipy quad; the disc irradiance against the closed form pi L a^2 / (a^2 + d^2);
 * capsule occlusion against Monte Carlo over the hemisphere.
 */
#include "vt_oracle.h"
#include <math.h>
#include <string.h>

#define INV_PI (1.0 / M_PI)

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void norm3(double *a) {
  double l = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  a[0] /= l; a[1] /= l; a[2] /= l;
}
static inline double clamp01(double x) { return x < 0 ? 0 : x > 1 ? 1 : x; }

void or_mat_unpack(const float r[16], or_mat *m) {
  memset(m, 0, sizeof(*m));
  m->c[0] = r[0]; m->c[1] = r[1]; m->c[2] = r[2]; m->alpha = r[3];
  m->ior = r[4]; m->kind = (int)r[5]; m->fbar = r[6]; m->roughness = r[7];
  m->tau[0] = r[8]; m->tau[1] = r[9]; m->tau[2] = r[10]; m->alpha_c = r[11];
  m->fbar_computed = r[12];
}

double or_fresnel_dielectric(double cosi, double eta) {
  double c = clamp01(cosi);
  double s2 = (1 - c * c) / (eta * eta);
  if (s2 >= 1) return 1;
  double ct = sqrt(1 - s2);
  double rs = (c - eta * ct) / (c + eta * ct), rp = (eta * c - ct) / (eta * c + ct);
  return 0.5 * (rs * rs + rp * rp);
}
double or_fresnel_average(double eta, int n) {
  double s = 0;
  for (int i = 0; i <= n; i++) {
    double mu = (double)i / n;
    int w = (i == 0 || i == n) ? 1 : (i % 2) ? 4 : 2;
    s += w * or_fresnel_dielectric(mu, eta) * mu;
  }
  return 2 * s / (3.0 * n);
}
double or_schlick(double f0, double c) { double m = pow(1 - clamp01(c), 5); return f0 + (1 - f0) * m; }
double or_ggx_d(double nh, double a) {
  if (nh <= 0) return 0;
  double a2 = a * a, d = nh * nh * (a2 - 1) + 1;
  return a2 / (M_PI * d * d);
}
double or_ggx_lambda(double c, double a) {
  double c2 = c * c > 1e-12 ? c * c : 1e-12;
  double t2 = (1 - c2) / c2;
  return 0.5 * (-1 + sqrt(1 + a * a * t2));
}
double or_ggx_g1(double c, double a) { return 1 / (1 + or_ggx_lambda(c, a)); }
double or_ggx_g2(double nl, double nv, double a) { return 1 / (1 + or_ggx_lambda(nl, a) + or_ggx_lambda(nv, a)); }

static void fresnel_rgb(const or_mat *M, double c, double F[3]) {
  if (M->kind == OR_CONDUCTOR) { for (int k = 0; k < 3; k++) F[k] = or_schlick(M->c[k], c); return; }
  double f = or_fresnel_dielectric(c, M->ior);
  F[0] = F[1] = F[2] = f;
}

void or_eval_bsdf(const or_mat *M, const double n[3], const double l[3], const double v[3], double f[3]) {
  double nl = dot3(n, l), nv = dot3(n, v);
  f[0] = f[1] = f[2] = 0;
  if (nl <= 0 || nv <= 0) return;
  if (M->kind == OR_LAMBERT) { for (int k = 0; k < 3; k++) f[k] = M->c[k] * INV_PI; return; }
  double h[3] = {l[0] + v[0], l[1] + v[1], l[2] + v[2]};
  norm3(h);
  double nh = dot3(n, h), vh = dot3(v, h);
  if (vh < 0) vh = 0;
  if (M->kind == OR_COATED) {
    double coat = or_ggx_d(nh, M->alpha_c) * or_ggx_g2(nl, nv, M->alpha_c) * or_fresnel_dielectric(vh, M->ior) / (4 * nl * nv);
    double t = (1 - or_fresnel_dielectric(nl, M->ior)) * (1 - or_fresnel_dielectric(nv, M->ior));
    double metal = or_ggx_d(nh, M->alpha) * or_ggx_g2(nl, nv, M->alpha) / (4 * nl * nv);
    for (int c = 0; c < 3; c++) f[c] = coat + t * pow(M->tau[c], 1 / nl + 1 / nv) * metal * or_schlick(M->c[c], vh);
    return;
  }
  double spec = or_ggx_d(nh, M->alpha) * or_ggx_g2(nl, nv, M->alpha) / (4 * nl * nv);
  double F[3];
  fresnel_rgb(M, vh, F);
  for (int c = 0; c < 3; c++) f[c] = spec * F[c];
  if (M->kind == OR_DIELECTRIC) {
    double kd = INV_PI * (1 - or_fresnel_dielectric(nl, M->ior)) * (1 - or_fresnel_dielectric(nv, M->ior)) / (1 - M->fbar);
    for (int c = 0; c < 3; c++) f[c] += M->c[c] * kd;
  }
}

void or_basis(const double n[3], double t[3], double b[3]) {
  double s = n[2] >= 0 ? 1 : -1;
  double a = -1 / (s + n[2]), bb = n[0] * n[1] * a;
  t[0] = 1 + s * n[0] * n[0] * a; t[1] = s * bb; t[2] = -s * n[0];
  b[0] = bb; b[1] = s + n[1] * n[1] * a; b[2] = -n[1];
}
static void to_world(const double t[3], const double b[3], const double n[3], const double x[3], double out[3]) {
  for (int k = 0; k < 3; k++) out[k] = t[k] * x[0] + b[k] * x[1] + n[k] * x[2];
}

void or_sample_vndf(const double ve[3], double a, double u1, double u2, double hout[3]) {
  double vh[3] = {a * ve[0], a * ve[1], ve[2]};
  norm3(vh);
  double lensq = vh[0] * vh[0] + vh[1] * vh[1];
  double T1[3] = {1, 0, 0};
  if (lensq > 0) { double s = sqrt(lensq); T1[0] = -vh[1] / s; T1[1] = vh[0] / s; T1[2] = 0; }
  double T2[3] = {vh[1] * T1[2] - vh[2] * T1[1], vh[2] * T1[0] - vh[0] * T1[2], vh[0] * T1[1] - vh[1] * T1[0]};
  double r = sqrt(u1), phi = 2 * M_PI * u2;
  double t1 = r * cos(phi), t2 = r * sin(phi);
  double s = 0.5 * (1 + vh[2]);
  t2 = (1 - s) * sqrt(fmax(0, 1 - t1 * t1)) + s * t2;
  double t3 = sqrt(fmax(0, 1 - t1 * t1 - t2 * t2));
  double nh[3];
  for (int k = 0; k < 3; k++) nh[k] = t1 * T1[k] + t2 * T2[k] + t3 * vh[k];
  hout[0] = a * nh[0]; hout[1] = a * nh[1]; hout[2] = fmax(1e-7, nh[2]);
  norm3(hout);
}

double or_spec_prob(const or_mat *M, double nv) {
  if (M->kind == OR_LAMBERT) return 0;
  if (M->kind == OR_COATED) {
    double Fc = or_fresnel_dielectric(nv, M->ior), mf = (M->c[0] + M->c[1] + M->c[2]) / 3;
    return Fc / (Fc + (1 - Fc) * mf);
  }
  if (M->kind == OR_CONDUCTOR) return 1;
  double F = or_fresnel_dielectric(nv, M->ior);
  double d = (M->c[0] * 0.2126 + M->c[1] * 0.7152 + M->c[2] * 0.0722) * (1 - F);
  double s = F > 0.04 ? F : 0.04;
  double p = s / (s + d);
  return p < 0.05 ? 0.05 : p > 0.95 ? 0.95 : p;
}

double or_pdf_bsdf(const or_mat *M, const double n[3], const double v[3], const double l[3]) {
  double nl = dot3(n, l), nv = dot3(n, v);
  if (nl <= 0 || nv <= 0) return 0;
  if (M->kind == OR_LAMBERT) return nl * INV_PI;
  double h[3] = {l[0] + v[0], l[1] + v[1], l[2] + v[2]};
  norm3(h);
  double nh = dot3(n, h), vh = dot3(v, h);
  double ps = or_spec_prob(M, nv);
  if (M->kind == OR_COATED) {
    if (!(vh > 0)) return 0;
    double pc = or_ggx_g1(nv, M->alpha_c) * or_ggx_d(nh, M->alpha_c) / (4 * nv);
    double pm = or_ggx_g1(nv, M->alpha) * or_ggx_d(nh, M->alpha) / (4 * nv);
    return ps * pc + (1 - ps) * pm;
  }
  double pdf_spec = vh > 0 ? or_ggx_g1(nv, M->alpha) * fmax(0, vh) * or_ggx_d(nh, M->alpha) / nv / (4 * vh) : 0;
  return ps * pdf_spec + (1 - ps) * nl * INV_PI;
}

double or_sample_bsdf(const or_mat *M, const double n[3], const double v[3], const double u[3], double l[3]) {
  double nv = dot3(n, v);
  if (nv <= 0) return 0;
  double t[3], b[3];
  or_basis(n, t, b);
  double ps = or_spec_prob(M, nv);
  double ve[3] = {dot3(t, v), dot3(b, v), dot3(n, v)};
  if (M->kind == OR_COATED || (M->kind != OR_LAMBERT && u[0] < ps)) {
    double a = M->kind == OR_COATED ? (u[0] < ps ? M->alpha_c : M->alpha) : M->alpha;
    double hl[3], h[3];
    or_sample_vndf(ve, a, u[1], u[2], hl);
    to_world(t, b, n, hl, h);
    double vh = dot3(v, h);
    for (int k = 0; k < 3; k++) l[k] = 2 * vh * h[k] - v[k];
  } else {
    double r = sqrt(u[1]), phi = 2 * M_PI * u[2];
    double x[3] = {r * cos(phi), r * sin(phi), sqrt(fmax(0, 1 - u[1]))};
    to_world(t, b, n, x, l);
  }
  if (dot3(n, l) <= 0) return 0;
  return or_pdf_bsdf(M, n, v, l);
}

double or_radinv2(uint32_t i) {
  double b = 0, f = 0.5;
  while (i) { if (i & 1) b += f; i >>= 1; f *= 0.5; }
  return b;
}

void or_albedo_table(const or_mat *M, int bins, int samples, double *spec, double *diff, double hh[3], double hh_spec[3]) {
  const double n[3] = {0, 0, 1};
  for (int k = 0; k < bins; k++) {
    double mu = (k + 0.5) / bins, v[3] = {sqrt(1 - mu * mu), 0, mu};
    double acc[3] = {0, 0, 0};
    if (M->kind == OR_LAMBERT) {
      for (int c = 0; c < 3; c++) { spec[3 * k + c] = 0; diff[3 * k + c] = M->c[c]; }
      continue;
    }
    if (M->kind == OR_COATED) {
      for (int i = 0; i < samples; i++) {
        double u[3] = {(i + 0.5) / samples, or_radinv2((uint32_t)i), fmod(or_radinv2((uint32_t)i ^ 0x5555u), 1.0)};
        double l[3];
        double pdf = or_sample_bsdf(M, n, v, u, l);
        if (!(pdf > 0)) continue;
        double f[3];
        or_eval_bsdf(M, n, l, v, f);
        for (int c = 0; c < 3; c++) acc[c] += f[c] * l[2] / pdf;
      }
      for (int c = 0; c < 3; c++) { spec[3 * k + c] = acc[c] / samples; diff[3 * k + c] = 0; }
      continue;
    }
    for (int i = 0; i < samples; i++) {
      double u1 = (i + 0.5) / samples, u2 = or_radinv2((uint32_t)i);
      double h[3];
      or_sample_vndf(v, M->alpha, u1, u2, h);
      double vh = dot3(v, h);
      double l[3] = {2 * vh * h[0] - v[0], 2 * vh * h[1] - v[1], 2 * vh * h[2] - v[2]};
      if (l[2] <= 0) continue;
      double w = or_ggx_g2(l[2], mu, M->alpha) / or_ggx_g1(mu, M->alpha);
      double F[3];
      fresnel_rgb(M, vh, F);
      for (int c = 0; c < 3; c++) acc[c] += F[c] * w;
    }
    for (int c = 0; c < 3; c++) {
      spec[3 * k + c] = acc[c] / samples;
      diff[3 * k + c] = M->kind == OR_CONDUCTOR ? 0
        : M->c[c] * (1 - or_fresnel_dielectric(mu, M->ior)) * (1 - M->fbar_computed) / (1 - M->fbar);
    }
  }
  for (int c = 0; c < 3; c++) { hh[c] = 0; hh_spec[c] = 0; }
  for (int k = 0; k < bins; k++) {
    double mu = (k + 0.5) / bins;
    for (int c = 0; c < 3; c++) {
      hh[c] += 2 * mu * (spec[3 * k + c] + diff[3 * k + c]) / bins;
      hh_spec[c] += 2 * mu * spec[3 * k + c] / bins;
    }
  }
}

void or_lookup_albedo(const double *spec, const double *diff, int bins, double mu, int which, double out[3]) {
  double x = mu * bins - 0.5;
  if (x < 0) x = 0;
  if (x > bins - 1) x = bins - 1;
  int i = (int)floor(x);
  double f = x - i;
  int j = i + 1 < bins - 1 ? i + 1 : bins - 1;
  for (int c = 0; c < 3; c++) {
    double gs = spec[3 * i + c] * (1 - f) + spec[3 * j + c] * f;
    double gd = diff[3 * i + c] * (1 - f) + diff[3 * j + c] * f;
    out[c] = which == 1 ? gs : which == 2 ? gd : gs + gd;
  }
}

void or_layer_normal(const double n[3], const double p[3], const double bd[3], double period, double depth, double out[3]) {
  double s = dot3(p, bd);
  double dh = -(depth / 2) * (2 * M_PI / period) * sin(2 * M_PI * s / period);
  double bn = dot3(bd, n);
  for (int k = 0; k < 3; k++) out[k] = n[k] - dh * (bd[k] - bn * n[k]);
  norm3(out);
}

/* ------------------------------------------------------------------ gpu/shade_ref.js */
double or_profile(double theta, const double *knots, const double *p, int n) {
  if (theta >= knots[n - 1]) return theta == knots[n - 1] ? p[n - 1] : 0;
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) { int m = (lo + hi) >> 1; if (knots[m] <= theta) lo = m; else hi = m; }
  double den = knots[hi] - knots[lo];
  double f = clamp01((theta - knots[lo]) / (den > 1e-12 ? den : 1e-12));
  return p[lo] + (p[hi] - p[lo]) * f;
}

double or_capsule_occ(const double p[3], const double n[3], const double a[3], const double b[3], double r) {
  double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double L = sqrt(dot3(ab, ab));
  if (L < 1e-9 || r <= 0) return 0;
  double e[3] = {ab[0] / L, ab[1] / L, ab[2] / L};
  double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
  double s0 = dot3(ap, e);
  double perp[3] = {ap[0] - s0 * e[0], ap[1] - s0 * e[1], ap[2] - s0 * e[2]};
  double pl = sqrt(dot3(perp, perp));
  double h = sqrt(pl * pl + r * r);
  double ep[3];
  double den = pl > 1e-12 ? pl : 1e-12;
  for (int k = 0; k < 3; k++) ep[k] = -perp[k] / den;
  if (pl < 1e-12) {
    double t[3] = {fabs(e[0]) < 0.9 ? 1 : 0, fabs(e[0]) < 0.9 ? 0 : 1, 0};
    ep[0] = e[1] * t[2] - e[2] * t[1]; ep[1] = e[2] * t[0] - e[0] * t[2]; ep[2] = e[0] * t[1] - e[1] * t[0];
    norm3(ep);
  }
  double np = dot3(n, ep), na = dot3(n, e);
  double f1 = atan((0 - s0) / h), f2 = atan((L - s0) / h);
  double f0 = atan2(na, np);
  if (f0 - M_PI / 2 > f1) f1 = f0 - M_PI / 2;
  if (f0 + M_PI / 2 < f2) f2 = f0 + M_PI / 2;
  if (f2 <= f1) return 0;
  double I = np * (0.5 * (f2 - f1) + 0.25 * (sin(2 * f2) - sin(2 * f1))) + na * 0.5 * (pow(sin(f2), 2) - pow(sin(f1), 2));
  double o = (2 * r / h) * I / M_PI;
  return o < 0 ? 0 : o > 1 ? 1 : o;
}

void or_gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < n; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0;
    for (int it = 0; it < 100; it++) {
      double p1 = 1, p2 = 0;
      for (int j = 1; j <= n; j++) { double p3 = p2; p2 = p1; p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j; }
      pp = n * (z * p1 - p2) / (z * z - 1);
      double z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) < 1e-15) break;
    }
    x[i] = -z; w[i] = 2 / ((1 - z * z) * pp * pp);
  }
}

double or_disc_irradiance(const double x[3], const double n[3], const double c[3], const double m[3],
                          double a, double L, int nr, int nphi) {
  double gx[256], gw[256];
  if (nr > 256) nr = 256;
  or_gauss_legendre(nr, gx, gw);
  double t[3] = {fabs(m[0]) < 0.9 ? 1 : 0, fabs(m[0]) < 0.9 ? 0 : 1, 0};
  double u[3] = {m[1] * t[2] - m[2] * t[1], m[2] * t[0] - m[0] * t[2], m[0] * t[1] - m[1] * t[0]};
  norm3(u);
  double w[3] = {m[1] * u[2] - m[2] * u[1], m[2] * u[0] - m[0] * u[2], m[0] * u[1] - m[1] * u[0]};
  double E = 0;
  for (int i = 0; i < nr; i++) {
    double rho = 0.5 * a * (gx[i] + 1), wr = 0.5 * a * gw[i];
    for (int j = 0; j < nphi; j++) {
      double ph = (2 * M_PI * (j + 0.5)) / nphi;
      double q[3], d[3];
      for (int k = 0; k < 3; k++) { q[k] = c[k] + rho * (cos(ph) * u[k] + sin(ph) * w[k]); d[k] = q[k] - x[k]; }
      double d2 = dot3(d, d), dl = sqrt(d2);
      double cosX = dot3(n, d) / dl, cosQ = -dot3(m, d) / dl;
      if (cosX > 0 && cosQ > 0) E += L * cosX * cosQ / d2 * rho * wr * (2 * M_PI / nphi);
    }
  }
  return E;
}
