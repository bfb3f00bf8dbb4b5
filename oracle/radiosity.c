/* Oracle (test infrastructure only): node radiosity.
 *
 * or_jacobi_dense: the dense Jacobi iteration x <- e + F x of the precision configuration, FP64 arithmetic
 * on FP32 F, stopped when ||e + F x - x||_F <= tol ||e||_F. Pinned by a direct FP64 LAPACK solve of
 * (I - F) x = e (numpy) and the uniform closed box x = e / (1 - rho).
 * or_jacobi_gather / or_jacobi_update: the reference's sparse gather step, gpu/wgsl/kernels.wgsl.js:416-461
 * (gatherDir :271-278), FP64 arithmetic on the FP32 tables. Pinned by the uniform-field identity stated at
 * kernels.wgsl.js:414-415 (aux = pi L_in for a uniform incident field with Lambertian tables) and by the
 * fixed point of a two-node exchange solved in closed form.
 */
#include "vt_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

int or_jacobi_dense(const float *F, const float *e, uint32_t n, uint32_t r, double tol, int max_iter,
                    double *x, double *res_out) {
  double *y = malloc(sizeof(double) * (size_t)n * r);
  double en = 0;
  for (size_t i = 0; i < (size_t)n * r; i++) { x[i] = e[i]; en += (double)e[i] * e[i]; }
  en = sqrt(en);
  int it = 0;
  double res = INFINITY;
  for (;;) {
    /* y = e + F x ; residual = ||y - x|| */
    double rr = 0;
#pragma omp parallel for reduction(+ : rr) schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
      const float *row = F + (size_t)i * n;
      for (uint32_t c = 0; c < r; c++) {
        double s = e[(size_t)i * r + c];
        for (uint32_t j = 0; j < n; j++) s += (double)row[j] * x[(size_t)j * r + c];
        y[(size_t)i * r + c] = s;
        double d = s - x[(size_t)i * r + c];
        rr += d * d;
      }
    }
    res = sqrt(rr) / (en > 0 ? en : 1);
    if (res <= tol || it >= max_iter) break;
    memcpy(x, y, sizeof(double) * (size_t)n * r);
    it++;
  }
  if (res_out) *res_out = res;
  free(y);
  return it;
}

static const float GOLD = 0.6180339887498949f;

void or_gather_dir(uint32_t i, uint32_t j, uint32_t K, const float nf[3], float out[3]) {
  double n[3] = {nf[0], nf[1], nf[2]};
  double ln = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int k = 0; k < 3; k++) n[k] /= ln;
  double t[3], b[3];
  or_basis(n, t, b);
  float rot = or_u01(or_pcg(i * 2654435761u + 17u));
  double rr = sqrt(((double)j + 0.5) / K);
  float fr = or_kr(j, GOLD, rot);
  double ph = 2 * M_PI * fr;
  double d[3];
  double z = sqrt(fmax(0, 1 - rr * rr));
  for (int k = 0; k < 3; k++) d[k] = t[k] * (rr * cos(ph)) + b[k] * (rr * sin(ph)) + n[k] * z;
  double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int k = 0; k < 3; k++) out[k] = (float)(d[k] / l);
}

static void alb(const float *spec, const float *diff, int bins, uint32_t m, double mu, double out[3]) {
  double x = mu * bins - 0.5;
  if (x < 0) x = 0;
  if (x > bins - 1) x = bins - 1;
  int i = (int)floor(x);
  int j = i + 1 < bins - 1 ? i + 1 : bins - 1;
  double f = x - i;
  const float *s = spec + (size_t)m * bins * 3, *d = diff + (size_t)m * bins * 3;
  for (int c = 0; c < 3; c++)
    out[c] = ((double)s[3 * i + c] * (1 - f) + (double)s[3 * j + c] * f) + ((double)d[3 * i + c] * (1 - f) + (double)d[3 * j + c] * f);
}

void or_jacobi_gather(uint32_t n_active, const uint32_t *active, const uint32_t *gtable, uint32_t K,
                      const float *verts, const uint32_t *node_mat, const float *alb_spec, const float *alb_diff,
                      int bins, const float *hh, const float *node_state, int slot_in, float *aux) {
  for (uint32_t k = 0; k < n_active; k++) {
    uint32_t i = active[k];
    const float *nf = verts + 8 * (size_t)i + 4;
    uint32_t mi = node_mat[i];
    double acc[3] = {0, 0, 0};
    for (uint32_t j = 0; j < K; j++) {
      uint32_t h = gtable[(size_t)i * K + j];
      if (h == 0xffffffffu) continue;
      float dir[3];
      or_gather_dir(i, j, K, nf, dir);
      double muIn = sqrt(fmax(0, 1 - ((double)j + 0.5) / K));
      const float *nh = verts + 8 * (size_t)h + 4;
      double lh = sqrt((double)nh[0] * nh[0] + (double)nh[1] * nh[1] + (double)nh[2] * nh[2]);
      double muOut = -((double)dir[0] * nh[0] + (double)dir[1] * nh[1] + (double)dir[2] * nh[2]) / lh;
      muOut = muOut < 0 ? 0 : muOut > 1 ? 1 : muOut;
      double ai[3], ah[3];
      alb(alb_spec, alb_diff, bins, mi, muIn, ai);
      alb(alb_spec, alb_diff, bins, node_mat[h], muOut, ah);
      const float *Ls = node_state + 16 * (size_t)h + 4 * slot_in;
      for (int c = 0; c < 3; c++) acc[c] += ai[c] * ah[c] * Ls[c];
    }
    for (int c = 0; c < 3; c++) {
      double h3 = fmax(hh[3 * mi + c], 1e-4);
      aux[4 * (size_t)k + c] = (float)(acc[c] * (M_PI / K) / h3);
    }
    aux[4 * (size_t)k + 3] = 0;
  }
}

void or_jacobi_update(uint32_t n_active, const uint32_t *active, const uint32_t *nbr, const float *nbr_w,
                      uint32_t nbr_k, const float *aux, const uint32_t *node_mat, const float *hh,
                      float *node_state, int slot_out) {
  for (uint32_t k = 0; k < n_active; k++) {
    uint32_t i = active[k];
    double Ef[3] = {0, 0, 0}, ws = 0;
    for (uint32_t j = 0; j < nbr_k; j++) {
      double w = nbr_w[(size_t)k * nbr_k + j];
      if (w <= 0) break;
      uint32_t a = nbr[(size_t)k * nbr_k + j];
      for (int c = 0; c < 3; c++) Ef[c] += (double)aux[4 * (size_t)a + c] * w;
      ws += w;
    }
    if (ws > 0) for (int c = 0; c < 3; c++) Ef[c] /= ws;
    else for (int c = 0; c < 3; c++) Ef[c] = aux[4 * (size_t)k + c];
    float *st = node_state + 16 * (size_t)i;
    uint32_t m = node_mat[i];
    double V = st[3];
    for (int c = 0; c < 3; c++) {
      double h3 = fmax(hh[3 * m + c], 1e-4);
      double Lc = ((double)st[12 + c] + Ef[c] * V * h3) * (1 / M_PI) / h3;
      st[4 * slot_out + c] = (float)Lc;
    }
    st[4 * slot_out + 3] = 1;
    for (int c = 0; c < 3; c++) st[c] = (float)Ef[c];
  }
}
