/* Oracle (test infrastructure only): node radiosity.
 *
 * or_jacobi_dense: the dense Jacobi iteration x <- e + F x of the precision configuration, FP64 arithmetic
 * on FP32 F, stopped when ||e + F x - x||_F <= tol ||e||_F. Pinned by a direct FP64 LAPACK solve of
 * (I - F) x = e (numpy) and the uniform closed box x = e / (1 - rho).
 * or_jacobi_gather / or_jacobi_update: the reference's sparse gather step, gpu/wgsl/kernels.wgsl.js:416-461
 * (gatherDir :271-278), FP64 arithmetic on the FP32 tables. Pinned by the uniform-field identity stated at
 * kernels.wgsl.js:414-415 (aux = pi L_in for a uniform incident field with Lambertian tables) and by the
 * fixed point of a two-node exchange solved in closed form.
 * or_baked_occ / or_baked_jacobi: the lite renderer's gatherOcc and baked Jacobi (lite/shaders.js:234-283).
 * Pinned by a bounded numerical minimiser of the segment distance, capsules placed by hand on two facing planes
 * of nodes, the first step E_dir / pi, the closed box's fixed point e / (pi (1 - rho)) and a GMRES solve of the
 * linear system on the twin's real baked table.
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
  /* nodes are independent (each writes its own aux entry): spread over the OpenMP default team */
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t kk = 0; kk < (int64_t)n_active; kk++) {
    const uint32_t k = (uint32_t)kk;
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

/* ---------------------------------------------------------------- the lite renderer's baked node table */
/* closest distance of the segments p0-p1 and q0-q1 (segSeg, lite/shaders.js:86-104), FP64 */
static double seg_seg(const double p0[3], const double p1[3], const double q0[3], const double q1[3]) {
  double d1[3], d2[3], r[3];
  for (int k = 0; k < 3; k++) { d1[k] = p1[k] - p0[k]; d2[k] = q1[k] - q0[k]; r[k] = p0[k] - q0[k]; }
  const double a = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2], e = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
  const double f = d2[0] * r[0] + d2[1] * r[1] + d2[2] * r[2];
  double s = 0, t = 0;
  if (a <= 1e-14) {
    t = fmin(fmax(f / fmax(e, 1e-14), 0), 1);
  } else {
    const double c = d1[0] * r[0] + d1[1] * r[1] + d1[2] * r[2];
    if (e <= 1e-14) {
      s = fmin(fmax(-c / a, 0), 1);
    } else {
      const double b = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
      const double den = a * e - b * b;
      if (den > 1e-14) s = fmin(fmax((b * f - c * e) / den, 0), 1);
      t = (b * s + f) / e;
      if (t < 0) { t = 0; s = fmin(fmax(-c / a, 0), 1); }
      else if (t > 1) { t = 1; s = fmin(fmax((b - c) / a, 0), 1); }
    }
  }
  double w2 = 0;
  for (int k = 0; k < 3; k++) {
    const double w = (p0[k] + d1[k] * s) - (q0[k] + d2[k] * t);
    w2 += w * w;
  }
  return sqrt(w2);
}

/* gatherOcc (lite/shaders.js:234-258): per gather ray of node i (from p_i + 1 mm n_i to the hit node) the
 * transmission byte, 255 or the smallest round(255 rho) of the capsules [a.xyz, r, b.xyz, rho] within r of it;
 * dist_out (optional): per ray the smallest capsule distance minus that capsule's radius (for near-threshold
 * bookkeeping; +inf for a miss or no capsule) */
void or_baked_occ(uint32_t nodes, uint32_t K, const uint16_t *gather, const float *node_a, const float *node_n,
                  const float *caps, uint32_t n_caps, uint8_t *occ, double *dist_out) {
  for (size_t i = 0; i < nodes; i++) {
    const double p[3] = {(double)node_a[4 * i] + 1e-3 * node_n[4 * i], (double)node_a[4 * i + 1] + 1e-3 * node_n[4 * i + 1],
                         (double)node_a[4 * i + 2] + 1e-3 * node_n[4 * i + 2]};
    for (size_t k = 0; k < K; k++) {
      const uint32_t hit = gather[i * K + k];
      int t = 255;
      double dm = INFINITY;
      if (hit != 0xffffu) {
        const double e[3] = {node_a[4 * (size_t)hit], node_a[4 * (size_t)hit + 1], node_a[4 * (size_t)hit + 2]};
        for (uint32_t c = 0; c < n_caps; c++) {
          const float *C = caps + 8 * (size_t)c;
          const double qa[3] = {C[0], C[1], C[2]}, qb[3] = {C[4], C[5], C[6]};
          const double d = seg_seg(p, e, qa, qb) - C[3];
          if (d < dm) dm = d;
          if (d < 0) {
            /* the byte as the reference quantises it: a binary32 product, rounded half to even */
            const float rho = fminf(fmaxf(C[7], 0.0f), 1.0f);
            const float prod = 255.0f * rho;
            const int v = (int)nearbyintf(prod);
            if (v < t) t = v;
          }
        }
      }
      occ[i * K + k] = (uint8_t)t;
      if (dist_out) dist_out[i * K + k] = dm;
    }
  }
}

/* one step of the baked Jacobi (lite/shaders.js:260-283): E_ind = pi/K sum_k (byte_k / 255) B_in(hit_k),
 * B_out = (E_dir + hh(mesh) E_ind) / pi; B, E_dir, E_ind 4 floats per node, hh 4 floats per mesh */
void or_baked_jacobi(uint32_t nodes, uint32_t K, const uint16_t *gather, const uint8_t *occ, const float *node_n,
                     const float *hh, const float *e_dir, const float *b_in, float *b_out, float *e_ind) {
  for (size_t i = 0; i < nodes; i++) {
    double acc[3] = {0, 0, 0};
    for (size_t k = 0; k < K; k++) {
      const uint32_t hit = gather[i * K + k];
      if (hit == 0xffffu) continue;
      const double t = occ[i * K + k] / 255.0;
      for (int c = 0; c < 3; c++) acc[c] += t * b_in[4 * (size_t)hit + c];
    }
    uint32_t m;
    memcpy(&m, node_n + 4 * i + 3, 4);
    for (int c = 0; c < 3; c++) {
      const double E = acc[c] * (M_PI / K);
      e_ind[4 * i + c] = (float)E;
      b_out[4 * i + c] = (float)(((double)e_dir[4 * i + c] + (double)hh[4 * (size_t)m + c] * E) / M_PI);
    }
    e_ind[4 * i + 3] = 0;
    b_out[4 * i + 3] = 0;
  }
}
