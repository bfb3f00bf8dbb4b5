/* Oracle (test infrastructure only): the reference BVH2 and its CPU traversal.
 *
 * Restates gpu/bvh.js:18-279 operation for operation: FP32 storage (Float32Array) with FP64 arithmetic (JS
 * numbers), the same bin index formula, sweep order, tie-breaks, in-place partition, node allocation order
 * (a LIFO stack, children allocated when their parent is popped) and the same traversal order (leaves of a
 * node tested in child order, interior children pushed far-first, V8's stable sort on equal entry distance).
 * Pinned by golden node arrays, orders and traversal results made by running gpu/bvh.js under Node
 * (tests/golden/make_golden.mjs), by brute force (gpu/bvh.js:253-261) and by the tree invariants.
 */
#include "vt_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BIG 1e30

typedef struct { double lo[3], hi[3], clo[3], chi[3]; } bounds_t;
typedef struct {
  const float *tbmin, *tbmax, *cen;
  uint32_t *order;
  int bins, leaf;
  double trav, tric;
  uint32_t *bcount, *rcount;
  double *bmin, *bmax, *rarea;
} builder_t;

/* Math.min / Math.max on non-NaN values: -0 counts as less than +0 */
static inline double dmin(double a, double b) {
  if (a < b) return a;
  if (b < a) return b;
  return signbit(a) ? a : b;
}
static inline double dmax(double a, double b) {
  if (a > b) return a;
  if (b > a) return b;
  return signbit(a) ? b : a;
}

static void range_bounds(const builder_t *B, int64_t s, int64_t e, bounds_t *R) {
  for (int k = 0; k < 3; k++) { R->lo[k] = BIG; R->hi[k] = -BIG; R->clo[k] = BIG; R->chi[k] = -BIG; }
  for (int64_t i = s; i < e; i++) {
    uint32_t t = B->order[i];
    for (int k = 0; k < 3; k++) {
      double a = B->tbmin[3 * t + k], b = B->tbmax[3 * t + k], c = B->cen[3 * t + k];
      if (a < R->lo[k]) R->lo[k] = a;
      if (b > R->hi[k]) R->hi[k] = b;
      if (c < R->clo[k]) R->clo[k] = c;
      if (c > R->chi[k]) R->chi[k] = c;
    }
  }
}
static double area3(const double lo[3], const double hi[3]) {
  double dx = dmax(0, hi[0] - lo[0]), dy = dmax(0, hi[1] - lo[1]), dz = dmax(0, hi[2] - lo[2]);
  return dx * dy + dy * dz + dz * dx;
}
static inline int bin_of(double c, double lo, double sc, int bins) {
  int b = (int)floor((c - lo) * sc);
  if (b >= bins) b = bins - 1;
  return b;
}

/* gpu/bvh.js:66-116 */
static int64_t split(builder_t *B, int64_t s, int64_t e, const bounds_t *R) {
  int64_t cnt = e - s;
  if (cnt <= 1) return -1;
  double best = INFINITY;
  int best_axis = -1, best_bin = -1;
  const int bins = B->bins;
  double pA = area3(R->lo, R->hi);
  for (int ax = 0; ax < 3; ax++) {
    double lo = R->clo[ax], ext = R->chi[ax] - lo;
    if (!(ext > 0)) continue;
    for (int b = 0; b < bins; b++) {
      B->bcount[b] = 0;
      for (int k = 0; k < 3; k++) { B->bmin[3 * b + k] = BIG; B->bmax[3 * b + k] = -BIG; }
    }
    double sc = bins / ext;
    for (int64_t i = s; i < e; i++) {
      uint32_t t = B->order[i];
      int b = bin_of(B->cen[3 * t + ax], lo, sc, bins);
      B->bcount[b]++;
      for (int k = 0; k < 3; k++) {
        if (B->tbmin[3 * t + k] < B->bmin[3 * b + k]) B->bmin[3 * b + k] = B->tbmin[3 * t + k];
        if (B->tbmax[3 * t + k] > B->bmax[3 * b + k]) B->bmax[3 * b + k] = B->tbmax[3 * t + k];
      }
    }
    double lo3[3] = {BIG, BIG, BIG}, hi3[3] = {-BIG, -BIG, -BIG};
    uint32_t c = 0;
    for (int b = bins - 1; b > 0; b--) {
      c += B->bcount[b];
      for (int k = 0; k < 3; k++) { lo3[k] = dmin(lo3[k], B->bmin[3 * b + k]); hi3[k] = dmax(hi3[k], B->bmax[3 * b + k]); }
      B->rcount[b] = c;
      B->rarea[b] = c ? area3(lo3, hi3) : 0;
    }
    for (int k = 0; k < 3; k++) { lo3[k] = BIG; hi3[k] = -BIG; }
    c = 0;
    for (int b = 0; b < bins - 1; b++) {
      c += B->bcount[b];
      for (int k = 0; k < 3; k++) { lo3[k] = dmin(lo3[k], B->bmin[3 * b + k]); hi3[k] = dmax(hi3[k], B->bmax[3 * b + k]); }
      if (c == 0 || B->rcount[b + 1] == 0) continue;
      double cost = B->trav + B->tric * ((double)c * area3(lo3, hi3) + (double)B->rcount[b + 1] * B->rarea[b + 1]) / pA;
      if (cost < best) { best = cost; best_axis = ax; best_bin = b; }
    }
  }
  if (best_axis < 0) return cnt > (int64_t)B->leaf * 4 ? s + (cnt >> 1) : -1;
  if (cnt <= B->leaf && best >= B->tric * (double)cnt) return -1;
  double lo = R->clo[best_axis], sc = bins / (R->chi[best_axis] - lo);
  int64_t i = s, j = e - 1;
  while (i <= j) {
    uint32_t t = B->order[i];
    int b = bin_of(B->cen[3 * t + best_axis], lo, sc, bins);
    if (b <= best_bin) i++;
    else { B->order[i] = B->order[j]; B->order[j] = t; j--; }
  }
  if (i == s || i == e) return s + (cnt >> 1);
  return i;
}

static void write_child(float *nodes, uint32_t node, int slot, const double lo[3], const double hi[3],
                        uint32_t ref, uint32_t count) {
  float *o = nodes + (size_t)node * OR_NODE_WORDS + slot * 8;
  uint32_t *u = (uint32_t *)o;
  o[0] = (float)lo[0]; o[1] = (float)lo[1]; o[2] = (float)lo[2]; u[3] = ref;
  o[4] = (float)hi[0]; o[5] = (float)hi[1]; o[6] = (float)hi[2]; u[7] = count;
}

typedef struct { uint32_t node; int64_t s, e, m; int depth; } frame_t;

int or_build_bvh(const float *pos, const uint32_t *idx, uint32_t ntri, int leaf_size, int bins,
                 float *nodes_out, uint32_t cap_nodes, uint32_t *order_out, or_bvh_info *info) {
  const int64_t n = ntri;
  float *tbmin = malloc(sizeof(float) * 3 * (n + 1)), *tbmax = malloc(sizeof(float) * 3 * (n + 1));
  float *cen = malloc(sizeof(float) * 3 * (n + 1));
  for (int64_t t = 0; t < n; t++) {
    uint32_t a = idx[3 * t] * 3, b = idx[3 * t + 1] * 3, c = idx[3 * t + 2] * 3;
    for (int k = 0; k < 3; k++) {
      double pa = pos[a + k], pb = pos[b + k], pc = pos[c + k];
      double lo = dmin(dmin(pa, pb), pc), hi = dmax(dmax(pa, pb), pc);
      tbmin[3 * t + k] = (float)lo; tbmax[3 * t + k] = (float)hi; cen[3 * t + k] = (float)(0.5 * (lo + hi));
    }
  }
  for (int64_t i = 0; i < n; i++) order_out[i] = (uint32_t)i;
  builder_t B = {tbmin, tbmax, cen, order_out, bins, leaf_size, 1.0, 1.0,
                 calloc(bins, 4), calloc(bins, 4), calloc(bins * 3, 8), calloc(bins * 3, 8), calloc(bins, 8)};
  uint32_t node_count = 0;
  int rc = 0, max_depth = 1;
  frame_t *stack = malloc(sizeof(frame_t) * (size_t)(n + 2));
  int64_t sp = 0;
#define ALLOC() (node_count < cap_nodes ? node_count++ : (rc = -1, 0u))
  uint32_t root = ALLOC();
  bounds_t R;
  range_bounds(&B, 0, n, &R);
  int64_t m0 = n ? split(&B, 0, n, &R) : -1;
  if (rc == 0) {
    if (n == 0) {
      double far[3] = {BIG, BIG, BIG};
      write_child(nodes_out, root, 0, far, far, 0, 1);
      write_child(nodes_out, root, 1, far, far, 0, 1);
    } else if (m0 < 0) {
      int64_t h = n > 1 ? n >> 1 : n;
      bounds_t A;
      range_bounds(&B, 0, h, &A);
      write_child(nodes_out, root, 0, A.lo, A.hi, 0, (uint32_t)h);
      if (n > 1) { range_bounds(&B, h, n, &A); write_child(nodes_out, root, 1, A.lo, A.hi, (uint32_t)h, (uint32_t)(n - h)); }
      else write_child(nodes_out, root, 1, R.lo, R.hi, 0, (uint32_t)n);
    } else {
      stack[sp++] = (frame_t){root, 0, n, m0, 1};
    }
  }
  while (sp > 0 && rc == 0) {
    frame_t f = stack[--sp];
    if (f.depth > max_depth) max_depth = f.depth;
    int64_t ra[2] = {f.s, f.m}, rb[2] = {f.m, f.e};
    for (int slot = 0; slot < 2; slot++) {
      int64_t a = ra[slot], b = rb[slot];
      bounds_t C;
      range_bounds(&B, a, b, &C);
      int64_t mm = split(&B, a, b, &C);
      if (mm < 0) write_child(nodes_out, f.node, slot, C.lo, C.hi, (uint32_t)a, (uint32_t)(b - a));
      else {
        uint32_t child = ALLOC();
        if (rc) break;
        write_child(nodes_out, f.node, slot, C.lo, C.hi, child, 0);
        stack[sp++] = (frame_t){child, a, b, mm, f.depth + 1};
      }
    }
  }
#undef ALLOC
  if (info) {
    info->node_count = node_count;
    info->max_depth = max_depth;
    for (int k = 0; k < 3; k++) { info->bmin[k] = (float)R.lo[k]; info->bmax[k] = (float)R.hi[k]; }
  }
  free(stack); free(tbmin); free(tbmax); free(cen);
  free(B.bcount); free(B.rcount); free(B.bmin); free(B.bmax); free(B.rarea);
  return rc;
}

void or_pack_triangles(const float *pos, const uint32_t *idx, const uint32_t *order, uint32_t n, float *T) {
  for (uint32_t k = 0; k < n; k++) {
    uint32_t t = order[k], a = idx[3 * t] * 3, b = idx[3 * t + 1] * 3, c = idx[3 * t + 2] * 3;
    float *o = T + (size_t)k * 16;
    o[0] = pos[a]; o[1] = pos[a + 1]; o[2] = pos[a + 2];
    /* JS: difference of two float32 values in FP64, stored into a Float32Array */
    o[4] = (float)((double)pos[b] - pos[a]); o[5] = (float)((double)pos[b + 1] - pos[a + 1]); o[6] = (float)((double)pos[b + 2] - pos[a + 2]);
    o[8] = (float)((double)pos[c] - pos[a]); o[9] = (float)((double)pos[c + 1] - pos[a + 1]); o[10] = (float)((double)pos[c + 2] - pos[a + 2]);
  }
}

double or_intersect_tri(const float *T, uint32_t k, const double o[3], const double d[3], double tmin,
                        double tmax, double *uo, double *vo) {
  const float *b = T + (size_t)k * 16;
  double e1x = b[4], e1y = b[5], e1z = b[6], e2x = b[8], e2y = b[9], e2z = b[10];
  double px = d[1] * e2z - d[2] * e2y, py = d[2] * e2x - d[0] * e2z, pz = d[0] * e2y - d[1] * e2x;
  double det = e1x * px + e1y * py + e1z * pz;
  if (fabs(det) < 1e-30) return INFINITY;
  double inv = 1 / det;
  double sx = o[0] - b[0], sy = o[1] - b[1], sz = o[2] - b[2];
  double u = (sx * px + sy * py + sz * pz) * inv;
  if (u < 0 || u > 1) return INFINITY;
  double qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
  double v = (d[0] * qx + d[1] * qy + d[2] * qz) * inv;
  if (v < 0 || u + v > 1) return INFINITY;
  double t = (e2x * qx + e2y * qy + e2z * qz) * inv;
  if (t <= tmin || t >= tmax) return INFINITY;
  if (uo) { *uo = u; *vo = v; }
  return t;
}

static double slab(const float *N, const double o[3], const double inv[3], size_t x, double tmin, double tmax) {
  double t0 = tmin, t1 = tmax;
  for (int k = 0; k < 3; k++) {
    double a = (N[x + k] - o[k]) * inv[k], b = (N[x + 4 + k] - o[k]) * inv[k];
    if (a > b) { double s = a; a = b; b = s; }
    if (a > t0) t0 = a;
    if (b < t1) t1 = b;
    if (t0 > t1) return INFINITY;
  }
  return t0;
}

void or_traverse(const float *nodes, const float *T, const double o[3], const double d[3], double tmin,
                 double tmax, int any_hit, uint32_t root, or_hit *out) {
  const uint32_t *U = (const uint32_t *)nodes;
  double inv[3] = {1 / d[0], 1 / d[1], 1 / d[2]};
  double best = tmax, bu = 0, bv = 0;
  int64_t tri = -1;
  uint64_t visits = 0, tests = 0;
  uint32_t stack[256];
  int sp = 0;
  stack[sp++] = root;
  while (sp > 0) {
    uint32_t node = stack[--sp];
    visits++;
    size_t base = (size_t)node * OR_NODE_WORDS;
    double tc[2] = {slab(nodes, o, inv, base, tmin, best), slab(nodes, o, inv, base + 8, tmin, best)};
    double pt[2];
    uint32_t pr[2];
    int np = 0;
    for (int c = 0; c < 2; c++) {
      if (!(tc[c] < best)) continue;
      uint32_t ref = U[base + c * 8 + 3], cnt = U[base + c * 8 + 7];
      if (cnt > 0) {
        for (uint32_t k = ref; k < ref + cnt; k++) {
          double u, v;
          tests++;
          double t = or_intersect_tri(T, k, o, d, tmin, best, &u, &v);
          if (t < best) {
            best = t; tri = k; bu = u; bv = v;
            if (any_hit) { out->t = t; out->tri = tri; out->u = bu; out->v = bv; out->node_visits = visits; out->tri_tests = tests; return; }
          }
        }
      } else { pt[np] = tc[c]; pr[np] = ref; np++; }
    }
    /* push.sort((a, b) => b[0] - a[0]) is stable: swap only when the second is strictly farther */
    if (np == 2 && pt[1] > pt[0]) { uint32_t r = pr[0]; pr[0] = pr[1]; pr[1] = r; }
    for (int i = 0; i < np; i++) if (sp < 256) stack[sp++] = pr[i];
  }
  out->t = tri >= 0 ? best : INFINITY;
  out->tri = tri; out->u = bu; out->v = bv;
  out->node_visits = visits; out->tri_tests = tests;
}

void or_brute_force(const float *T, uint32_t count, const double o[3], const double d[3], double tmin,
                    double tmax, or_hit *out) {
  double best = tmax, bu = 0, bv = 0;
  int64_t tri = -1;
  for (uint32_t k = 0; k < count; k++) {
    double u, v;
    double t = or_intersect_tri(T, k, o, d, tmin, best, &u, &v);
    if (t < best) { best = t; tri = k; bu = u; bv = v; }
  }
  out->t = tri >= 0 ? best : INFINITY;
  out->tri = tri; out->u = bu; out->v = bv; out->node_visits = 0; out->tri_tests = count;
}

double or_sah_cost(const float *N, uint32_t node_count, const float bmin[3], const float bmax[3],
                   double trav_cost, double tri_cost) {
  const uint32_t *U = (const uint32_t *)N;
  double ext[3] = {(double)bmax[0] - bmin[0], (double)bmax[1] - bmin[1], (double)bmax[2] - bmin[2]};
  double dx = (double)N[4] - N[0], dy = (double)N[5] - N[1], dz = (double)N[6] - N[2];
  double rootA = dx < 0 ? 0 : dx * dy + dy * dz + dz * dx;
  double A0 = ext[0] * ext[1] + ext[1] * ext[2] + ext[2] * ext[0];
  if (!(A0 != 0)) A0 = rootA;
  if (!(A0 != 0)) A0 = 1;
  double cost = 0;
  for (uint32_t i = 0; i < node_count; i++)
    for (int c = 0; c < 2; c++) {
      size_t o = (size_t)i * OR_NODE_WORDS + 8 * c;
      double ex = (double)N[o + 4] - N[o], ey = (double)N[o + 5] - N[o + 1], ez = (double)N[o + 6] - N[o + 2];
      double a = (ex < 0 ? 0 : ex * ey + ey * ez + ez * ex) / A0;
      uint32_t cnt = U[o + 7];
      cost += cnt > 0 ? a * cnt * tri_cost : a * trav_cost;
    }
  return cost + trav_cost;
}

void or_trace_batch(const float *nodes, const float *T, const float *rays, uint32_t n, int any_hit,
                    double *hits, uint64_t *visits, int nthreads) {
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const float *r = rays + 8 * i;
    double o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
    or_hit h;
    or_traverse(nodes, T, o, d, r[3], r[7], any_hit, 0, &h);
    hits[4 * i] = h.t; hits[4 * i + 1] = (double)h.tri; hits[4 * i + 2] = h.u; hits[4 * i + 3] = h.v;
    if (visits) { visits[2 * i] = h.node_visits; visits[2 * i + 1] = h.tri_tests; }
  }
}

void or_brute_batch(const float *T, uint32_t count, const float *rays, uint32_t n, double *hits, int nthreads) {
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 16)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const float *r = rays + 8 * i;
    double o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
    or_hit h;
    or_brute_force(T, count, o, d, r[3], r[7], &h);
    hits[4 * i] = h.t; hits[4 * i + 1] = (double)h.tri; hits[4 * i + 2] = h.u; hits[4 * i + 3] = h.v;
  }
}
