/* Oracle (test infrastructure only): the camera estimator, FP64 geometry and shading on FP32 scene data.
 *
 * The estimator is the build's composition of the reference's pieces (the reference itself is a hybrid
 * that reads a radiosity cache at static surfaces; the configurations ask for paths of depth D with NEE):
 *   sample pattern   pixSeeds / subSample, gpu/wgsl/kernels.wgsl.js:34-53 (binary32, bit-exact with the GPU)
 *   camera ray       gpu/lens.js:54-62 unproject (FP64, 20 Newton steps) of the binary32 sample position
 *   ray queries      gpu/bvh.js:224-250 traverse (closest, any hit), Moller-Trumbore gpu/bvh.js:189-206
 *   surface          gpu/wgsl/trace.wgsl.js:330-358 (two-sided normals, 2e-6 m origin offset)
 *   emitter sample   gpu/wgsl/shade.wgsl.js:31-47 (area CDF, sqrt-u point on the triangle)
 *   NEE + MIS        gpu/wgsl/shade.wgsl.js:160-188 (solid-angle pdf d^2/(cos_e A), power heuristic,
 *                    shadow ray to dist - 2e-5); emitter hit by a BSDF ray weighted as :224-236
 *   BSDF             gpu/bsdf.js (+ the Lambertian kind)
 * Path of depth D: vertices v1..vD. NEE with MIS at v1..v(D-1), NEE without MIS at vD (no continuation from
 * vD, so its light paths are carried by NEE alone); a BSDF continuation from vk (k < D) that hits an emitter
 * adds its MIS-weighted emission and ends the path. An emitter seen by the camera ray adds its radiance.
 * Random numbers: the reference's Kronecker dimensions at v1; counter hashes u01(hash3(hs, i, 256 b + dim))
 * from v2 on (the same on the GPU). No Russian roulette.
 * Pins (tests/test_oracle_pins.py): white furnace inside a closed Lambertian sphere (L = Le rho^k sums), the
 * on-axis square-emitter irradiance closed form through NEE, MIS continuations against NEE alone; the adjoint
 * against central finite differences of this forward estimator (tests/test_adjoint.py).
 */
#include "vt_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define F_EMITTER 2u
#define F_STATIC 1u
#define O_NO_INDIRECT 4u
#define O_DIRECT_ONLY 8u
#define O_NO_CAPSULE 32u
#define O_NORMALS 64u
#define O_NO_MIS 128u
#define O_REFERENCE (1u << 16)
#define O_CULL_SHADOWS (1u << 17)
#define O_REFINE_SHADOW 1024u
#define O_SHADOW_CACHE 8192u
static const float R2A = 0.7548776662466927f, R2B = 0.5698402909980532f;
static const float KL0 = 0.41421356237f, KL1 = 0.73205080757f, KL2 = 0.23606797750f;
static const float KB0 = 0.64575131106f, KB1 = 0.31662479036f, KB2 = 0.60555127546f;
static const float GOLD_F = 0.6180339887498949f;   /* sub-sample time dimension (kernels.wgsl.js:44) */

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void norm3(double *a) {
  double l = sqrt(dot3(a, a));
  a[0] /= l; a[1] /= l; a[2] /= l;
}
static inline double maxc(const double *a) { return fmax(a[0], fmax(a[1], a[2])); }

typedef struct { uint32_t hs; float s[9]; } pix_t;

static void pix_seeds(const or_camera *C, uint32_t x, uint32_t y, pix_t *P) {
  P->hs = or_hash3(x, y, C->seed * 64u + C->cam_index);
  for (int k = 0; k < 9; k++) P->s[k] = or_u01(or_pcg(P->hs + (uint32_t)k));
}
static inline double rnd(const pix_t *P, uint32_t i, uint32_t b, uint32_t dim) {
  return or_u01(or_hash3(P->hs, i, 256u * b + dim));
}

static void cam_dir(const or_camera *C, const pix_t *P, uint32_t x, uint32_t y, uint32_t i, double d[3]) {
  float uv[2];
  {
    volatile float a = (float)x + or_kr(i, R2A, P->s[0]);
    volatile float b = (float)y + or_kr(i, R2B, P->s[1]);
    uv[0] = a - 0.5f; uv[1] = b - 0.5f;
  }
  double dc[3];
  if (C->mode == 0) {
    or_lens L = {C->lens[0], C->lens[1], C->lens[2], C->lens[3], C->lens[8], {C->lens[4], C->lens[5], C->lens[6], C->lens[7]}};
    double duv[2] = {uv[0], uv[1]};
    or_unproject(&L, duv, dc);
  } else {
    dc[0] = ((double)uv[0] - C->lens[2]) / C->lens[0];
    dc[1] = ((double)uv[1] - C->lens[3]) / C->lens[1];
    dc[2] = 1;
    norm3(dc);
  }
  for (int k = 0; k < 3; k++) d[k] = (double)C->R[k] * dc[0] + (double)C->R[3 + k] * dc[1] + (double)C->R[6 + k] * dc[2];
  norm3(d);
}

void or_camera_ray(const or_camera *C, uint32_t x, uint32_t y, uint32_t i, float o[3], float d[3]) {
  pix_t P;
  pix_seeds(C, x, y, &P);
  double dd[3];
  cam_dir(C, &P, x, y, i, dd);
  for (int k = 0; k < 3; k++) { o[k] = C->pos[k]; d[k] = (float)dd[k]; }
}

typedef struct {
  double p[3], ng[3], ns[3], b[3];
  double nsi[3];    /* the interpolated vertex normal in the world, before any flip (an emitter's outward normal) */
  uint32_t vi[3], mesh, mat, flags, light;
  int ns_from_ng;   /* the shading normal fell back to the geometric one (grazing view of the interpolated normal) */
} surf_t;

/* ------------------------------------------------------------------ posed bodies
 * pose of a body at exposure time t: slerp of the quaternions (the shorter arc, normalised lerp when the ends
 * nearly agree: gpu/renderer.js:85-91 slerpJS, common.wgsl.js:125-133) and lerp of the positions; R holds the
 * rotation columns (common.wgsl.js:117-123 quatToMat) */
typedef struct { double R[9], p[3]; } pose_t;

static void body_pose(const or_body *B, double t, pose_t *P) {
  double a[4], b[4], q[4], c = 0;
  for (int k = 0; k < 4; k++) { a[k] = B->q0[k]; b[k] = B->q1[k]; }
  double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3]);
  double lb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2] + b[3] * b[3]);
  for (int k = 0; k < 4; k++) { a[k] /= la; b[k] /= lb; c += a[k] * b[k]; }
  if (c < 0) { for (int k = 0; k < 4; k++) b[k] = -b[k]; c = -c; }
  if (c > 0.9995) {
    double l = 0;
    for (int k = 0; k < 4; k++) { q[k] = a[k] + (b[k] - a[k]) * t; l += q[k] * q[k]; }
    l = sqrt(l);
    for (int k = 0; k < 4; k++) q[k] /= l;
  } else {
    double th = acos(c), sn = sin(th);
    for (int k = 0; k < 4; k++) q[k] = (a[k] * sin((1 - t) * th) + b[k] * sin(t * th)) / sn;
  }
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  const double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w),
                       2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w),
                       2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)};
  memcpy(P->R, R, sizeof R);
  for (int k = 0; k < 3; k++) P->p[k] = B->p0[k] + ((double)B->p1[k] - B->p0[k]) * t;
}
static void rot_apply(const double R[9], const double v[3], double out[3]) {
  for (int k = 0; k < 3; k++) out[k] = R[k] * v[0] + R[3 + k] * v[1] + R[6 + k] * v[2];
}
static void rot_apply_t(const double R[9], const double v[3], double out[3]) {
  for (int k = 0; k < 3; k++) out[k] = R[3 * k] * v[0] + R[3 * k + 1] * v[1] + R[3 * k + 2] * v[2];
}

/* traceSceneSkip (trace.wgsl.js:144-164; skip: bit i = ignore body i) without its culling tests (box and capsule only skip work): the static
 * BVH, then every posed body in its own frame at time t; a rigid pose keeps ray lengths, so t carries over.
 * h->inst names the body of the hit (-1: static); *P receives its pose. Visits add up over every BVH walked. */
static void trace_scene(const or_scene *S, const double o[3], const double d[3], double tmin, double tmax, int any,
                        double time, uint32_t skip, or_hit *h, pose_t *P) {
  or_traverse(S->nodes, S->tris, o, d, tmin, tmax, any, 0, h);
  h->inst = -1;
  if (any && h->tri >= 0) return;
  for (uint32_t b = 0; b < S->body_count; b++) {
    const or_body *B = &S->bodies[b];
    if (B->hidden || ((skip >> b) & 1u)) continue;
    pose_t Q;
    body_pose(B, time, &Q);
    double ol[3], dl[3], rel[3] = {o[0] - Q.p[0], o[1] - Q.p[1], o[2] - Q.p[2]};
    rot_apply_t(Q.R, rel, ol);
    rot_apply_t(Q.R, d, dl);
    const double best = h->tri >= 0 ? h->t : tmax;
    or_hit hb;
    or_traverse(B->nodes, B->tris, ol, dl, tmin, best, any, 0, &hb);
    const uint64_t nv = h->node_visits + hb.node_visits, nt = h->tri_tests + hb.tri_tests;
    if (hb.tri >= 0 && hb.t < best) {
      *h = hb;
      h->inst = (int32_t)b;
      if (P) *P = Q;
    }
    h->node_visits = nv;
    h->tri_tests = nt;
    if (any && h->tri >= 0) return;
  }
}

static void surface(const or_scene *S, const or_hit *h, const pose_t *P, const double o[3], const double d[3],
                    surf_t *s) {
  const float *r = (h->inst >= 0 ? S->bodies[h->inst].tris : S->tris) + (size_t)h->tri * 16;
  const uint32_t *ru = (const uint32_t *)r;
  s->vi[0] = ru[3]; s->vi[1] = ru[7]; s->vi[2] = ru[11];
  s->mesh = ru[12];
  const uint32_t *m = S->meshes + 4 * (size_t)s->mesh;
  s->mat = m[0]; s->flags = m[1]; s->light = m[2];
  s->b[0] = 1 - h->u - h->v; s->b[1] = h->u; s->b[2] = h->v;
  double e1[3] = {r[4], r[5], r[6]}, e2[3] = {r[8], r[9], r[10]};
  s->ng[0] = e1[1] * e2[2] - e1[2] * e2[1]; s->ng[1] = e1[2] * e2[0] - e1[0] * e2[2]; s->ng[2] = e1[0] * e2[1] - e1[1] * e2[0];
  norm3(s->ng);
  for (int k = 0; k < 3; k++) {
    s->ns[k] = 0;
    for (int j = 0; j < 3; j++) s->ns[k] += s->b[j] * S->verts[8 * (size_t)s->vi[j] + 4 + k];
  }
  norm3(s->ns);
  if (h->inst >= 0) {
    double a[3];
    rot_apply(P->R, s->ng, a);
    memcpy(s->ng, a, sizeof a);
    rot_apply(P->R, s->ns, a);
    memcpy(s->ns, a, sizeof a);
  }
  memcpy(s->nsi, s->ns, sizeof s->nsi);
  if (dot3(s->ng, d) > 0) for (int k = 0; k < 3; k++) s->ng[k] = -s->ng[k];
  if (dot3(s->ns, s->ng) < 0) for (int k = 0; k < 3; k++) s->ns[k] = -s->ns[k];
  double md[3] = {-d[0], -d[1], -d[2]};
  s->ns_from_ng = dot3(s->ns, md) < 1e-4;
  if (s->ns_from_ng) memcpy(s->ns, s->ng, sizeof(s->ng));
  for (int k = 0; k < 3; k++) s->p[k] = o[k] + d[k] * h->t;
}

float or_board_mask(const uint8_t *tex, uint32_t w, uint32_t h, float u, float v) {
  const float fw = (float)w, fh = (float)h;
  float x = u * fw - 0.5f, y = v * fh - 0.5f;
  x = x < -1.0f ? -1.0f : (x > fw ? fw : x);
  y = y < -1.0f ? -1.0f : (y > fh ? fh : y);
  const float x0 = floorf(x), y0 = floorf(y);
  const float ax = x - x0, ay = y - y0;
  int i0 = (int)x0, i1 = (int)x0 + 1, j0 = (int)y0, j1 = (int)y0 + 1;
  const int lx = (int)w - 1, ly = (int)h - 1;
  i0 = i0 < 0 ? 0 : (i0 > lx ? lx : i0); i1 = i1 < 0 ? 0 : (i1 > lx ? lx : i1);
  j0 = j0 < 0 ? 0 : (j0 > ly ? ly : j0); j1 = j1 < 0 ? 0 : (j1 > ly ? ly : j1);
  const float s = 1.0f / 255.0f;
  const float t00 = (float)tex[(size_t)j0 * w + i0] * s, t10 = (float)tex[(size_t)j0 * w + i1] * s;
  const float t01 = (float)tex[(size_t)j1 * w + i0] * s, t11 = (float)tex[(size_t)j1 * w + i1] * s;
  const float top = t00 * (1.0f - ax) + t10 * ax, bot = t01 * (1.0f - ax) + t11 * ax;
  return top * (1.0f - ay) + bot * ay;
}

/* the material record at a hit, mixed between ink and paper on a textured mesh (shade.wgsl.js:258-265):
 * c = ink (1 - m) + paper m and alpha likewise, m the mask at the interpolated texture coordinates */
static void hit_material(const or_scene *S, const surf_t *s, const float *rec, or_mat *M) {
  or_mat_unpack(rec, M);
  if (!(s->flags & 128u) || !S->uvs) return;
  double uv[2] = {0, 0};
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 2; k++) uv[k] += s->b[j] * S->uvs[2 * (size_t)s->vi[j] + k];
  static const uint8_t paper = 255;
  const float m = S->tex ? or_board_mask(S->tex, S->tex_w, S->tex_h, (float)uv[0], (float)uv[1])
                         : or_board_mask(&paper, 1, 1, (float)uv[0], (float)uv[1]);
  const float k = 1.0f - m;
  for (int c = 0; c < 3; c++) M->c[c] = (float)((float)M->tau[c] * k + (float)M->c[c] * m);
  M->alpha = (float)((float)M->alpha_c * k + (float)M->alpha * m);
}

/* angular profile of light `light` toward the unit direction w leaving the emitter (shade.wgsl.js:48-52 emitted:
 * p(theta) with theta = acos(w . axis), or_profile over the light's knots); 1 for a Lambertian light */
static double light_profile(const or_scene *S, uint32_t light, const double w[3]) {
  if (!S->light_prof) return 1;
  const float *a = S->light_prof + 4 * (size_t)light;
  uint32_t cnt;
  memcpy(&cnt, a + 3, 4);
  if (cnt == 0) return 1;
  const double ax[3] = {a[0], a[1], a[2]};
  const double th = acos(fmin(1.0, fmax(-1.0, dot3(w, ax))));
  double kt[64], kp[64];
  for (uint32_t k = 0; k < cnt && k < 64; k++) {
    kt[k] = S->knots[2 * ((size_t)light * 64 + k)];
    kp[k] = S->knots[2 * ((size_t)light * 64 + k) + 1];
  }
  return or_profile(th, kt, kp, (int)(cnt < 64 ? cnt : 64));
}

/* outward emitter normal at a hit: the interpolated vertex normal in the world, unflipped (shade.wgsl.js:217-220) */
static void emitter_normal(const or_scene *S, const surf_t *s, double n[3]) {
  (void)S;
  memcpy(n, s->nsi, 3 * sizeof(double));
}

/* The emitter's frame at exposure time t: the pose of the body the light rides on (the reference's outlet on the
 * scope, shade.wgsl.js:45, 55), 0 (and no pose) for a light in the static scene. */
static int emit_pose(const or_scene *S, double t, pose_t *P) {
  if (!S->emit_posed) return 0;
  body_pose(&S->bodies[S->emit_body], t, P);
  return 1;
}
/* a world direction in the emitter's frame (R^T w), where the light axes are */
static void emit_local(const or_scene *S, double t, const double w[3], double out[3]) {
  pose_t P;
  if (!emit_pose(S, t, &P)) { memcpy(out, w, 3 * sizeof(double)); return; }
  rot_apply_t(P.R, w, out);
}
/* light_profile toward the world direction w at exposure time t */
static double light_profile_at(const or_scene *S, uint32_t light, const double w[3], double t) {
  double wl[3];
  emit_local(S, t, w, wl);
  return light_profile(S, light, wl);
}
/* the emitter centre (emit_c, emitter frame) in the world at exposure time t */
static void emit_centre(const or_scene *S, double t, double c[3]) {
  pose_t P;
  const double cl[3] = {S->emit_c[0], S->emit_c[1], S->emit_c[2]};
  if (!emit_pose(S, t, &P)) { memcpy(c, cl, sizeof cl); return; }
  rot_apply(P.R, cl, c);
  for (int k = 0; k < 3; k++) c[k] += P.p[k];
}

static void offset_ray(const double p[3], const double ng[3], const double dir[3], double out[3]) {
  double s = dot3(ng, dir) > 0 ? 1 : -1;
  for (int k = 0; k < 3; k++) out[k] = p[k] + ng[k] * (s * 2e-6);
}

/* shade.wgsl.js:31-47: area-CDF pick, uniform point, posed at exposure time t with the light's body; p and the
 * emitter normal n in the world; returns the emitter record */
static const float *sample_emitter(const or_scene *S, const double u[3], double t, double p[3], double n[3]) {
  uint32_t lo = 0, hi = S->emit_count - 1;
  for (int it = 0; it < 40; it++) {
    if (lo >= hi) break;
    uint32_t m = (lo + hi) / 2;
    if ((double)S->emit[16 * (size_t)m + 3] < u[0]) lo = m + 1; else hi = m;
  }
  const float *e = S->emit + 16 * (size_t)lo;
  double su = sqrt(u[1]);
  for (int k = 0; k < 3; k++) {
    p[k] = (double)e[k] + (double)e[4 + k] * (su * (1 - u[2])) + (double)e[8 + k] * (su * u[2]);
    n[k] = e[12 + k];
  }
  pose_t P;
  if (emit_pose(S, t, &P)) {
    double a[3];
    rot_apply(P.R, p, a);
    for (int k = 0; k < 3; k++) p[k] = a[k] + P.p[k];
    rot_apply(P.R, n, a);
    memcpy(n, a, sizeof a);
  }
  return e;
}

/* gradient sink of the path-replay adjoint (FP64): d(delta . L)/d(material colour), d/d(light radiance) */
typedef struct {
  double *mat, *light;   /* n_mat * 3, n_light * 3 */
  double delta[3];
  double Ltot[3];
} grad_t;

/* d f / d c (c = rho_d of dielectric and Lambert, F0 of conductor and coated), FP64 twin of the GPU's */
static void eval_bsdf_dc(const or_mat *M, const double n[3], const double l[3], const double v[3], double out[3]) {
  double nl = dot3(n, l), nv = dot3(n, v);
  out[0] = out[1] = out[2] = 0;
  if (nl <= 0 || nv <= 0) return;
  if (M->kind == OR_LAMBERT) { out[0] = out[1] = out[2] = 1 / M_PI; return; }
  if (M->kind == OR_DIELECTRIC) {
    double kd = (1 / M_PI) * (1 - or_fresnel_dielectric(nl, M->ior)) * (1 - or_fresnel_dielectric(nv, M->ior)) / (1 - M->fbar);
    out[0] = out[1] = out[2] = kd;
    return;
  }
  double h[3] = {l[0] + v[0], l[1] + v[1], l[2] + v[2]};
  norm3(h);
  double nh = dot3(n, h), vh = fmax(0, dot3(v, h));
  double one_m = 1 - pow(1 - fmin(fmax(vh, 0), 1), 5);
  if (M->kind == OR_CONDUCTOR) {
    double spec = or_ggx_d(nh, M->alpha) * or_ggx_g2(nl, nv, M->alpha) / (4 * nl * nv);
    out[0] = out[1] = out[2] = spec * one_m;
    return;
  }
  double t = (1 - or_fresnel_dielectric(nl, M->ior)) * (1 - or_fresnel_dielectric(nv, M->ior));
  double metal = or_ggx_d(nh, M->alpha) * or_ggx_g2(nl, nv, M->alpha) / (4 * nl * nv);
  for (int c = 0; c < 3; c++) out[c] = pow(M->tau[c], 1 / nl + 1 / nv) * t * metal * one_m;
}

/* seq (optional): a running hash of every ray's result (hit triangle or miss), so a re-run with a perturbed
 * camera ray can tell whether the path kept its triangles and visibility; ray0 (optional): the camera ray
 * (o, d) in place of the camera's own */
static void seq_add(uint64_t *seq, int64_t tri) {
  if (seq) *seq = (*seq ^ (uint64_t)(tri + 2)) * 0x100000001b3ull;
}

/* A path's trace results, for the camera gradient's interior term: recorded on the unperturbed path (mode 1: each
 * extension ray's triangle and body, each shadow ray's outcome, tagged by vertex and light sample), then replayed
 * on the perturbed ones (mode 2): an extension hit is the ray's intersection with the recorded triangle's plane
 * (barycentrics may leave the triangle) and a shadow ray keeps its recorded outcome, so the derivative holds the
 * triangles and visibility fixed, as the forward-mode tangents do. A replay whose events stop matching (a branch
 * on a value flipped) sets `bad` and traces normally from there. */
#define PATH_EVENTS 256
typedef struct {
  int mode, n, bad;
  double min_cos;    /* record: the smallest |d . n_g| of the path's hits */
  /* each vertex's continuation direction and pdf (record); a replay continues along them (sampling detached) */
  int n_cont;
  double cont_l[16][3], cont_pdf[16];
  uint32_t tag[PATH_EVENTS];
  int64_t tri[PATH_EVENTS];
  int32_t inst[PATH_EVENTS];
} path_rec_t;

/* the ray's hit on the recorded triangle's plane (no range test on u, v, t) */
static void plane_hit(const or_scene *S, int64_t tri, int32_t inst, const double o[3], const double d[3], double time,
                      or_hit *h, pose_t *P) {
  double ol[3] = {o[0], o[1], o[2]}, dl[3] = {d[0], d[1], d[2]};
  const float *T = S->tris;
  if (inst >= 0) {
    const or_body *B = &S->bodies[inst];
    body_pose(B, time, P);
    double rel[3] = {o[0] - P->p[0], o[1] - P->p[1], o[2] - P->p[2]};
    rot_apply_t(P->R, rel, ol);
    rot_apply_t(P->R, d, dl);
    T = B->tris;
  }
  const float *b = T + (size_t)tri * 16;
  const double e1[3] = {b[4], b[5], b[6]}, e2[3] = {b[8], b[9], b[10]};
  const double pv[3] = {dl[1] * e2[2] - dl[2] * e2[1], dl[2] * e2[0] - dl[0] * e2[2], dl[0] * e2[1] - dl[1] * e2[0]};
  const double inv = 1 / dot3(e1, pv);
  const double sv[3] = {ol[0] - b[0], ol[1] - b[1], ol[2] - b[2]};
  const double qv[3] = {sv[1] * e1[2] - sv[2] * e1[1], sv[2] * e1[0] - sv[0] * e1[2], sv[0] * e1[1] - sv[1] * e1[0]};
  memset(h, 0, sizeof *h);
  h->tri = tri;
  h->inst = inst;
  h->u = dot3(sv, pv) * inv;
  h->v = dot3(dl, qv) * inv;
  h->t = dot3(e2, qv) * inv;
}

/* one event of the record / replay; returns 1 when the replayed event stands in for tracing */
static int path_event(path_rec_t *pr, uint32_t tag, int64_t *tri, int32_t *inst) {
  if (!pr || pr->bad) return 0;
  if (pr->mode == 1) {
    if (pr->n >= PATH_EVENTS) { pr->bad = 1; return 0; }
    pr->tag[pr->n] = tag; pr->tri[pr->n] = *tri; pr->inst[pr->n] = *inst;
    pr->n++;
    return 0;
  }
  if (pr->n >= PATH_EVENTS || pr->tag[pr->n] != tag) { pr->bad = 1; return 0; }
  *tri = pr->tri[pr->n];
  *inst = pr->inst[pr->n];
  pr->n++;
  return 1;
}

static double sample_radiance_ex(const or_scene *S, const or_camera *C, const pix_t *P, uint32_t x, uint32_t y,
                                 uint32_t i, double out[3], uint64_t *rays, uint64_t *visits, grad_t *G,
                                 const double *ray0, uint64_t *seq, path_rec_t *pr) {
  double d[3], o[3] = {C->pos[0], C->pos[1], C->pos[2]};
  if (ray0) {
    memcpy(o, ray0, sizeof o);
    memcpy(d, ray0 + 3, sizeof d);
  } else {
    cam_dir(C, P, x, y, i, d);
  }
  double thr[3] = {1, 1, 1}, L[3] = {0, 0, 0}, pB = 0;
  const float tk = or_kr(i, GOLD_F, P->s[2]);
  const double time = (double)(C->shutter[0] + (C->shutter[1] - C->shutter[0]) * tk);
  pose_t pose;
  const uint32_t nl = C->light_samples < 1 ? 1 : C->light_samples > 16 ? 16 : C->light_samples;
  const uint32_t D = C->max_depth < 1 ? 1 : C->max_depth;
  const double A = S->emit_area;
  for (uint32_t b = 0; b < D; b++) {
    or_hit h;
    {
      int64_t rt = 0;
      int32_t ri = -1;
      if (pr && pr->mode == 2 && path_event(pr, b << 8, &rt, &ri)) {
        if (rt >= 0) plane_hit(S, rt, ri, o, d, time, &h, &pose);
        else { memset(&h, 0, sizeof h); h.tri = -1; h.inst = -1; }
      } else {
        trace_scene(S, o, d, 0, INFINITY, 0, time, b == 0 ? C->skip_bodies : 0u, &h, &pose);
        rt = h.tri; ri = h.inst;
        if (pr && pr->mode == 1) path_event(pr, b << 8, &rt, &ri);
      }
    }
    (*rays)++; visits[0] += h.node_visits; visits[1] += h.tri_tests;
    visits[2]++; visits[3] += h.node_visits; visits[4] += h.tri_tests;
    seq_add(seq, h.tri);
    if (h.tri < 0) break;
    surf_t s;
    surface(S, &h, &pose, o, d, &s);
    if (pr && pr->mode == 1) {
      const double c = fabs(dot3(d, s.ng));
      if (c < pr->min_cos) pr->min_cos = c;
    }
    if (pr) {
      /* discrete choices of the vertex that the replay must reproduce: the shading-normal fallback */
      int64_t v = s.ns_from_ng;
      int32_t none = -1;
      const int64_t rec = v;
      if (path_event(pr, (b << 8) | 0xfe, &v, &none) && v != rec) pr->bad = 1;
    }
    if (s.flags & F_EMITTER) {
      double ne[3], md[3] = {-d[0], -d[1], -d[2]};
      emitter_normal(S, &s, ne);
      double cosE = dot3(md, ne);
      if (cosE > 0) {
        const float *Le = S->lights + 4 * (size_t)s.light;
        const double pe = light_profile_at(S, s.light, md, time);
        double w = 1;
        if (b > 0) {
          double pL = h.t * h.t / (cosE * A);
          w = pB * pB / (pB * pB + nl * nl * pL * pL);
        }
        for (int k = 0; k < 3; k++) L[k] += thr[k] * Le[k] * pe * w;
        if (G) for (int k = 0; k < 3; k++) G->light[3 * (size_t)s.light + k] += G->delta[k] * thr[k] * w * pe;
      }
      break;
    }
    or_mat M, Ms;
    hit_material(S, &s, S->materials + 16 * (size_t)s.mat, &M);
    Ms = M;
    if (S->sample_materials) hit_material(S, &s, S->sample_materials + 16 * (size_t)s.mat, &Ms);
    double wo[3] = {-d[0], -d[1], -d[2]};
    const int last = b + 1 == D;
    double Ln[3] = {0, 0, 0};
    for (uint32_t j = 0; j < nl; j++) {
      double u[3];
      if (b == 0) {
        uint32_t li = i * nl + j;
        u[0] = or_kr(li, KL0, P->s[3]); u[1] = or_kr(li, KL1, P->s[4]); u[2] = or_kr(li, KL2, P->s[5]);
      } else {
        u[0] = rnd(P, i, b, 3 * j); u[1] = rnd(P, i, b, 3 * j + 1); u[2] = rnd(P, i, b, 3 * j + 2);
      }
      double pl[3], en[3];
      const float *e = sample_emitter(S, u, time, pl, en);
      double wi[3] = {pl[0] - s.p[0], pl[1] - s.p[1], pl[2] - s.p[2]};
      double d2 = dot3(wi, wi), dist = sqrt(d2);
      for (int k = 0; k < 3; k++) wi[k] /= dist;
      double cosE = -dot3(wi, en);
      if (cosE <= 0) continue;
      double cosS = dot3(s.ns, wi);
      if (cosS <= 0 || dot3(s.ng, wi) <= 0) continue;
      uint32_t lid;
      memcpy(&lid, e + 11, 4);
      const float *L0 = S->lights + 4 * (size_t)lid;
      const double mwi[3] = {-wi[0], -wi[1], -wi[2]};
      const double pe = light_profile_at(S, lid, mwi, time);
      const double Le[3] = {L0[0] * pe, L0[1] * pe, L0[2] * pe};
      if (fmax(Le[0], fmax(Le[1], Le[2])) <= 0) continue;
      double f[3];
      or_eval_bsdf(&M, s.ns, wi, wo, f);
      if (maxc(f) <= 0) continue;
      double so[3];
      offset_ray(s.p, s.ng, wi, so);
      or_hit sh;
      {
        int64_t rt = 0;
        int32_t ri = -1;
        const uint32_t tag = (b << 8) | (j + 1);
        if (pr && pr->mode == 2 && path_event(pr, tag, &rt, &ri)) {
          memset(&sh, 0, sizeof sh);
          sh.tri = rt;
        } else {
          trace_scene(S, so, wi, 0, dist - 2e-5, 1, time, S->shadow_skip, &sh, NULL);
          rt = sh.tri >= 0 ? 0 : -1;
          if (pr && pr->mode == 1) path_event(pr, tag, &rt, &ri);
        }
      }
      (*rays)++; visits[0] += sh.node_visits; visits[1] += sh.tri_tests;
      seq_add(seq, sh.tri >= 0 ? 0 : -1);
      if (sh.tri >= 0) continue;
      double pW = d2 / (cosE * A);
      double w = 1;
      if (!last) {
        double pb = or_pdf_bsdf(&Ms, s.ns, wo, wi);
        w = (double)nl * nl * pW * pW / ((double)nl * nl * pW * pW + pb * pb);
      }
      const double kk = cosS * w / pW;
      for (int k = 0; k < 3; k++) Ln[k] += f[k] * Le[k] * kk;
      if (G) {
        double dc[3];
        eval_bsdf_dc(&M, s.ns, wi, wo, dc);
        for (int k = 0; k < 3; k++) {
          G->light[3 * (size_t)lid + k] += G->delta[k] * thr[k] / nl * f[k] * kk * pe;
          G->mat[3 * (size_t)s.mat + k] += G->delta[k] * thr[k] / nl * dc[k] * Le[k] * kk;
        }
      }
    }
    for (int k = 0; k < 3; k++) L[k] += thr[k] * Ln[k] / nl;
    if (last) break;
    double ub[3];
    if (b == 0) { ub[0] = or_kr(i, KB0, P->s[6]); ub[1] = or_kr(i, KB1, P->s[7]); ub[2] = or_kr(i, KB2, P->s[8]); }
    else { ub[0] = rnd(P, i, b, 200); ub[1] = rnd(P, i, b, 201); ub[2] = rnd(P, i, b, 202); }
    double l[3];
    double pdf;
    if (pr && pr->mode == 2 && !pr->bad) {
      /* replay: the recorded direction and pdf (sampling detached), whatever the moved vertex would sample */
      if (b >= (uint32_t)pr->n_cont) { pr->bad = 1; break; }
      memcpy(l, pr->cont_l[b], sizeof l);
      pdf = pr->cont_pdf[b];
    } else {
      pdf = or_sample_bsdf(&Ms, s.ns, wo, ub, l);
      if (pr && pr->mode == 1 && b < 16) {   /* pdf 0 too: the replay ends where the record did */
        memcpy(pr->cont_l[b], l, sizeof l);
        pr->cont_pdf[b] = pdf;
        pr->n_cont = (int)b + 1;
      }
    }
    if (!(pdf > 0) || dot3(s.ng, l) <= 0) break;
    double f[3];
    or_eval_bsdf(&M, s.ns, l, wo, f);
    double c = dot3(s.ns, l) / pdf;
    if (G) {
      double dc[3];
      eval_bsdf_dc(&M, s.ns, l, wo, dc);
      for (int k = 0; k < 3; k++) {
        double a = f[k] * c;
        if (a > 0) G->mat[3 * (size_t)s.mat + k] += G->delta[k] * (dc[k] * c / a) * (G->Ltot[k] - L[k]);
      }
    }
    for (int k = 0; k < 3; k++) thr[k] *= f[k] * c;
    if (maxc(thr) <= 0) break;
    offset_ray(s.p, s.ng, l, o);
    memcpy(d, l, sizeof(l));
    pB = pdf;
  }
  memcpy(out, L, sizeof(L));
  return 0;
}

static double sample_radiance(const or_scene *S, const or_camera *C, const pix_t *P, uint32_t x, uint32_t y,
                              uint32_t i, double out[3], uint64_t *rays, uint64_t *visits, grad_t *G) {
  return sample_radiance_ex(S, C, P, x, y, i, out, rays, visits, G, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------ the reference camera estimator
 * shadeHit / directLight / secondary / cachedVisibility / capsuleVis (gpu/wgsl/shade.wgsl.js:83-334) on the node
 * cache of or_scene, in FP64. */
static void albedo_at(const or_scene *S, uint32_t m, double mu, double out[3]) {
  double x = mu * 32 - 0.5;
  if (x < 0) x = 0;
  if (x > 31) x = 31;
  int i = (int)floor(x);
  int j = i + 1 < 31 ? i + 1 : 31;
  double f = x - i;
  const float *b = S->alb + (size_t)m * 32 * 8;
  for (int c = 0; c < 3; c++)
    out[c] = ((double)b[8 * i + c] * (1 - f) + (double)b[8 * j + c] * f) + ((double)b[8 * i + 4 + c] * (1 - f) + (double)b[8 * j + 4 + c] * f);
}

static void body_pose_mid(const or_body *B, pose_t *P) { body_pose(B, 0.5, P); }

/* cavityCapsule (gpu/scene_gpu.js:486-503) on the posed capsule a-b of radius *r: the part with s >= s0 */
static void cavity_cut(const or_body *B, double a[3], double b[3], double *r) {
  if (!B->cavity) return;
  if (B->cavity_r > 0) *r = B->cavity_r;
  double ax[3] = {B->cavity_axis[0], B->cavity_axis[1], B->cavity_axis[2]};
  norm3(ax);
  double sa = 0, sb = 0;
  for (int k = 0; k < 3; k++) {
    sa += (a[k] - B->cavity_origin[k]) * ax[k];
    sb += (b[k] - B->cavity_origin[k]) * ax[k];
  }
  const double s0 = B->cavity_s0;
  if (sa < s0 && sb < s0) { *r = 0; return; }
  double a0[3], b0[3];
  memcpy(a0, a, sizeof a0);
  memcpy(b0, b, sizeof b0);
  if (sa < s0) for (int k = 0; k < 3; k++) a[k] = a0[k] + (b0[k] - a0[k]) * ((s0 - sa) / (sb - sa));
  else if (sb < s0) for (int k = 0; k < 3; k++) b[k] = b0[k] + (a0[k] - b0[k]) * ((s0 - sb) / (sa - sb));
}

double or_capsule_vis(const or_scene *S, const double p[3], const double n[3]) {
  double v = 1;
  for (uint32_t b = 0; b < S->body_count; b++) {
    const or_body *B = &S->bodies[b];
    if (B->hidden || !(B->occ_r > 0)) continue;
    pose_t Q;
    body_pose_mid(B, &Q);
    double a[3], bb[3], r = B->occ_r;
    const double oa[3] = {B->occ_a[0], B->occ_a[1], B->occ_a[2]}, ob[3] = {B->occ_b[0], B->occ_b[1], B->occ_b[2]};
    rot_apply(Q.R, oa, a);
    rot_apply(Q.R, ob, bb);
    for (int k = 0; k < 3; k++) { a[k] += Q.p[k]; bb[k] += Q.p[k]; }
    cavity_cut(B, a, bb, &r);
    if (!(r > 0)) continue;
    v -= or_capsule_occ(p, n, a, bb, r) * (1 - B->rho_eff);
  }
  return v < 0 ? 0 : v > 1 ? 1 : v;
}

/* segment o + t d, t in [0, 1], within r of segment a-b (segNearCapsule, trace.wgsl.js:122-136) */
static int seg_near(const double o[3], const double d[3], const double a[3], const double b[3], double r) {
  double v[3], w[3];
  for (int k = 0; k < 3; k++) { v[k] = b[k] - a[k]; w[k] = o[k] - a[k]; }
  const double A = dot3(d, d), B = dot3(d, v), Cc = dot3(v, v), D = dot3(d, w), E = dot3(v, w);
  const double den = A * Cc - B * B;
  double s = 0;
  if (den > 1e-20 * A * Cc) s = fmin(fmax((B * E - Cc * D) / den, 0), 1);
  double t = Cc > 1e-30 ? (B * s + E) / Cc : 0;
  if (t < 0) { t = 0; s = fmin(fmax(-D / fmax(A, 1e-30), 0), 1); }
  else if (t > 1) { t = 1; s = fmin(fmax((B - D) / fmax(A, 1e-30), 0), 1); }
  double q[3];
  for (int k = 0; k < 3; k++) q[k] = (o[k] + d[k] * s) - (a[k] + v[k] * t);
  return dot3(q, q) <= r * r;
}

static double cached_visibility(const or_scene *S, const surf_t *s, double time) {
  const double a = S->node_state[16 * (size_t)s->vi[0] + 15], b = S->node_state[16 * (size_t)s->vi[1] + 15],
               c = S->node_state[16 * (size_t)s->vi[2] + 15];
  if (!(a == b && b == c && (a == 1.0 || a == 0.0))) return -1;
  double ec[3];
  emit_centre(S, time, ec);
  const double d[3] = {ec[0] - s->p[0], ec[1] - s->p[1], ec[2] - s->p[2]};
  for (uint32_t i = 0; i < S->body_count; i++) {
    const or_body *B = &S->bodies[i];
    if (B->hidden || ((S->shadow_skip >> i) & 1u)) continue;
    const double ca[3] = {B->bc_a[0], B->bc_a[1], B->bc_a[2]}, cb[3] = {B->bc_b[0], B->bc_b[1], B->bc_b[2]};
    if (seg_near(s->p, d, ca, cb, (double)B->bc_a[3] + S->emit_c[3] + 5e-4)) return -1;
  }
  return a;
}

/* renormalised barycentric weights over the solved nodes, then state word k interpolated (nodeWeights,
 * nodeEfree, nodeLout: shade.wgsl.js:211-231) */
static void node_interp(const or_scene *S, const surf_t *s, int k, double out[3]) {
  double w[3], sw = 0;
  for (int j = 0; j < 3; j++) { w[j] = s->b[j] * S->node_state[16 * (size_t)s->vi[j] + 4 * S->slot + 3]; sw += w[j]; }
  for (int j = 0; j < 3; j++) w[j] = sw <= 1e-6 ? s->b[j] : w[j] / sw;
  for (int c = 0; c < 3; c++) {
    out[c] = 0;
    for (int j = 0; j < 3; j++) out[c] += w[j] * S->node_state[16 * (size_t)s->vi[j] + 4 * k + c];
  }
}

static int is_static(const or_hit *h, const surf_t *s) { return (s->flags & F_STATIC) && h->inst < 0; }

/* one emitter sample's unshadowed contribution c and its shadow segment so + t wi, t < tmax; 0 when the sample
 * cannot light the point (lightEval, shade.wgsl.js:135-157) */
static int light_eval(const or_scene *S, const surf_t *s, const or_mat *M, const double wo[3], const double u[3],
                      double time, int mis, double nl, double c[3], double so[3], double wi[3], double *tmax) {
  double pl[3], en[3];
  const float *e = sample_emitter(S, u, time, pl, en);
  for (int k = 0; k < 3; k++) wi[k] = pl[k] - s->p[k];
  const double d2 = dot3(wi, wi), dist = sqrt(d2);
  for (int k = 0; k < 3; k++) wi[k] /= dist;
  const double cosE = -dot3(wi, en);
  if (cosE <= 0) return 0;
  const double cosS = dot3(s->ns, wi);
  if (cosS <= 0 || dot3(s->ng, wi) <= 0) return 0;
  uint32_t lid;
  memcpy(&lid, e + 11, 4);
  const double mwi[3] = {-wi[0], -wi[1], -wi[2]};
  const double pe = light_profile_at(S, lid, mwi, time);
  const float *L0 = S->lights + 4 * (size_t)lid;
  const double Le[3] = {L0[0] * pe, L0[1] * pe, L0[2] * pe};
  if (fmax(Le[0], fmax(Le[1], Le[2])) <= 0) return 0;
  double f[3];
  or_eval_bsdf(M, s->ns, wi, wo, f);
  if (maxc(f) <= 0) return 0;
  const double pW = d2 / (cosE * S->emit_area);
  double w = 1;
  if (mis) {
    const double pb = or_pdf_bsdf(M, s->ns, wo, wi);
    w = nl * nl * pW * pW / (nl * nl * pW * pW + pb * pb);
  }
  for (int k = 0; k < 3; k++) c[k] = f[k] * Le[k] * cosS * w / pW;
  offset_ray(s->p, s->ng, wi, so);
  *tmax = dist - 2e-5;
  return 1;
}

/* whether a visible body (not in `skip`) blocks the segment o + t d, t in (0, tmax), at `time` */
static int bodies_occlude(const or_scene *S, const double o[3], const double d[3], double tmax, double time,
                          uint32_t skip) {
  for (uint32_t b = 0; b < S->body_count; b++) {
    const or_body *B = &S->bodies[b];
    if (B->hidden || ((skip >> b) & 1u)) continue;
    pose_t Q;
    body_pose(B, time, &Q);
    double ol[3], dl[3], rel[3] = {o[0] - Q.p[0], o[1] - Q.p[1], o[2] - Q.p[2]};
    rot_apply_t(Q.R, rel, ol);
    rot_apply_t(Q.R, d, dl);
    or_hit hb;
    or_traverse(B->nodes, B->tris, ol, dl, 0, tmax, 1, 0, &hb);
    if (hb.tri >= 0) return 1;
  }
  return 0;
}

/* the nL light samples of a static surface under a still light (shade.wgsl.js:272-301): a lit sample counts when
 * neither the static scene nor a visible body blocks its segment; the key gains 0x20000 for a body's shadow, a
 * penumbra (some lit samples hidden, some not) or visible samples whose luminances differ by more than 20 % of
 * their midpoint. The renderer keeps the static answers per sub-sample; tracing them every time gives the same
 * image. Returns the sum of the visible samples. */
static void still_light(const or_scene *S, const surf_t *s, const or_mat *M, const double wo[3], double time,
                        const pix_t *P, uint32_t i, uint32_t nl, uint64_t *rays, uint32_t *key, double out[3]) {
  out[0] = out[1] = out[2] = 0;
  uint32_t nvis = 0, nlit = 0;
  double lmin = INFINITY, lmax = 0;
  for (uint32_t j = 0; j < nl; j++) {
    const uint32_t li = i * nl + j;
    const double u[3] = {or_kr(li, KL0, P->s[3]), or_kr(li, KL1, P->s[4]), or_kr(li, KL2, P->s[5])};
    double c[3], so[3], wi[3], tmax;
    if (!light_eval(S, s, M, wo, u, time, 0, nl, c, so, wi, &tmax)) continue;
    nlit++;
    (*rays)++;
    or_hit sh;
    or_traverse(S->nodes, S->tris, so, wi, 0, tmax, 1, 0, &sh);
    if (sh.tri >= 0) continue;
    if (bodies_occlude(S, so, wi, tmax, time, S->shadow_skip)) {
      *key |= 0x20000u;
      continue;
    }
    nvis++;
    for (int k = 0; k < 3; k++) out[k] += c[k];
    const double y = 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2];
    lmin = fmin(lmin, y);
    lmax = fmax(lmax, y);
  }
  if (nvis > 0 && nvis < nlit) *key |= 0x20000u;
  if (nvis > 1 && lmax - lmin > 0.2 * (lmax + lmin) * 0.5) *key |= 0x20000u;
}

static void direct_light(const or_scene *S, const surf_t *s, const or_mat *M, const double wo[3], double time,
                         const double u[3], int mis, double nl, double cached, uint64_t *rays, uint32_t *key,
                         double out[3]) {
  out[0] = out[1] = out[2] = 0;
  double pl[3], en[3];
  const float *e = sample_emitter(S, u, time, pl, en);
  double wi[3] = {pl[0] - s->p[0], pl[1] - s->p[1], pl[2] - s->p[2]};
  const double d2 = dot3(wi, wi), dist = sqrt(d2);
  for (int k = 0; k < 3; k++) wi[k] /= dist;
  const double cosE = -dot3(wi, en);
  if (cosE <= 0) return;
  const double cosS = dot3(s->ns, wi);
  if (cosS <= 0 || dot3(s->ng, wi) <= 0) return;
  uint32_t lid;
  memcpy(&lid, e + 11, 4);
  const double mwi[3] = {-wi[0], -wi[1], -wi[2]};
  const double pe = light_profile_at(S, lid, mwi, time);
  const float *L0 = S->lights + 4 * (size_t)lid;
  const double Le[3] = {L0[0] * pe, L0[1] * pe, L0[2] * pe};
  if (fmax(Le[0], fmax(Le[1], Le[2])) <= 0) return;
  double f[3];
  or_eval_bsdf(M, s->ns, wi, wo, f);
  if (maxc(f) <= 0) return;
  if (cached == 0) return;
  if (cached < 0) {
    double so[3];
    offset_ray(s->p, s->ng, wi, so);
    or_hit sh;
    *key |= 0x20000u;           /* a traced shadow ray marks the pixel for refinement (shade.wgsl.js:176-178) */
    trace_scene(S, so, wi, 0, dist - 2e-5, 1, time, S->shadow_skip, &sh, NULL);
    (*rays)++;
    if (sh.tri >= 0) return;
  }
  const double pW = d2 / (cosE * S->emit_area);
  double w = 1;
  if (mis) {
    const double pb = or_pdf_bsdf(M, s->ns, wo, wi);
    w = nl * nl * pW * pW / (nl * nl * pW * pW + pb * pb);
  }
  for (int k = 0; k < 3; k++) out[k] = f[k] * Le[k] * cosS * w / pW;
}

static void secondary(const or_scene *S, const double o[3], const double l[3], double time, double pdfB,
                      uint32_t opts, double nl, uint64_t *rays, double out[3]) {
  out[0] = out[1] = out[2] = 0;
  or_hit h;
  pose_t pose;
  trace_scene(S, o, l, 0, INFINITY, 0, time, 0u, &h, &pose);
  (*rays)++;
  if (h.tri < 0) return;
  surf_t s;
  surface(S, &h, &pose, o, l, &s);
  if (s.flags & F_EMITTER) {
    if (opts & O_NO_MIS) return;
    double ne[3];
    emitter_normal(S, &s, ne);
    const double ml[3] = {-l[0], -l[1], -l[2]};
    const double cosE = dot3(ml, ne);
    if (cosE <= 0) return;
    const double pe = light_profile_at(S, s.light, ml, time);
    const double pL = h.t * h.t / (cosE * S->emit_area);
    const double w = pdfB * pdfB / (pdfB * pdfB + nl * nl * pL * pL);
    for (int k = 0; k < 3; k++) out[k] = S->lights[4 * (size_t)s.light + k] * pe * w;
    return;
  }
  if (opts & (O_DIRECT_ONLY | O_NO_INDIRECT)) return;
  if (is_static(&h, &s)) {
    const double ml[3] = {-l[0], -l[1], -l[2]};
    double mu = dot3(s.ns, ml);
    mu = mu < 0 ? 0 : mu > 1 ? 1 : mu;
    double A[3], Ls[3];
    albedo_at(S, s.mat, mu, A);
    node_interp(S, &s, (int)S->slot, Ls);
    for (int k = 0; k < 3; k++) out[k] = A[k] * Ls[k];
    return;
  }
  for (int k = 0; k < 3; k++) out[k] = S->hh[4 * (size_t)s.mat + k] * S->e_avg[k] / M_PI;
}

/* key: the first hit's mesh (0xfffe on a miss) | 0x10000 on a posed body | 0x20000 once a shadow ray is traced
 * (gKey, shade.wgsl.js:10, 251-253); nl_in: the light samples of this pass */
static void ref_radiance(const or_scene *S, const or_camera *C, const pix_t *P, uint32_t x, uint32_t y, uint32_t i,
                         uint32_t nl_in, double out[3], uint64_t *rays, uint32_t *key) {
  out[0] = out[1] = out[2] = 0;
  *key = 0xfffeu;
  const float tk = or_kr(i, GOLD_F, P->s[2]);
  const double time = (double)(C->shutter[0] + (C->shutter[1] - C->shutter[0]) * tk);
  double d[3], o[3] = {C->pos[0], C->pos[1], C->pos[2]};
  cam_dir(C, P, x, y, i, d);
  or_hit h;
  pose_t pose;
  trace_scene(S, o, d, 0, INFINITY, 0, time, C->skip_bodies, &h, &pose);
  (*rays)++;
  if (h.tri < 0) return;
  surf_t s;
  surface(S, &h, &pose, o, d, &s);
  *key = (s.mesh & 0xffffu) | ((h.inst >= 0 && !(s.flags & F_STATIC)) ? 0x10000u : 0u);
  const double wo[3] = {-d[0], -d[1], -d[2]};
  const uint32_t opts = C->flags;
  if (opts & O_NORMALS) {
    for (int k = 0; k < 3; k++) out[k] = s.ns[k] * 0.5 + 0.5;
    return;
  }
  if (s.flags & F_EMITTER) {
    double ne[3];
    emitter_normal(S, &s, ne);
    if (dot3(wo, ne) <= 0) return;
    const double pe = light_profile_at(S, s.light, wo, time);
    for (int k = 0; k < 3; k++) out[k] = S->lights[4 * (size_t)s.light + k] * pe;
    return;
  }
  or_mat M;
  hit_material(S, &s, S->materials + 16 * (size_t)s.mat, &M);
  const int stat = is_static(&h, &s);
  const int mis = !stat && !(opts & O_NO_MIS);
  const uint32_t nl = nl_in < 1 ? 1 : nl_in > 16 ? 16 : nl_in;
  const double cached = (stat && (opts & O_CULL_SHADOWS) && !(opts & O_SHADOW_CACHE)) ? cached_visibility(S, &s, time) : -1;
  double L[3] = {0, 0, 0};
  if (S->emit_count > 0 && stat && (opts & O_SHADOW_CACHE))
    still_light(S, &s, &M, wo, time, P, i, nl, rays, key, L);
  else if (S->emit_count > 0)
    for (uint32_t j = 0; j < nl; j++) {
      const uint32_t li = i * nl + j;
      const double u[3] = {or_kr(li, KL0, P->s[3]), or_kr(li, KL1, P->s[4]), or_kr(li, KL2, P->s[5])};
      double c[3];
      direct_light(S, &s, &M, wo, time, u, mis, nl, cached, rays, key, c);
      for (int k = 0; k < 3; k++) L[k] += c[k];
    }
  for (int k = 0; k < 3; k++) L[k] /= nl;
  if (stat) {
    if (!(opts & (O_NO_INDIRECT | O_DIRECT_ONLY))) {
      double E[3], A[3];
      node_interp(S, &s, 0, E);
      const double V = (opts & O_NO_CAPSULE) ? 1.0 : or_capsule_vis(S, s.p, s.ns);
      double mu = dot3(s.ns, wo);
      if (mu < 0) mu = 0;
      albedo_at(S, s.mat, mu, A);
      for (int k = 0; k < 3; k++) L[k] += E[k] * V / M_PI * A[k];
    }
  } else {
    const double ub[3] = {or_kr(i, KB0, P->s[6]), or_kr(i, KB1, P->s[7]), or_kr(i, KB2, P->s[8])};
    double l[3];
    const double pdf = or_sample_bsdf(&M, s.ns, wo, ub, l);
    if (pdf > 0 && dot3(s.ng, l) > 0) {
      double f[3], so[3], sec[3];
      or_eval_bsdf(&M, s.ns, l, wo, f);
      const double c = dot3(s.ns, l) / pdf;
      offset_ray(s.p, s.ng, l, so);
      secondary(S, so, l, time, pdf, opts, nl, rays, sec);
      for (int k = 0; k < 3; k++) L[k] += f[k] * c * sec[k];
    }
  }
  memcpy(out, L, sizeof L);
}

/* ------------------------------------------------------------------ node radiosity cache */
static uint32_t gather_vertex(const or_scene *S, const float *npos, const float *nnorm, uint32_t i, uint32_t j,
                              uint32_t K, double bb[3], uint32_t vi[3], int *hit) {
  const double p[3] = {npos[4 * (size_t)i], npos[4 * (size_t)i + 1], npos[4 * (size_t)i + 2]};
  const float nf[3] = {nnorm[4 * (size_t)i], nnorm[4 * (size_t)i + 1], nnorm[4 * (size_t)i + 2]};
  float df[3];
  or_gather_dir(i, j, K, nf, df);
  const double dir[3] = {df[0], df[1], df[2]};
  const double o[3] = {p[0] + nf[0] * 2e-6, p[1] + nf[1] * 2e-6, p[2] + nf[2] * 2e-6};
  or_hit h;
  or_traverse(S->nodes, S->tris, o, dir, 0, INFINITY, 0, 0, &h);
  *hit = h.tri >= 0;
  if (h.tri < 0) return 0xffffffffu;
  const uint32_t *ru = (const uint32_t *)(S->tris + (size_t)h.tri * 16);
  vi[0] = ru[3]; vi[1] = ru[7]; vi[2] = ru[11];
  bb[0] = 1 - h.u - h.v; bb[1] = h.u; bb[2] = h.v;
  double nv = 0;
  for (int k = 0; k < 3; k++) {
    double c = 0;
    for (int q = 0; q < 3; q++) c += bb[q] * nnorm[4 * (size_t)vi[q] + k];
    nv += c * dir[k];
  }
  if (!(nv < 0)) return 0xffffffffu;
  if (bb[0] >= bb[1] && bb[0] >= bb[2]) return vi[0];
  return bb[1] >= bb[2] ? vi[1] : vi[2];
}

void or_gather_build(const or_scene *S, const float *npos, const float *nnorm, uint32_t V, uint32_t K,
                     uint32_t *gtable, int nthreads) {
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 64)
  for (int64_t r = 0; r < (int64_t)V * K; r++) {
    const uint32_t i = (uint32_t)(r / K), j = (uint32_t)(r % K);
    double bb[3];
    uint32_t vi[3];
    int hit;
    gtable[r] = npos[4 * (size_t)i + 3] > 0.5f ? gather_vertex(S, npos, nnorm, i, j, K, bb, vi, &hit) : 0xffffffffu;
  }
}

void or_gather_remap(const or_scene *S, const float *npos, const float *nnorm, uint32_t V, uint32_t K,
                     const float *miss, uint32_t *gtable, int nthreads) {
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 64)
  for (int64_t r = 0; r < (int64_t)V * K; r++) {
    const uint32_t v0 = gtable[r];
    if (v0 == 0xffffffffu || miss[v0] <= 0.5 * K) continue;
    const uint32_t i = (uint32_t)(r / K), j = (uint32_t)(r % K);
    double bb[3];
    uint32_t vi[3];
    int hit;
    gather_vertex(S, npos, nnorm, i, j, K, bb, vi, &hit);
    if (!hit) continue;
    uint32_t best = v0;
    double bw = -1;
    for (int q = 0; q < 3; q++)
      if (miss[vi[q]] <= 0.5 * K && bb[q] > bw) { best = vi[q]; bw = bb[q]; }
    gtable[r] = best;
  }
}

void or_node_direct(const or_scene *S, const float *npos, const float *nnorm, const uint32_t *nmat,
                    const uint32_t *active, uint32_t n_active, uint32_t samples, const float *alb, float *state,
                    int nthreads) {
  or_scene S2 = *S;
  S2.alb = alb;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 16)
  for (int64_t k = 0; k < (int64_t)n_active; k++) {
    const uint32_t i = active[k];
    const double p[3] = {npos[4 * (size_t)i], npos[4 * (size_t)i + 1], npos[4 * (size_t)i + 2]};
    const double n[3] = {nnorm[4 * (size_t)i], nnorm[4 * (size_t)i + 1], nnorm[4 * (size_t)i + 2]};
    const uint32_t m = nmat[i];
    const uint32_t hs = or_pcg(i * 2246822519u + 3u);
    const float s0 = or_u01(or_pcg(hs)), s1 = or_u01(or_pcg(hs + 1u)), s2 = or_u01(or_pcg(hs + 2u));
    double B[3] = {0, 0, 0};
    uint32_t nlit = 0, nvis = 0;
    const double o[3] = {p[0] + n[0] * 2e-6, p[1] + n[1] * 2e-6, p[2] + n[2] * 2e-6};
    for (uint32_t s = 0; s < samples && S->emit_count > 0; s++) {
      const double u[3] = {or_kr(s, KL0, s0), or_kr(s, KL1, s1), or_kr(s, KL2, s2)};
      double pl[3], en[3];
      const float *e = sample_emitter(S, u, 0.5, pl, en);
      double wi[3] = {pl[0] - p[0], pl[1] - p[1], pl[2] - p[2]};
      const double d2 = dot3(wi, wi), dist = sqrt(d2);
      for (int q = 0; q < 3; q++) wi[q] /= dist;
          const double cosE = -dot3(wi, en), cosN = dot3(n, wi);
      if (cosE <= 0 || cosN <= 0) continue;
      uint32_t lid;
      memcpy(&lid, e + 11, 4);
      const double mwi[3] = {-wi[0], -wi[1], -wi[2]};
      const double pe = light_profile_at(S, lid, mwi, 0.5);
      const float *L0 = S->lights + 4 * (size_t)lid;
      if (fmax(L0[0], fmax(L0[1], L0[2])) * pe <= 0) continue;
      nlit++;
      or_hit sh;
      trace_scene(S, o, wi, 0, dist - 2e-5, 1, 0.5, S->shadow_skip, &sh, NULL);
      if (sh.tri >= 0) continue;
      nvis++;
      double A[3];
      albedo_at(&S2, m, cosN, A);
      for (int q = 0; q < 3; q++) B[q] += L0[q] * pe * (cosE * cosN * S->emit_area / d2) * A[q];
    }
    const double inv = 1.0 / (samples ? samples : 1);
    float *st = state + 16 * (size_t)i;
    for (int q = 0; q < 3; q++) st[12 + q] = (float)(B[q] * inv);
    st[15] = nlit == 0 ? -1.0f : nvis == nlit ? 1.0f : nvis == 0 ? 0.0f : 0.5f;
    st[3] = (float)or_capsule_vis(S, p, n);
  }
}

/* one pixel's first pass: the sum over sub-samples [s0, s1) and, for the reference estimator, its key {mesh of
 * sub-sample 0 (0xfffe: a miss), a mesh change among the sub-samples (0x40000), body / traced-shadow bits} (camMain,
 * kernels.wgsl.js:172-213) */
static void first_pass(const or_scene *S, const or_camera *C, uint32_t x, uint32_t y, uint32_t s0, uint32_t s1,
                       double acc[3], uint32_t *key, uint64_t *r, uint64_t vv[5]) {
  pix_t P;
  pix_seeds(C, x, y, &P);
  acc[0] = acc[1] = acc[2] = 0;
  *key = 0;
  for (uint32_t i = s0; i < s1; i++) {
    double L[3];
    if ((C->flags & O_REFERENCE) && S->node_state) {
      uint32_t k;
      ref_radiance(S, C, &P, x, y, i, C->light_samples, L, r, &k);
      if (i == s0) *key = k;
      else if ((k & 0xffffu) != (*key & 0xffffu)) *key |= 0x40000u;
      *key |= k & 0x30000u;
    } else {
      sample_radiance(S, C, &P, x, y, i, L, r, vv, NULL);
    }
    for (int k = 0; k < 3; k++) acc[k] += L[k];
  }
}

void or_render_tile(const or_scene *S, const or_camera *C, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                    uint32_t spp_begin, uint32_t spp_end, float *out, uint64_t *rays, uint64_t *visits, int nthreads) {
  uint64_t nr = 0, v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0;
  const int refine = (C->flags & O_REFERENCE) && S->node_state && C->refine > 0 && spp_begin == 0;
  /* with the refinement pass, the first pass covers the tile and a 1-pixel rim (the 3x3 key test) */
  const uint32_t rx0 = refine && x0 > 0 ? x0 - 1 : x0, ry0 = refine && y0 > 0 ? y0 - 1 : y0;
  const uint32_t rx1 = refine ? (x0 + w + 1 < C->width ? x0 + w + 1 : C->width) : x0 + w;
  const uint32_t ry1 = refine ? (y0 + h + 1 < C->height ? y0 + h + 1 : C->height) : y0 + h;
  const uint32_t rw = rx1 - rx0, rh = ry1 - ry0;
  double *sum = malloc(sizeof(double) * 3 * (size_t)rw * rh);
  uint32_t *keys = malloc(sizeof(uint32_t) * (size_t)rw * rh);
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 1) reduction(+ : nr, v0, v1, v2, v3, v4) collapse(2)
  for (uint32_t yy = 0; yy < rh; yy++)
    for (uint32_t xx = 0; xx < rw; xx++) {
      uint64_t r = 0, vv[5] = {0, 0, 0, 0, 0};
      const size_t q = (size_t)yy * rw + xx;
      first_pass(S, C, rx0 + xx, ry0 + yy, spp_begin, spp_end, sum + 3 * q, keys + q, &r, vv);
      nr += r; v0 += vv[0]; v1 += vv[1]; v2 += vv[2]; v3 += vv[3]; v4 += vv[4];
    }
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 1) reduction(+ : nr) collapse(2)
  for (uint32_t yy = 0; yy < h; yy++)
    for (uint32_t xx = 0; xx < w; xx++) {
      const uint32_t x = x0 + xx, y = y0 + yy;
      const size_t q = (size_t)(y - ry0) * rw + (x - rx0);
      double acc[3] = {sum[3 * q], sum[3 * q + 1], sum[3 * q + 2]};
      uint32_t n = spp_end - spp_begin;
      if (refine) {
        /* camRefine (kernels.wgsl.js:218-252): an edge among the sub-samples or in the 3x3 block, a posed body,
         * or (option) traced shadow rays: `refine` more sub-samples, refine_light_samples light samples each */
        const uint32_t k = keys[q];
        int go = (k & 0x50000u) != 0 || ((k & 0x20000u) && (C->flags & O_REFINE_SHADOW));
        for (int dy = -1; dy <= 1 && !go; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            int xn = (int)x + dx, yn = (int)y + dy;
            xn = xn < 0 ? 0 : xn > (int)C->width - 1 ? (int)C->width - 1 : xn;
            yn = yn < 0 ? 0 : yn > (int)C->height - 1 ? (int)C->height - 1 : yn;
            if ((keys[(size_t)(yn - (int)ry0) * rw + (size_t)(xn - (int)rx0)] & 0xffffu) != (k & 0xffffu)) go = 1;
          }
        if (go) {
          pix_t P;
          pix_seeds(C, x, y, &P);
          uint64_t r = 0;
          for (uint32_t i = spp_end; i < spp_end + C->refine; i++) {
            double L[3];
            uint32_t kk;
            ref_radiance(S, C, &P, x, y, i, C->refine_light_samples ? C->refine_light_samples : C->light_samples, L,
                         &r, &kk);
            for (int c = 0; c < 3; c++) acc[c] += L[c];
          }
          nr += r;
          n += C->refine;
        }
      }
      const double inv = n ? 1.0 / n : 0;
      float *o = out + 4 * ((size_t)yy * w + xx);
      o[0] = (float)(acc[0] * inv); o[1] = (float)(acc[1] * inv); o[2] = (float)(acc[2] * inv); o[3] = 1;
    }
  free(sum);
  free(keys);
  if (rays) *rays = nr;
  if (visits) { visits[0] = v0; visits[1] = v1; visits[2] = v2; visits[3] = v3; visits[4] = v4; }
}

void or_backward_tile(const or_scene *S, const or_camera *C, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                      uint32_t spp_begin, uint32_t spp_end, const float *adj, double *grad_mat, double *grad_light,
                      int nthreads) {
  const size_t nm = 3 * (size_t)S->mat_count, nli = 3 * (size_t)S->light_count;
  const double sc = 1.0 / (double)(spp_end - spp_begin);
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    double *gm = calloc(nm, sizeof(double)), *gl = calloc(nli > 0 ? nli : 1, sizeof(double));
#pragma omp for schedule(dynamic, 1) collapse(2)
    for (uint32_t yy = 0; yy < h; yy++)
      for (uint32_t xx = 0; xx < w; xx++) {
        const float *a = adj + 4 * ((size_t)yy * w + xx);
        if (a[0] == 0 && a[1] == 0 && a[2] == 0) continue;
        pix_t P;
        pix_seeds(C, x0 + xx, y0 + yy, &P);
        for (uint32_t i = spp_begin; i < spp_end; i++) {
          double L[3];
          uint64_t r = 0, vv[5] = {0, 0, 0, 0, 0};
          sample_radiance(S, C, &P, x0 + xx, y0 + yy, i, L, &r, vv, NULL);
          grad_t G = {gm, gl, {a[0] * sc, a[1] * sc, a[2] * sc}, {L[0], L[1], L[2]}};
          sample_radiance(S, C, &P, x0 + xx, y0 + yy, i, L, &r, vv, &G);
        }
      }
#pragma omp critical
    {
      for (size_t k = 0; k < nm; k++) grad_mat[k] += gm[k];
      for (size_t k = 0; k < nli; k++) grad_light[k] += gl[k];
    }
    free(gm);
    free(gl);
  }
}

/* ---------------------------------------------------------------- camera-parameter gradients (interior term)
 * theta = (fx, fy, cx, cy, k1..k4, r, t): intrinsics of the KB4 (or pinhole) camera, a rotation r about the
 * world axes applied to R (R' = exp([r]x) R) and the camera centre t. For every sample, central differences
 * in FP64 of its radiance with the image point, the random numbers and the triangles held fixed: the ray of
 * theta +- h e_k is built in FP64 (gpu/lens.js:54-62 unproject) and the recorded path replayed on it: every hit
 * re-derived on its triangle's plane, shadow rays keeping their outcomes, and each continuation along the recorded
 * direction with the recorded pdf (sampling detached; resampling would turn the directions with the camera and
 * feed the change back through every later bounce off curved, interpolated-normal surfaces). A pair whose branches
 * change (a light sample or the shading-normal fallback) is retried with h shrunk up to three times; one still
 * changing is counted in *changed and left out (the boundary term, which is excluded). Gradient of
 * sum_p adj_p . I_p (I = mean over the sample range). */
static void rodrigues(const double r[3], double M[9]) {
  double th = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  double kx = 0, ky = 0, kz = 0, c = 1, sn = 0;
  if (th > 0) { kx = r[0] / th; ky = r[1] / th; kz = r[2] / th; c = cos(th); sn = sin(th); }
  double v = 1 - c;
  /* row-major M: M[3*i + j] */
  M[0] = c + kx * kx * v;      M[1] = kx * ky * v - kz * sn; M[2] = kx * kz * v + ky * sn;
  M[3] = ky * kx * v + kz * sn; M[4] = c + ky * ky * v;      M[5] = ky * kz * v - kx * sn;
  M[6] = kz * kx * v - ky * sn; M[7] = kz * ky * v + kx * sn; M[8] = c + kz * kz * v;
}

static void ray_of_theta(const or_camera *C, const double th[14], double u, double v, double ray[6]) {
  double dc[3];
  if (C->mode == 0) {
    or_lens L = {th[0], th[1], th[2], th[3], C->lens[8], {th[4], th[5], th[6], th[7]}};
    double duv[2] = {u, v};
    or_unproject(&L, duv, dc);
  } else {
    dc[0] = (u - th[2]) / th[0];
    dc[1] = (v - th[3]) / th[1];
    dc[2] = 1;
    norm3(dc);
  }
  double d0[3];
  for (int k = 0; k < 3; k++) d0[k] = (double)C->R[k] * dc[0] + (double)C->R[3 + k] * dc[1] + (double)C->R[6 + k] * dc[2];
  double M[9];
  rodrigues(th + 8, M);
  for (int k = 0; k < 3; k++) ray[3 + k] = M[3 * k] * d0[0] + M[3 * k + 1] * d0[1] + M[3 * k + 2] * d0[2];
  norm3(ray + 3);
  for (int k = 0; k < 3; k++) ray[k] = th[11 + k];
}

void or_camera_grad_tile(const or_scene *S, const or_camera *C, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                         uint32_t spp_begin, uint32_t spp_end, const float *adj, double rel_step, double *grad,
                         uint64_t *changed, int nthreads) {
  const double sc = 1.0 / (double)(spp_end - spp_begin);
  double th0[14];
  for (int k = 0; k < 8; k++) th0[k] = C->lens[k];
  th0[8] = th0[9] = th0[10] = 0;
  for (int k = 0; k < 3; k++) th0[11 + k] = C->pos[k];
  double step[14];
  for (int k = 0; k < 14; k++) step[k] = rel_step;
  for (int k = 0; k < 4; k++) step[k] = rel_step * fabs(th0[0]);
  uint64_t nch = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : nch)
  {
    double g[14] = {0};
#pragma omp for schedule(dynamic, 1) collapse(2)
    for (uint32_t yy = 0; yy < h; yy++)
      for (uint32_t xx = 0; xx < w; xx++) {
        const float *a = adj + 4 * ((size_t)yy * w + xx);
        if (a[0] == 0 && a[1] == 0 && a[2] == 0) continue;
        const uint32_t x = x0 + xx, y = y0 + yy;
        pix_t P;
        pix_seeds(C, x, y, &P);
        for (uint32_t i = spp_begin; i < spp_end; i++) {
          float uvf[2];
          {
            volatile float ua = (float)x + or_kr(i, R2A, P.s[0]);
            volatile float va = (float)y + or_kr(i, R2B, P.s[1]);
            uvf[0] = ua - 0.5f; uvf[1] = va - 0.5f;
          }
          double ray[6], L0[3];
          uint64_t r = 0, vv[5] = {0, 0, 0, 0, 0}, seq0 = 0;
          path_rec_t rec;
          memset(&rec, 0, sizeof rec);
          rec.mode = 1; rec.min_cos = 1.0;
          ray_of_theta(C, th0, uvf[0], uvf[1], ray);
          sample_radiance_ex(S, C, &P, x, y, i, L0, &r, vv, NULL, ray, &seq0, &rec);
          const int want_n = rec.n;
          for (int k = 0; k < 14; k++) {
            double hk = step[k];
            int ok = 0;
            double Lp[3], Lm[3];
            for (int tries = 0; tries < 4 && !ok; tries++, hk /= 16) {
              double th[14];
              memcpy(th, th0, sizeof th);
              uint64_t sp = 0, sm = 0;
              path_rec_t rp = rec, rm = rec;
              rp.mode = rm.mode = 2;
              rp.n = rm.n = 0;
              rp.bad = rm.bad = 0;
              th[k] = th0[k] + hk;
              ray_of_theta(C, th, uvf[0], uvf[1], ray);
              sample_radiance_ex(S, C, &P, x, y, i, Lp, &r, vv, NULL, ray, &sp, &rp);
              th[k] = th0[k] - hk;
              ray_of_theta(C, th, uvf[0], uvf[1], ray);
              sample_radiance_ex(S, C, &P, x, y, i, Lm, &r, vv, NULL, ray, &sm, &rm);
              /* the replays hold the recorded triangles and shadow outcomes; a pair whose branches changed is retried
                 with a smaller step */
              ok = !rec.bad && !rp.bad && !rm.bad && rp.n == want_n && rm.n == want_n;
              if (ok)
                for (int c = 0; c < 3; c++) g[k] += a[c] * sc * (Lp[c] - Lm[c]) / (2 * hk);
            }
            if (!ok) nch++;
          }
        }
      }
#pragma omp critical
    for (int k = 0; k < 14; k++) grad[k] += g[k];
  }
  if (changed) *changed = nch;
}

void or_visible(const or_scene *S, const or_camera *C, const float *pts, uint32_t n, double tol, double time,
                uint8_t *vis) {
  const double c[3] = {C->pos[0], C->pos[1], C->pos[2]};
  for (uint32_t k = 0; k < n; k++) {
    double v[3] = {pts[3 * (size_t)k] - c[0], pts[3 * (size_t)k + 1] - c[1], pts[3 * (size_t)k + 2] - c[2]};
    const double dist = sqrt(dot3(v, v));
    vis[k] = 1;
    if (!(dist > tol)) continue;
    for (int j = 0; j < 3; j++) v[j] /= dist;
    or_hit h;
    trace_scene(S, c, v, 0, dist - tol, 1, time, C->skip_bodies, &h, NULL);
    vis[k] = h.tri < 0;
  }
}
