/* CPU oracle for the camera-rendering path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU path, used by tests/, the smoke check and bench.py's
 * cpu_baseline leg as the checker. Nothing in the product library links or calls this code.
 *
 * Geometry and shading arithmetic is FP64 on FP32 scene data, as the reference's JS (JS numbers are
 * doubles; scene arrays are Float32Array). The sample pattern and the sensor model are evaluated in
 * FP32 exactly as the reference's WGSL states them, so the GPU path can be compared bit for bit.
 *
 * Every function cites the reference file:line it follows. Parity pins: see the header of each .c file.
 */
#ifndef VT_ORACLE_H
#define VT_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- rng (control/rng.js) */
uint32_t or_mulberry32_u32(uint32_t *state);                 /* control/rng.js:6-11, the 32-bit word */
double or_mulberry32_uniform(uint32_t *state);               /* control/rng.js:11, word / 2^32 */
void or_mulberry32_fill(uint32_t seed, double *out, size_t n);
typedef struct { uint32_t a; int has_spare; double spare; } or_rng;
void or_rng_init(or_rng *r, uint32_t seed);                  /* control/rng.js:16 */
double or_rng_uniform(or_rng *r, double lo, double hi);      /* control/rng.js:17 */
int64_t or_rng_int(or_rng *r, int64_t lo, int64_t hi);       /* control/rng.js:18 */
double or_rng_normal(or_rng *r, double mu, double sigma);    /* control/rng.js:19-27 (host libm) */

/* hashes of gpu/wgsl/common.wgsl.js:94-113 (u32 exact, u01 exact) */
uint32_t or_pcg(uint32_t v);
uint32_t or_hash3(uint32_t a, uint32_t b, uint32_t c);
float or_u01(uint32_t x);
float or_kr(uint32_t i, float a, float shift);

/* ---------------------------------------------------------------- lens (gpu/lens.js) */
typedef struct { double fx, fy, cx, cy, alpha, k[4]; } or_lens;
double or_distort(const or_lens *L, double th);                      /* gpu/lens.js:19-22 */
double or_distort_deriv(const or_lens *L, double th);                /* gpu/lens.js:25-28 */
double or_undistort_theta(const or_lens *L, double thd, int iters);  /* gpu/lens.js:31-40 */
void or_project(const or_lens *L, const double X[3], double uv[2]);  /* gpu/lens.js:43-51 */
void or_unproject(const or_lens *L, const double uv[2], double ray[3]); /* gpu/lens.js:54-62 */
double or_max_monotone_theta(const or_lens *L, double limit);        /* gpu/lens.js:65-72 */
double or_max_image_theta(const or_lens *L, int w, int h);           /* gpu/lens.js:75-83 */
/* FP32 twin of gpu/wgsl/lens.wgsl.js:9-29 (fixed 8 Newton steps), evaluated in binary32 */
void or_unproject_f32(const float lens[9], const float uv[2], float ray[3]);

/* ---------------------------------------------------------------- BVH (gpu/bvh.js) */
#define OR_NODE_WORDS 16
typedef struct { uint32_t node_count; int max_depth; float bmin[3], bmax[3]; } or_bvh_info;
/* gpu/bvh.js:18-161. nodes_out holds cap_nodes*16 words; returns 0, or -1 if cap_nodes is too small. */
int or_build_bvh(const float *pos, const uint32_t *idx, uint32_t ntri, int leaf_size, int bins,
                 float *nodes_out, uint32_t cap_nodes, uint32_t *order_out, or_bvh_info *info);
/* gpu/bvh.js:175-186: 16 floats per BVH slot (v0|e1|e2 geometry; w words untouched) */
void or_pack_triangles(const float *pos, const uint32_t *idx, const uint32_t *order, uint32_t n, float *out);
/* gpu/bvh.js:189-206; returns t or +inf, writes u, v */
double or_intersect_tri(const float *T, uint32_t k, const double o[3], const double d[3], double tmin,
                        double tmax, double *u, double *v);
typedef struct { double t, u, v; int64_t tri; uint64_t node_visits, tri_tests; int32_t inst; } or_hit;
/* gpu/bvh.js:224-250 (closest or any hit), counting node visits and triangle tests */
void or_traverse(const float *nodes, const float *T, const double o[3], const double d[3], double tmin,
                 double tmax, int any_hit, uint32_t root, or_hit *out);
void or_brute_force(const float *T, uint32_t count, const double o[3], const double d[3], double tmin,
                    double tmax, or_hit *out);                          /* gpu/bvh.js:253-261 */
double or_sah_cost(const float *nodes, uint32_t node_count, const float bmin[3], const float bmax[3],
                   double trav_cost, double tri_cost);                  /* gpu/bvh.js:264-279 */
/* batch helpers for tests: rays as 8 floats (o.xyz, tmin, d.xyz, tmax); hits as (t, tri, u, v) doubles */
void or_trace_batch(const float *nodes, const float *T, const float *rays, uint32_t n, int any_hit,
                    double *hits, uint64_t *visits, int nthreads);
void or_brute_batch(const float *T, uint32_t count, const float *rays, uint32_t n, double *hits, int nthreads);

/* ---------------------------------------------------------------- BSDF (gpu/bsdf.js) */
enum { OR_DIELECTRIC = 0, OR_CONDUCTOR = 1, OR_COATED = 2, OR_LAMBERT = 3 };
typedef struct {
  int kind;
  double c[3];          /* rho_d (dielectric, Lambert) or F0 (conductor, coated) */
  double alpha, roughness, ior, fbar, fbar_computed;
  double alpha_c, tau[3];
} or_mat;
/* packed material record shared with the upload format: 16 floats
 * [c.rgb, alpha][ior, kind, fbar, roughness][tau.rgb, alpha_c][-] */
void or_mat_unpack(const float rec[16], or_mat *m);
double or_fresnel_dielectric(double cosi, double eta);        /* gpu/bsdf.js:25-32 */
double or_fresnel_average(double eta, int n);                 /* gpu/bsdf.js:35-42 */
double or_schlick(double f0, double c);                       /* gpu/bsdf.js:44 */
double or_ggx_d(double nh, double a);                         /* gpu/bsdf.js:46-50 */
double or_ggx_lambda(double c, double a);                     /* gpu/bsdf.js:52-56 */
double or_ggx_g1(double c, double a);                         /* gpu/bsdf.js:58 */
double or_ggx_g2(double nl, double nv, double a);             /* gpu/bsdf.js:59 */
void or_eval_bsdf(const or_mat *M, const double n[3], const double l[3], const double v[3], double f[3]); /* :98-117 */
double or_spec_prob(const or_mat *M, double nv);              /* gpu/bsdf.js:147-157 */
double or_pdf_bsdf(const or_mat *M, const double n[3], const double v[3], const double l[3]); /* :160-173 */
void or_basis(const double n[3], double t[3], double b[3]);   /* gpu/bsdf.js:122-126 */
void or_sample_vndf(const double ve[3], double a, double u1, double u2, double h[3]); /* :131-144 */
/* gpu/bsdf.js:176-198; returns pdf (0 = no sample), writes l */
double or_sample_bsdf(const or_mat *M, const double n[3], const double v[3], const double u[3], double l[3]);
double or_radinv2(uint32_t i);                                /* gpu/bsdf.js:203 */
/* gpu/bsdf.js:210-254: spec, diff [bins*3]; hh, hh_spec [3] */
void or_albedo_table(const or_mat *M, int bins, int samples, double *spec, double *diff, double hh[3], double hh_spec[3]);
void or_lookup_albedo(const double *spec, const double *diff, int bins, double mu, int which, double out[3]); /* :257-266 */
void or_layer_normal(const double n[3], const double p[3], const double bd[3], double period, double depth, double out[3]); /* :274-280 */

/* ---------------------------------------------------------------- shade_ref (gpu/shade_ref.js) */
double or_profile(double theta, const double *knots, const double *p, int n);          /* :5-14 */
double or_capsule_occ(const double p[3], const double n[3], const double a[3], const double b[3], double r); /* :21-38 */
void or_gauss_legendre(int n, double *x, double *w);                                     /* :64-78 */
double or_disc_irradiance(const double x[3], const double n[3], const double c[3], const double m[3],
                          double a, double L, int nr, int nphi);                         /* :44-62 */

/* ---------------------------------------------------------------- sensor (gpu/wgsl/sensor.wgsl.js, gpu/sensor.js) */
/* binary32 math written for bit-identical host and device results (published methods, see sensor.c) */
float or_expf(float x);
float or_logf(float x);
float or_cosf(float x);
float or_erfa(float z);                                       /* A&S 7.1.26, sensor.wgsl.js:57-61 */
typedef struct {
  uint32_t width, height, cfa, frame;          /* cfa: 0 RGGB, 1 GRBG, 2 GBRG, 3 BGGR */
  float optics_k, esf_sigma, exposure, gain;   /* pi T / (4 N^2), sigma (px), t (s), analogue gain */
  float m[9];                                  /* response matrix rows (e-/s per W m^-2) */
  float read_noise, full_well, dark, conv_gain;
  float max_dn, black, noise_on; uint32_t seed;
  uint32_t cam_index;
  float glare_k, glare_sigma; uint32_t glare_f; /* veiling glare: fraction (0 off), sigma (level px), level F */
} or_sensor;
/* the blurred veiling-glare level (sensor.wgsl.js:87-127): ceil(w/F) x ceil(h/F) x 3 floats */
void or_glare_level(const or_sensor *s, const float *radiance, const float *ri, float *level);

/* the ISP after the raw DN (sensor.wgsl.js:210-427), isp.c */
typedef struct {
  uint32_t w, h, cfa, demosaic, nr, tone, chroma420;   /* demosaic 0 bilinear 1 edge-aware 2 Hamilton-Adams;
                                                          tone 0 sRGB 1 gamma 2 linear 3 sRGB + S-curve */
  float max_dn, black, read_noise, conv_gain, gain;    /* from the raw stage */
  float wb[3], ccm[9];
  float nr_sigma, nr_scale, gamma, s_c, black_off, sharpen_amount;
  int sharpen_rad;                                     /* 1..4 */
} or_isp_p;
void or_isp_green(const or_isp_p *P, const uint16_t *dn, float *g);
void or_isp_rb(const or_isp_p *P, const uint16_t *dn, const float *g, float *rgb, uint32_t *stats);
void or_isp_nr(const or_isp_p *P, const float *src, float *dst);
float or_isp_tone(const or_isp_p *P, float x);
void or_isp_out(const or_isp_p *P, const float *src, uint8_t *rgba, uint8_t *yuv, uint32_t *stats);
void or_isp(const or_isp_p *P, const uint16_t *dn, uint8_t *rgba, uint8_t *yuv, uint32_t *stats, float *lin);
/* sensorOptics + sensorExpose (sensor.wgsl.js:68-84, 143-208) for one camera: radiance rgba f32 [w*h*4],
 * ri [w*h] -> dn [w*h] (u16). Glare off (fraction 0). */
void or_sensor_expose(const or_sensor *s, const float *radiance, const float *ri, uint16_t *dn);
/* gpu/sensor.js:103-124 for the default vignetting table [0, 90 deg] -> [1, 1] */
/* relative illumination x vignetting per pixel (gpu/sensor.js:103-124): vn knots vt_rad (ascending) / vv, linear
 * between them, the end values outside; vn 0 = no vignetting */
void or_ri_map(const or_lens *L, int w, int h, double pupil_exp, int vn, const double *vt_rad, const double *vv,
               float *out);
double or_raw_noise_dn(double e, double read_noise, double conv_gain, double gain); /* gpu/sensor.js:180-182 */

/* ---------------------------------------------------------------- path tracer (the build's estimator) */
typedef struct {
  const float *nodes; uint32_t node_count;     /* BVH2, gpu/bvh.js layout */
  const float *tris; uint32_t tri_count;       /* 16 words per slot: v0|i0, e1|i1, e2|i2, mesh, flags */
  const float *verts; uint32_t vert_count;     /* 8 words per vertex: pos|mesh, normal|area */
  const uint32_t *meshes; uint32_t mesh_count; /* 4 words per mesh: material, flags, light, - */
  const float *materials; uint32_t mat_count;  /* 16 floats per material (or_mat_unpack) */
  const float *lights; uint32_t light_count;   /* 4 floats per light: radiance rgb, - */
  const float *emit; uint32_t emit_count;      /* 16 floats per emitter triangle: v0|cdf, e1|area, e2|light, n|- */
  float emit_area;
  const struct or_body *bodies; uint32_t body_count;  /* posed bodies (may be 0) */
  uint32_t shadow_skip;                                /* bit i: shadow rays ignore body i */
  /* angular emission profiles (optional, NULL = every light Lambertian): per light 4 words (unit axis xyz, knot
   * count as uint32); knots: 64 (theta, p) pairs per light. L_e(w) = radiance * or_profile(theta(w, axis)). */
  const float *light_prof; const float *knots;
  /* node radiosity cache for the reference estimator (NULL: none): 16 floats per vertex (E_free + V, slot-1 and
   * slot-2 sources, B_dir + visibility code); albedo tables, 32 bins x 8 floats [spec.rgb, -, diff.rgb, -] per
   * material; hemispherical albedo, 4 floats per material; the current source slot; eAvg; the emitters'
   * centroid + radius (shadow culling) */
  const float *node_state; const float *alb; const float *hh;
  uint32_t slot; float e_avg[3]; float emit_c[4];
  /* material table of the path tracer's sampling (NULL: `materials`): the lobe choice, the BSDF pdf of a
   * continuation and the MIS weights read it, the BSDF values read `materials`. A finite difference of the
   * colours in `materials` with this table held at the unperturbed colours is the derivative the adjoint
   * computes (sampling detached). */
  const float *sample_materials;
  /* emit_posed 1: the light rides on body emit_body; the emit records, the light axes and emit_c are in its frame
   * and every emitter sample is posed with it at its exposure time (shade.wgsl.js:31-47, 55) */
  uint32_t emit_posed, emit_body;
  /* the printed board (gpu/wgsl/shade.wgsl.js:258-265): texture coordinates, 2 floats per vertex (NULL: none), and
   * the mask, tex_w x tex_h bytes, 255 = paper (NULL: one paper texel). A mesh flagged 128 (TEXTURED) mixes rho_d
   * and alpha between the ink (material words 8-11) and the paper by the mask at the hit. */
  const float *uvs; const uint8_t *tex; uint32_t tex_w, tex_h;
} or_scene;
/* the print mask at texture coordinates (u, v), bilinear with clamp-to-edge addressing, texel (i, j) centred at
 * ((i + 0.5) / w, (j + 0.5) / h) (a linear sampler, gpu/renderer.js:398), in binary32, 1 = paper */
float or_board_mask(const uint8_t *tex, uint32_t w, uint32_t h, float u, float v);
/* A posed body (the reference's BLAS, gpu/scene_gpu.js:313-356, placed by setPoses gpu/renderer.js:619-633):
 * BVH2 over its triangles in its own frame, slots as in or_scene.tris (vertex ids index or_scene.verts, whose
 * positions and normals for these vertices are in the body frame), and its pose at shutter open (q0, p0) and
 * close (q1, p1), quaternions (x, y, z, w). At exposure time t it sits at slerp(q0, q1, t), lerp(p0, p1, t). */
typedef struct or_body {
  const float *nodes; const float *tris; uint32_t tri_count;
  float q0[4], q1[4], p0[3], p1[3]; uint32_t hidden;
  /* occlusion capsule in the body frame (radius occ_r, 0 = none) and rho_eff (capsuleVis); the bounding capsule
   * in the world (a.xyz + radius, b.xyz) that the shadow culling tests */
  float occ_a[3], occ_b[3], occ_r, rho_eff;
  float bc_a[4], bc_b[4];
  /* cavityCapsule (gpu/scene_gpu.js:486-503): cavity 1 cuts the posed occlusion capsule at s = cavity_s0 along the
   * unit inward axis from cavity_origin (world); cavity_r > 0 replaces the radius */
  uint32_t cavity;
  float cavity_origin[3], cavity_axis[3], cavity_s0, cavity_r;
} or_body;
typedef struct {
  float R[9];            /* columns: camera x, y, z axes in the world frame */
  float pos[3]; uint32_t mode;
  float lens[9];         /* fx fy cx cy k1 k2 k3 k4 alpha */
  uint32_t width, height, spp, seed, cam_index, light_samples, max_depth, flags;
  float shutter[2];      /* exposure interval on the pose timeline: sub-sample i is at
                            shutter[0] + (shutter[1] - shutter[0]) kr(i, 1/golden ratio, s2) */
  uint32_t skip_bodies;  /* bit i: rays from the camera centre ignore body i (_skipMask, gpu/renderer.js:881-886) */
  uint32_t refine, refine_light_samples;   /* reference estimator: camRefine (kernels.wgsl.js:218-252) */
} or_camera;
/* radiance of pixels [x0, x0+w) x [y0, y0+h) of a camera, spp sub-samples each, mean written as rgba
 * (a = 1) into out[(y - y0) * w + (x - x0)]. Counts traced rays (camera + extension + shadow) into *rays;
 * visits (optional, 5 counts): node visits and triangle tests of all rays, then the camera and extension rays
 * alone (rays, node visits, triangle tests). */
void or_render_tile(const or_scene *S, const or_camera *C, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                    uint32_t spp_begin, uint32_t spp_end, float *out, uint64_t *rays, uint64_t *visits, int nthreads);
/* path-replay adjoint of the tile: adds d(sum_p adj_p . I_p)/d(material colour) (mat_count*3) and
 * d/d(light radiance) (light_count*3) to the outputs; adj holds w*h*4 floats (rgb used) */
void or_backward_tile(const or_scene *S, const or_camera *C, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                      uint32_t spp_begin, uint32_t spp_end, const float *adj, double *grad_mat, double *grad_light,
                      int nthreads);
/* camera-parameter gradient (interior term) by FP64 central differences per sample; grad[14] accumulates
 * (fx fy cx cy k1..k4, rotation about the world axes, camera centre); *changed counts (sample, parameter)
 * pairs whose perturbed paths changed triangles (left out) */
void or_camera_grad_tile(const or_scene *S, const or_camera *C, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                         uint32_t spp_begin, uint32_t spp_end, const float *adj, double rel_step, double *grad,
                         uint64_t *changed, int nthreads);
/* camera ray of sub-sample i of pixel (x, y): o, d (FP32 sample pattern, FP64 unproject, rounded to FP32) */
/* visible (gpu/renderer.js:1238-1258): vis[k] = 1 when no surface lies between the camera centre and tol short of
 * point k (n x 3 floats), the posed bodies at exposure time `time` and the camera's skip_bodies ignored */
void or_visible(const or_scene *S, const or_camera *C, const float *pts, uint32_t n, double tol, double time,
                uint8_t *vis);
/* ---------------------------------------------------------------- node radiosity cache (the reference estimator's)
 * nodes are vertices: npos 4 floats each (position, w = 1 for a static node), nnorm 4 floats (unit normal). */
/* gatherBuild (kernels.wgsl.js:283-306) against the static BVH: gtable[i K + j] = nearest vertex of the hit
 * triangle or 0xffffffff (escape, back face, or a vertex that is no static node) */
void or_gather_build(const or_scene *S, const float *npos, const float *nnorm, uint32_t V, uint32_t K,
                     uint32_t *gtable, int nthreads);
/* gatherRemap (:308-324): entries at inactive nodes (miss > K/2) move to the hit triangle's highest-barycentric
 * active vertex */
void or_gather_remap(const or_scene *S, const float *npos, const float *nnorm, uint32_t V, uint32_t K,
                     const float *miss, uint32_t *gtable, int nthreads);
/* nodeDirect (:348-406) of the listed active nodes: state[16 i + 12..15] = B_dir + visibility code,
 * state[16 i + 3] = capsule visibility (posed bodies at mid-exposure); alb as in or_scene */
void or_node_direct(const or_scene *S, const float *npos, const float *nnorm, const uint32_t *nmat,
                    const uint32_t *active, uint32_t n_active, uint32_t samples, const float *alb, float *state,
                    int nthreads);
/* capsuleVis (shade.wgsl.js:83-94) at a world point with normal n over the scene's visible bodies */
double or_capsule_vis(const or_scene *S, const double p[3], const double n[3]);
void or_camera_ray(const or_camera *C, uint32_t x, uint32_t y, uint32_t i, float o[3], float d[3]);

/* ---------------------------------------------------------------- radiosity */
/* dense Jacobi x <- e + F x (FP64), F row-major n x n (float), e and x n x r; returns sweeps, residual */
int or_jacobi_dense(const float *F, const float *e, uint32_t n, uint32_t r, double tol, int max_iter,
                    double *x, double *res_out);
/* the reference's sparse gather step (kernels.wgsl.js:416-461) over a gather table; see radiosity.c */
void or_jacobi_gather(uint32_t n_active, const uint32_t *active, const uint32_t *gtable, uint32_t K,
                      const float *verts, const uint32_t *node_mat, const float *alb_spec, const float *alb_diff,
                      int bins, const float *hh, const float *node_state, int slot_in, float *aux);
void or_jacobi_update(uint32_t n_active, const uint32_t *active, const uint32_t *nbr, const float *nbr_w,
                      uint32_t nbr_k, const float *aux, const uint32_t *node_mat, const float *hh,
                      float *node_state, int slot_out);
/* the lite renderer's baked node table (lite/shaders.js:234-283): gatherOcc and one Jacobi step */
void or_baked_occ(uint32_t nodes, uint32_t K, const uint16_t *gather, const float *node_a, const float *node_n,
                  const float *caps, uint32_t n_caps, uint8_t *occ, double *dist_out);
void or_baked_jacobi(uint32_t nodes, uint32_t K, const uint16_t *gather, const uint8_t *occ, const float *node_n,
                     const float *hh, const float *e_dir, const float *b_in, float *b_out, float *e_ind);
void or_gather_dir(uint32_t i, uint32_t j, uint32_t K, const float n[3], float out[3]);

#ifdef __cplusplus
}
#endif
#endif
