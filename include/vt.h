/* vt.h — C ABI of the MI355X camera-rendering library (libvt.so).
 *
 * Drop-in boundary for the reference's `TwinRenderer` (gpu/renderer.js:200-1279): scene upload, camera set,
 * render, sensor, readback. Plain pointers and sizes only. Every function returns an int status (VT_OK = 0,
 * negative on error); vt_last_error() gives the context's message. The caller owns all host memory; inputs
 * are copied before a call returns; outputs go to caller buffers whose exact size is checked (VT_ERR_SIZE).
 * The context owns all device memory. A HIP error latches the context into a lost state (every later call
 * returns VT_ERR_LOST), as `device.lost` does for the reference (gpu/renderer.js:241). One host thread uses a
 * context at a time.
 */
#ifndef VT_H
#define VT_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VT_OK = 0,
  VT_ERR_ARG = -1,     /* bad argument (null pointer, out-of-range index, malformed scene) */
  VT_ERR_SIZE = -2,    /* caller buffer size does not match the documented size */
  VT_ERR_HIP = -3,     /* a HIP call failed; the context is now lost */
  VT_ERR_LOST = -4,    /* the context was lost earlier */
  VT_ERR_STATE = -5,   /* call out of order (no scene, camera not set, nothing rendered) */
  VT_ERR_NOMEM = -6,
};

/* Triangle-record words (gpu/scene_gpu.js:351-373) and mesh flags (gpu/wgsl/common.wgsl.js:36-42). */
#define VT_NODE_WORDS 16
#define VT_TRI_WORDS 16
#define VT_VERT_WORDS 8
#define VT_MAT_WORDS 16
#define VT_EMIT_WORDS 16
#define VT_MESH_STATIC 1u
#define VT_MESH_EMITTER 2u
/* the mesh samples the print mask (MESH_FLAG.TEXTURED, gpu/scene_gpu.js:14, 324-325): its rho_d and GGX alpha are
 * mixed between the ink (material words 8-11) and the paper (words 0-3) by the mask at the hit's texture coordinates */
#define VT_MESH_TEXTURED 128u
#define VT_MAX_CAMERAS 64

typedef struct vt_ctx vt_ctx;

typedef struct {
  int device;            /* HIP device ordinal */
  uint32_t flags;        /* reserved, 0 */
} vt_create_info;

/* The prepared scene in the reference's upload layout (gpu/renderer.js:271-331, gpu/scene_gpu.js:461-469):
 *   nodes      BVH2, 16 words per node (gpu/bvh.js:5-8); NULL: the library builds it from tris/verts
 *   tris       16 words per BVH slot: [v0.xyz, i0][e1.xyz, i1][e2.xyz, i2][mesh, flags, -, -]
 *   verts      8 words per vertex: [pos.xyz, mesh id (uint32; the node cache's material)][normal.xyz, -]
 *   meshes     4 words per mesh: material, flags (VT_MESH_*), light index (emitters), -
 *   materials  16 floats per material: [c.rgb, alpha][ior, kind, fbar, roughness][tau.rgb, alpha_c]
 *              [fbar_computed, -, -, -]; kind 0 dielectric, 1 conductor, 2 coated conductor, 3 Lambert
 *              (gpu/bsdf.js:68-88 prepareMaterial); a printed material (charuco_print) keeps the paper's rho_d
 *              and alpha in words 0-3 and the ink's rho_d rgb and alpha in words 8-11 (gpu/renderer.js:293)
 *   lights     4 floats per light: radiance rgb, -
 *   emit       16 floats per emitter triangle: [v0.xyz, cdf][e1.xyz, area][e2.xyz, light (u32 bits)][n.xyz, -]
 *              (gpu/renderer.js:303-306, with the light index in place of the shared L0)
 * Triangles referenced by leaf ranges must be stored in BVH-slot order (the order gpu/bvh.js returns). */
typedef struct {
  const float *nodes; uint32_t node_count;
  const float *tris; uint32_t tri_count;
  const float *verts; uint32_t vert_count;
  const uint32_t *meshes; uint32_t mesh_count;
  const float *materials; uint32_t material_count;
  const float *lights; uint32_t light_count;
  const float *emit; uint32_t emit_count;
  float emit_area;       /* total emitter area (m^2), SceneU.L0.w (gpu/renderer.js:646) */
  /* posed bodies (the reference's BLAS, gpu/renderer.js:538-571): tri_body[k] = the body triangle k belongs to
   * (0xffffffff: the static scene); a body's vertices are in its own frame and vt_set_instances places it. With
   * bodies, nodes must be NULL (every BVH is built here); every body and the static scene need triangles. Emitter
   * triangles lie in the static scene, or all in the body that carries the light (emit_posed). tri_body may be
   * NULL when body_count is 0. */
  const uint32_t *tri_body; uint32_t body_count;
  uint32_t shadow_skip;  /* bit i: shadow rays ignore body i (a light's own housing; SceneU skip mask,
                            gpu/renderer.js:649) */
  /* emit_posed 1: the light rides on body emit_body, as the reference's outlet rides on the scope, whose pose
   * places every emitter sample at its exposure time (shade.wgsl.js:31-47, 55; SceneU.info.y,
   * gpu/renderer.js:652-653). The emit records, the light axes of vt_set_light_profile and the emitter triangles
   * are then in that body's frame, and every emitter triangle must belong to that body. 0: the light is static
   * scene geometry. */
  uint32_t emit_posed, emit_body;
  /* texture coordinates, 2 floats (u, v) per vertex, or NULL; required when a mesh is VT_MESH_TEXTURED (the
   * reference carries them in the dynamic vertices' spare words, gpu/scene_gpu.js:390-392). (0, 0) is the
   * mask's top-left corner. */
  const float *uvs;
} vt_scene;

#define VT_MAX_BODIES 24
/* Pose of body i over the exposure (setPoses, gpu/renderer.js:619-633): rotation as a unit quaternion
 * (x, y, z, w) and position, at shutter open (q0, p0) and close (q1, p1); a sub-sample at exposure time t sees
 * slerp(q0, q1, t), lerp(p0, p1, t). hidden = 1 leaves the body out (a body without a pose). */
typedef struct {
  float q0[4], q1[4];
  float p0[3], p1[3];
  uint32_t hidden;
  /* occlusion capsule in the body's frame (segment occ_a-occ_b, radius occ_r; 0 = none) and the effective
   * reflectance rho_eff of the body: the reference estimator dims a static point's cached indirect light by
   * 1 - (1 - rho_eff) O(capsule) (capsuleVis, gpu/wgsl/shade.wgsl.js:57-94), the capsule posed at mid-exposure */
  float occ_a[3], occ_b[3], occ_r, rho_eff;
  /* cavityCapsule (gpu/scene_gpu.js:474-503, applied at gpu/renderer.js:680-686): only the part of the posed
   * occlusion capsule inside the cavity counts. With cavity = 1 the mid-exposure capsule is cut at the plane
   * s = cavity_s0, s(X) = (X - cavity_origin) . cavity_axis (world frame; the axis points into the cavity): an end
   * with s < cavity_s0 moves along the segment onto the plane, and a capsule wholly below it is dropped;
   * cavity_r > 0 replaces the radius (a cannula's sheath). The reference's rules, for a body through a port
   * (origin = the port's pivot, axis = -axis_out): a cannula s0 = 0 and r = cannula_outer_radius, an instrument or
   * the scope s0 = cannula_toe_insertion; lens fronts and boards have no capsule (occ_r = 0). cavity = 0: uncut. */
  uint32_t cavity;
  float cavity_origin[3], cavity_axis[3], cavity_s0, cavity_r;
} vt_instance;

/* One tracking camera, CamU (gpu/wgsl/common.wgsl.js:79-90; gpu/renderer.js:888-898). */
typedef struct {
  float R[9];            /* columns: camera x, y, z axes in the world frame (R_D_C) */
  float pos[3];          /* camera centre in the world frame */
  uint32_t mode;         /* 0 KB4, 1 pinhole */
  float lens[9];         /* fx fy cx cy k1 k2 k3 k4 alpha (gpu/lens.js:8-16 lensParams) */
  uint32_t width, height, spp, seed;
  uint32_t cam_index;    /* camera index in the pixel-seed hash (gpu/wgsl/kernels.wgsl.js:35) */
  uint32_t light_samples;/* emitter samples per path vertex, 1..16 */
  uint32_t max_depth;    /* path vertices, >= 1 */
  uint32_t flags;        /* option bits VT_OPT_* (CamU.flags.y, gpu/wgsl/common.wgsl.js:45-58) */
  float scale;           /* output scale (CamU.expo.z), usually 1 */
  float shutter[2];      /* exposure interval on the pose timeline (CamU.expo.xy), within [0, 1]: sub-sample i
                            is at time shutter[0] + (shutter[1] - shutter[0]) frac(s2 + i / golden ratio) */
  uint32_t skip_bodies;  /* bit i: this camera's own rays (from its centre) ignore body i, e.g. its lens front
                            (_skipMask, gpu/renderer.js:881-886) */
  /* reference estimator: the refinement pass (camRefine, gpu/wgsl/kernels.wgsl.js:218-252; DEFAULTS.refine and
   * refineLightSamples, gpu/renderer.js:33-35): pixels on a mesh edge or a posed body (and, with
   * VT_OPT_REFINE_SHADOWS, those whose shadow rays were traced) take `refine` more sub-samples with
   * refine_light_samples emitter samples each. 0 = no refinement. Needs a whole frame (spp_begin 0, no accumulate). */
  uint32_t refine, refine_light_samples;
} vt_camera;

enum { VT_RENDER_BATCHED = 0, VT_RENDER_PER_CAMERA = 1, VT_RENDER_WAVEFRONT = 2, VT_RENDER_AUTO = 3 };
typedef struct {
  uint32_t mode;         /* VT_RENDER_BATCHED: every listed camera in one persistent launch over a global ray
                            queue; VT_RENDER_PER_CAMERA: one plain launch per camera; VT_RENDER_WAVEFRONT: the
                            frame's paths in batches advanced one vertex at a time (trace, shade, shadow, gather
                            kernels), live paths radix-sorted by ray direction and origin between vertices;
                            VT_RENDER_AUTO: the form vt_autotune found fastest (batched until it has run) */
  uint32_t spp_begin;    /* sub-sample range [spp_begin, spp_end); spp_end 0 = each camera's spp */
  uint32_t spp_end;
  uint32_t accumulate;   /* 0: clear the cameras' sums first; 1: add to them (progressive ranges) */
  uint32_t chunk;        /* queue chunk in 64-item waves (0 = tuned default); wavefront: idle lanes of a wave
                            that make it fetch new rays, 1..64 (0 = 16) */
  uint32_t block;        /* threads per block: 64, 128 or 256 (0 = tuned default) */
  uint32_t stack_lds;    /* traversal stack entries per lane kept in LDS, 8..24 (0 = tuned default); the rest of
                            the 48-entry stack spills to global memory. Wavefront: 4, 8 or 16 (0 = 8) */
  uint32_t queues;       /* batched: 1 global ray queue, or 8 queues (one per XCD, over contiguous eighths of the
                            frame, drained by the blocks of that XCD first); wavefront: the same for each trace
                            launch's rays; 0 = tuned default */
  uint32_t graph;        /* batched and wavefront: 1 = run the frame's launches as a hipGraph, captured on first
                            use and replayed while the cameras, sample range and launch shape stay the same (the
                            wavefront graph does not time its extension traces one by one) */
  uint32_t batch;        /* wavefront: paths per batch, rounded up to 64 (0 = 134,217,728, halved until the batch's
                            buffers fit an eighth of the device's memory and its shadow slots 32 bits) */
  uint32_t sort_bits;    /* wavefront: the sort of the live paths between vertices; 0 = tuned default, 1 = the
                            radix sort of the whole 23-bit key (direction cell, origin Morton code), 4..13 = an
                            unstable bin sort by the key's top bits (kernels only; a graph uses it, 12 bits when
                            the default is the radix sort). The image does not depend on it. */
  uint32_t shadow_sort;  /* wavefront: 1 = trace each vertex's shadow rays in the order of their (direction cell,
                            origin Morton code) keys, sorted like the paths; 2 = in slot order; 0 = tuned default.
                            The image does not depend on it. */
} vt_render_opts;

/* Result of vt_autotune: the chosen launch shapes (also made the context's defaults) and the probe timings. */
typedef struct {
  uint32_t block, stack_lds, chunk, queues;   /* persistent batched form */
  uint32_t candidates;   /* launch shapes timed (both forms) */
  double probe_mrays_best;    /* Mrays/s of the chosen persistent shape on the probe */
  double probe_mrays_default; /* Mrays/s of the built-in default (256 threads, 16 LDS entries, chunk 4, 1 queue) */
  double probe_fraction;      /* share of the frame's waves one persistent-form probe traces (wavefront: 8x) */
  uint32_t mode;              /* the faster form: VT_RENDER_BATCHED or VT_RENDER_WAVEFRONT (used by VT_RENDER_AUTO) */
  uint32_t wf_batch, wf_stack_lds, wf_refill;  /* wavefront form: paths per batch, LDS stack, refill threshold */
  double probe_mrays_wavefront;               /* Mrays/s of the chosen wavefront shape on the same probe */
  uint32_t wf_sort_bits;                      /* wavefront: 0 the whole-key radix sort, else bins of that many bits */
} vt_tune_result;

/* Sensor parameters, the 256-B sensor uniform (gpu/sensor.js:158-177), raw-DN part. */
#define VT_MAX_VIGNETTING 8
typedef struct {
  uint32_t cfa;          /* 0 RGGB, 1 GRBG, 2 GBRG, 3 BGGR */
  uint32_t frame, seed;
  float optics_k;        /* pi T / (4 N^2) */
  float esf_sigma;       /* optics blur sigma (px), after the pixel-aperture part */
  float exposure, gain;
  float m[9];            /* response matrix rows (e-/s per W m^-2) */
  float read_noise, full_well, dark, conv_gain;
  float max_dn, black;
  uint32_t noise_on;
  float pupil_exp;       /* relative-illumination cos exponent (gpu/sensor.js:119) */
  /* veiling glare (sensor.wgsl.js:87-141, the uniform's glare word, gpu/sensor.js:172-173): fraction k (0 = off),
   * the Gaussian's sigma in pixels of the 1/F level, the level factor F (1..32) */
  float glare_fraction, glare_sigma;
  uint32_t glare_level;
  /* vignetting V(theta) (physics lens.relative_illumination vignetting_theta_deg / vignetting; gpu/sensor.js:110-114):
   * linear between `vignetting_count` knots (field angle ascending, degrees), the end values outside; 0 knots: V = 1.
   * The relative-illumination map is the KB4 term times V. */
  uint32_t vignetting_count;
  float vignetting_theta_deg[VT_MAX_VIGNETTING], vignetting[VT_MAX_VIGNETTING];
} vt_sensor_params;

/* The ISP after the raw DN (sensor.wgsl.js:224-427; the uniform's wb, c0-c2, tone, flags and tone2 words,
 * gpu/sensor.js:169-176). */
typedef struct {
  uint32_t demosaic;     /* 0 bilinear, 1 edge-aware (OpenCV EA), 2 Hamilton-Adams */
  uint32_t nr;           /* 1: 3x3 bilateral noise reduction (used when nr_sigma_px > 0) */
  uint32_t tone;         /* 0 sRGB, 1 gamma, 2 linear, 3 sRGB + contrast S-curve */
  uint32_t chroma420;    /* 1: RGB from 4:2:0 chroma and the YUV planes; 0: 4:4:4 RGB, no YUV planes */
  float wb[3];           /* white-balance gains */
  float ccm[9];          /* colour correction matrix rows */
  float gamma;           /* tone 1 */
  float s_curve_c;       /* tone 3: contrast exponent c */
  float black_offset;    /* tone 3: black offset as a fraction of white */
  float nr_sigma_px;     /* bilateral spatial sigma (px) */
  float nr_range_scale;  /* bilateral range sigma = scale x the modelled noise */
  float sharpen_amount;  /* unsharp mask on luma, 0 = off */
  float sharpen_radius;  /* box radius (px), rounded and clamped to 1..4 */
} vt_isp_params;

/* readback kinds: LINEAR, RAW and RI as above; the ISP's outputs (gpu/renderer.js:1190-1231):
 *   VT_KIND_RGB8    uint8 rgb, width*height*3 bytes
 *   VT_KIND_RGBA8   uint8 rgba (a = 255), width*height*4 bytes
 *   VT_KIND_YUV420  uint8 planes Y (width*height), Cb, Cr (ceil(width/2)*ceil(height/2) each); 4:2:0 only
 *   VT_KIND_STATS   uint32[4]: the luma sum of the 8-bit Y, the grey-world sums of round(clamp(c)*1023) r g b
 *                   (the AE / AWB statistics, gpu/renderer.js:1075-1089) */
enum { VT_KIND_LINEAR = 0, VT_KIND_RAW = 1, VT_KIND_RI = 2, VT_KIND_RGB8 = 3, VT_KIND_RGBA8 = 4,
       VT_KIND_YUV420 = 5, VT_KIND_STATS = 6 };

typedef struct {
  double rays;           /* rays traced by the last vt_render (camera, extension and shadow) */
  double samples;        /* path samples of the last vt_render */
  double ms_render;      /* GPU time of the last vt_render (all its launches) */
  double ms_sensor;      /* GPU time of the last vt_sensor */
  double ms_kernel_avg;  /* average duration of the last vt_render's trace launches */
  uint32_t launches;     /* trace launches of the last vt_render */
  uint32_t stack_overflows; /* traversal stack overflows since upload (must be 0) */
  double ms_backward;    /* GPU time of the last vt_backward */
  double rays_backward;  /* rays traced by the last vt_backward (two walks per sample) */
  uint32_t graph_used;   /* 1 if the last vt_render ran as a hipGraph */
  uint32_t graph_builds; /* hipGraph captures since the context was created */
  double rays_extension; /* wavefront: camera and extension rays of the last vt_render */
  double rays_shadow;    /* wavefront: shadow rays of the last vt_render */
  double ms_ext_avg;     /* wavefront: average duration of its extension-trace launches */
  uint32_t ext_launches; /* wavefront: extension-trace launches of the last vt_render */
  double ms_isp;         /* GPU time of the last vt_isp */
} vt_frame_stats;

/* ---------------------------------------------------------------- lifetime (TwinRenderer.create / destroy,
 * gpu/renderer.js:202-261, 1278) */
int vt_create(const vt_create_info *info, vt_ctx **out);
void vt_destroy(vt_ctx *ctx);
const char *vt_last_error(const vt_ctx *ctx);

/* Binned-SAH BVH2 of gpu/bvh.js:18-161 (16 bins, leaf 4, same node numbering and triangle order), native.
 * cap_nodes >= ntri + 2 always suffices. order_out gets ntri entries. Replaces the JS buildBVH call in
 * prepareScene (gpu/scene_gpu.js:305-307). No context needed. */
int vt_build_bvh(const float *pos, uint32_t nvert, const uint32_t *idx, uint32_t ntri, int leaf_size, int bins,
                 float *nodes_out, uint32_t cap_nodes, uint32_t *node_count, uint32_t *order_out);
/* weld + subdivide of prepareScene's static meshes (gpu/scene_gpu.js:66-143, called at :268-270): vertices merged
 * at (1e-7 m, 1e-4) in order of first appearance, collapsed triangles dropped, then conforming edge bisection until
 * no edge exceeds max_edge (at most max_passes passes; the reference uses 24). Outputs in FP64 as the reference's
 * arrays hold them (positions, normals) and the triangles. The counts are always written; VT_ERR_SIZE when a
 * capacity is short (call again with room). No context needed. */
int vt_weld_subdivide(const float *pos, const float *nrm, uint32_t nvert, const uint32_t *idx, uint32_t ntri,
                      double max_edge, uint32_t max_passes, double *pos_out, double *nrm_out, uint32_t cap_vert,
                      uint32_t *idx_out, uint32_t cap_tri, uint32_t *nvert_out, uint32_t *ntri_out);

/* ---------------------------------------------------------------- scene and cameras */
/* Buffers of _buildBuffers (gpu/renderer.js:271-331): copies the scene in, re-lays the BVH2 out as the
 * device's 128-B four-wide nodes, and uploads. Replaces any earlier scene. */
int vt_upload_scene(vt_ctx *ctx, const vt_scene *scene);
/* CamU packing of _packCam (gpu/renderer.js:888-898) for camera slot `cam` (< VT_MAX_CAMERAS). */
int vt_set_camera(vt_ctx *ctx, uint32_t cam, const vt_camera *camera);
/* setPoses (gpu/renderer.js:619-633, _writeSceneU 640-705): one pose record per body of the uploaded scene
 * (count = body_count); also derives each body's swept world box and bounding capsule for culling. Bodies start
 * hidden after an upload. */
int vt_set_instances(vt_ctx *ctx, const vt_instance *instances, uint32_t count);

/* Replace the material table or the light radiances of the uploaded scene in place (same counts as the upload),
 * e.g. between the steps of a calibration loop; the geometry and BVH stay. No reference counterpart: the
 * reference rebuilds its params buffer in _buildBuffers (gpu/renderer.js:271-331). */
int vt_update_materials(vt_ctx *ctx, const float *materials, uint32_t material_count);

/* Angular emission profile of one light (physics.yaml light.profile; shade.wgsl.js:12-27 profile, :48-52 emitted):
 * L_e(w) = radiance * p(theta), theta the angle between the outgoing direction w and the light's axis, p linear in
 * theta between `count` knots (theta ascending, radians; p >= 0), 0 past the last knot; the emitter still emits on
 * its front side only. count = 0 makes the light Lambertian again. Profiles survive vt_update_lights and are
 * cleared by vt_upload_scene. The axis is in the emitter's frame: the world for a static light,
 * its body's frame for a light that rides on a body (vt_scene.emit_posed). */
#define VT_MAX_KNOTS 64
int vt_set_light_profile(vt_ctx *ctx, uint32_t light, const float axis[3], const float *theta, const float *p,
                         uint32_t count);
/* The printed board's mask (_loadBoardTexture, gpu/renderer.js:395-412; sampled in shadeHit,
 * gpu/wgsl/shade.wgsl.js:258-265): width x height 8-bit texels, row 0 the image's top row, 255 = paper, 0 = ink
 * (the PNG's red channel). Textured meshes sample it bilinearly with clamp-to-edge addressing, texel (i, j)
 * centred at ((i + 0.5) / width, (j + 0.5) / height). Until it is set (and after vt_upload_scene) the mask is
 * one paper texel, as the reference's 1x1 white fallback. At most 16384 x 16384. */
int vt_set_board_texture(vt_ctx *ctx, const uint8_t *mask, uint32_t width, uint32_t height);
int vt_update_lights(vt_ctx *ctx, const float *lights, uint32_t light_count);

/* ---------------------------------------------------------------- per frame */
/* renderCamera / renderFrame (gpu/renderer.js:974-996): enqueue the listed cameras' path tracing. */
int vt_render(vt_ctx *ctx, const uint32_t *cams, uint32_t n, const vt_render_opts *opts);
/* Autotuner (no reference counterpart; the reference fixes its workgroup shape): times the persistent batched
 * launch of the listed cameras over block {64, 128, 256} x LDS stack {8, 16, 24} x chunk {1, 4, 16} x queues
 * {1, 8}, and the wavefront form at its default batch over LDS stack {4, 8, 16} x refill {8, 16, 32}, shadow rays
 * in slot or key order, one or eight trace queues and four sorts between vertices, each on a probe of every
 * `probe_stride`-th wave of the frame (0 = 32); picks the fastest of each form in rays per second, makes
 * them the context's defaults for fields of vt_render_opts left 0, and the faster form the one VT_RENDER_AUTO
 * runs. The wavefront's probes cover eight times more waves than the persistent form's (its rate grows with the
 * paths it sorts together), and the form is judged at that size (out->mode; probe_mrays_best and _default are
 * the persistent form's small probes). opts: the sample range (as for vt_render). The listed cameras' accumulated images are cleared. Blocks. */
int vt_autotune(vt_ctx *ctx, const uint32_t *cams, uint32_t n, const vt_render_opts *opts, uint32_t probe_stride,
                vt_tune_result *out);
/* sensor(name) raw stage (gpu/renderer.js:1091-1099; sensorOptics + sensorExpose fused into one kernel). */
int vt_sensor(vt_ctx *ctx, uint32_t cam, const vt_sensor_params *params);
/* sensor(name) ISP stages after the raw DN (gpu/renderer.js:1091-1099, the demosaicG .. sensorOutput passes of
 * :1034-1037): demosaic, WB + CCM, AE / AWB statistics, noise reduction, tone, sharpening, BT.601 YCbCr. Uses the
 * camera's last vt_sensor output and its parameters (black, max DN, noise model, CFA). */
int vt_isp(vt_ctx *ctx, uint32_t cam, const vt_isp_params *isp);
/* readback(name, kind) (gpu/renderer.js:1191-1231). Blocks until the data is ready.
 *   VT_KIND_LINEAR  float32 rgba, width*height*16 bytes (mean radiance times scale; a = 1)
 *   VT_KIND_RAW     uint16 DN, width*height*2 bytes (the reference returns these integers as float32)
 *   VT_KIND_RI      float32 relative illumination, width*height*4 bytes */
int vt_readback(vt_ctx *ctx, uint32_t cam, int kind, void *dst, size_t dst_bytes);
int vt_sync(vt_ctx *ctx);
/* timings() (gpu/renderer.js:1271-1274): statistics of the last render / sensor call. */
int vt_frame_stats_get(vt_ctx *ctx, vt_frame_stats *out);

/* ---------------------------------------------------------------- adjoint (no reference counterpart: the
 * differentiable path of the twin's calibration) */
/* Gradients of the scalar sum_cams sum_pixels adj . I (adj = dLoss/dI, the adjoint image; I the linear image
 * vt_readback returns) with respect to the material colours (rho_d, or F0 of conductors; 3 floats per material)
 * and the light radiances (3 floats per light), by path-replay backpropagation: each sample's path is re-traced
 * from its seeded random numbers, first for its value, then accumulating the derivative of every term (sampling
 * and MIS weights detached; interior term only, visibility discontinuities excluded). adj: one pointer per listed
 * camera to width*height*4 floats (rgb used). opts: the sample range as for vt_render. Buffers of vt_grads are
 * overwritten; their counts must equal the scene's material and light counts. cameras (optional): one row of
 * VT_CAMERA_PARAMS per listed camera, the derivatives with respect to that camera's intrinsics and pose, by
 * forward-mode tangents of the camera ray carried along the replayed path with its continuation directions and
 * pdfs held at their sampled values (sampling detached): every hit re-derived on its triangle's plane as the
 * vertices move, and the shading, emission and NEE of the moved points (same random numbers and triangles;
 * interior term only: pixel-footprint and silhouette edges excluded). */
/* Camera parameters of the adjoint, per listed camera: fx fy cx cy k1 k2 k3 k4 (the KB4 intrinsics of
 * vt_camera.lens), then the pose: rx ry rz (a small rotation about the world axes applied to R, R' = exp([r]x) R)
 * and tx ty tz (the camera centre in the world frame). */
#define VT_CAMERA_PARAMS 14
typedef struct {
  float *materials; uint32_t material_count;
  float *lights; uint32_t light_count;
  float *cameras; uint32_t camera_count;   /* VT_CAMERA_PARAMS per listed camera, or NULL */
} vt_grads;
int vt_backward(vt_ctx *ctx, const uint32_t *cams, uint32_t n, const float *const *adj, const vt_render_opts *opts,
                vt_grads *out);

/* visible (gpu/renderer.js:1238-1258): out[k] = 1 when no surface lies between camera cam's centre and tol short of
 * world point k (points: n x 3 floats), over the static scene and the posed bodies at exposure time `time` in
 * [0, 1], the camera's skip_bodies ignored; a point within tol of the centre is visible. The reference walks its
 * BVH on the CPU; here one any-hit launch answers all n. Blocks until done. */
int vt_visible(vt_ctx *ctx, uint32_t cam, const float *points, uint32_t n, float tol, float time, uint8_t *out);

/* Copy bandwidth: y = x over two `bytes`-sized device buffers, `iters` times, read + written bytes per second
 * (GB/s). With bytes well under 256 MB the working set stays in the Infinity Cache; the roofline of the trace
 * kernels is reported against it as well as against HBM. No reference counterpart. */
int vt_copy_bandwidth(vt_ctx *ctx, uint64_t bytes, uint32_t iters, double *gbs);
/* Read bandwidth: one `bytes`-sized device buffer read `iters` times by a grid-stride kernel (16-B loads, four in
 * flight per lane, nothing written), bytes read per second (GB/s). With bytes well over 256 MB this is the
 * streaming-read rate HBM delivers to a kernel, the practical ceiling of read-bound kernels such as the dense
 * sweeps. No reference counterpart. */
int vt_read_bandwidth(vt_ctx *ctx, uint64_t bytes, uint32_t iters, double *gbs);

/* ---------------------------------------------------------------- test kernels (the reference's TESTS entry
 * points, gpu/wgsl/kernels.wgsl.js:487-551, that tests/gpu_test.html compared against the JS) */
/* traceTest (:534-543): rays 8 floats each (o.xyz, tmin, d.xyz, tmax) -> hits 4 words each
 * (t as float, tri as uint32 (0xffffffff = miss), u, v). */
int vt_trace_test(vt_ctx *ctx, const float *rays, uint32_t n, int any_hit, float *hits);
/* lensTest (:494-505): uv[2n] -> unit camera rays rays[3n] by the FP32 KB4 inverse (8 Newton steps). */
int vt_lens_test(vt_ctx *ctx, const float lens[9], const float *uv, uint32_t n, float *rays);
/* camera rays of sub-samples: for each (x, y, i) triple in xyi[3n] of camera `cam`, o.xyz d.xyz in out[6n] */
int vt_camera_ray_test(vt_ctx *ctx, uint32_t cam, const uint32_t *xyi, uint32_t n, float *out);
/* bsdfTest (:507-523): cases of 16 floats (n.xyz, -, l.xyz, -, v.xyz, -, u0 u1 u2, material) ->
 * 8 floats each (f.rgb, pdf(v, l), sampled l.xyz, sampled pdf). */
int vt_bsdf_test(vt_ctx *ctx, const float *cases, uint32_t n, float *out);
/* capsuleOcc (gpu/wgsl/shade.wgsl.js:57-81): cases of 16 floats (p.xyz, -, n.xyz, -, a.xyz, r, b.xyz, -) ->
 * the cosine-weighted occluded fraction, one float each */
int vt_capsule_test(vt_ctx *ctx, const float *cases, uint32_t n, float *out);
/* The fused sensor epilogue on caller-given radiance (rgba float32 w*h*4) and RI map (w*h) -> DN (w*h). */
int vt_sensor_test(vt_ctx *ctx, const vt_sensor_params *params, uint32_t width, uint32_t height,
                   const float *radiance, const float *ri, uint16_t *dn);
/* vt_sensor_test on caller DEVICE buffers, allocating nothing. No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * radiance (w*h float4 rgba, a = the pixel's sample count), ri (w*h floats) and dn (w*h uint16) are device memory of
 * the context's device (host memory is refused, VT_ERR_ARG); the work is enqueued on the context's stream (vt_sync
 * waits for it). The veiling-glare levels, its only scratch, live in a static device array of the library
 * (VT_SENSOR_GLARE_MAX float4 per level, two levels: 8 MiB held on every GPU the library runs on until the process
 * ends); a glare level ceil(w/F) x ceil(h/F) above VT_SENSOR_GLARE_MAX is refused with VT_ERR_SIZE. Calls that
 * overlap on one GPU (two contexts or threads at once) share that array and race. */
#define VT_SENSOR_GLARE_MAX 262144
int vt_sensor_raw_device(vt_ctx *ctx, const vt_sensor_params *params, uint32_t width, uint32_t height,
                         const float *radiance, const float *ri, uint16_t *dn);
/* The ISP on caller-given raw DN (w*h uint16): rgba (w*h*4 bytes), yuv (NULL, or the planes as VT_KIND_YUV420
 * when chroma420), stats (uint32[4], as VT_KIND_STATS). */
int vt_isp_test(vt_ctx *ctx, const vt_sensor_params *sensor, const vt_isp_params *isp, uint32_t width,
                uint32_t height, const uint16_t *dn, uint8_t *rgba, uint8_t *yuv, uint32_t *stats);

/* Camera option bits. The reference's own (common.wgsl.js:45-58; they act in the reference estimator):
 * NO_INDIRECT / DIRECT_ONLY drop the cached indirect light, NO_CAPSULE the bodies' occlusion of it, NORMALS
 * renders shading normals, NO_MIS leaves direct light to light sampling alone. Added here: REFERENCE selects the
 * reference's camera estimator (shadeHit, gpu/wgsl/shade.wgsl.js:250-334: first hit, nL light samples, static
 * surfaces read the node radiosity cache, posed bodies take one BSDF bounce) instead of the depth-D path tracer;
 * it needs vt_radiosity_build. CULL_SHADOWS skips the shadow rays whose answer the hit triangle's three nodes
 * give (cachedVisibility, :108-130). */
#define VT_OPT_NO_INDIRECT 4u
#define VT_OPT_DIRECT_ONLY 8u
#define VT_OPT_NO_CAPSULE 32u
#define VT_OPT_NORMALS 64u
#define VT_OPT_NO_MIS 128u
#define VT_OPT_REFINE_SHADOWS 1024u   /* refine pixels whose shadow rays were traced (refineShadows) */
/* reference estimator, fixed cameras: keep each sub-sample's static first hit and trace only the bodies on later
 * frames (hitCache, gpu/renderer.js:47, 940-967); exact. Reset when the camera's rays or the scene change. */
#define VT_OPT_HIT_CACHE 2048u
/* reference estimator, still light: static surfaces test each light sample against the static scene (kept per
 * sub-sample and light sample while the camera, its rays and the scene hold) and then against the bodies; a pixel's
 * shadow key marks a body's shadow, a penumbra or light samples that disagree by more than 20 %
 * (SHADOW_CACHE, gpu/wgsl/shade.wgsl.js:272-301; shadowCache, gpu/renderer.js:48, 951-956) */
#define VT_OPT_SHADOW_CACHE 8192u
#define VT_OPT_REFERENCE (1u << 16)
#define VT_OPT_CULL_SHADOWS (1u << 17)

/* ---------------------------------------------------------------- node radiosity on the uploaded scene
 * The reference's radiosity cache (gather table gpu/renderer.js:456-530, kernels.wgsl.js:271-344; direct light
 * and Jacobi :348-461, renderer.js:707-852). Nodes are the scene's vertices; a vertex of a static-flagged,
 * non-body triangle is a static node (its material: that of its first such triangle). Build: gather_k
 * cosine-distributed rays per static node against the static scene (gatherDir), each entry the nearest vertex of
 * the hit triangle (0xffffffff for an escape or a back face); a node is active when at most half its rays
 * escape; entries at inactive nodes move to the hit triangle's highest-barycentric active vertex; nbr_k
 * same-facing (n.n > 0.9) neighbours within smooth_radius with Gaussian weights (sigma smooth_sigma; 0 = none);
 * the directional-albedo tables of every material (32 bins, 1024 VNDF Hammersley samples, bsdf.js:210-254). */
typedef struct {
  uint32_t gather_k;         /* gather rays per node (64) */
  uint32_t node_samples;     /* emitter samples per node per update (4) */
  uint32_t nbr_k;            /* smoothing neighbours (16, at most 32) */
  float smooth_radius;       /* m (0.006) */
  float smooth_sigma;        /* m (0.003; 0 = no smoothing) */
} vt_radiosity_opts;
typedef struct {
  uint32_t nodes, static_nodes, active;
  double ms_build;
} vt_radiosity_info;
int vt_radiosity_build(vt_ctx *ctx, const vt_radiosity_opts *opts, vt_radiosity_info *info);
/* The interior tree (_buildInteriorBVH, gpu/renderer.js:533-571), after vt_radiosity_build: enable 1 builds (once
 * per node cache) a second static BVH without the triangles whose three nodes' gather rays escape in >= 90 % of
 * cases (bench triangles always kept) and makes it the root of every later camera, shadow, node and trace walk;
 * enable 0 returns to the whole scene (the reference's fullBVH). vt_visible always walks the whole scene.
 * interior_tris (optional): the static triangles the walks now see. */
int vt_interior_bvh(vt_ctx *ctx, uint32_t enable, uint32_t *interior_tris);
/* One node update (_ensureRadiosity / converge): with direct != 0, nodeDirect (node_samples emitter samples per
 * active node at mid-exposure, shadow rays through the static scene and the posed bodies, the capsule visibility
 * of its indirect light); then `iters` Jacobi steps (jacobiGather + jacobiUpdate, ping-pong source slots), warm
 * started from the previous state. Enqueue only. */
int vt_radiosity_update(vt_ctx *ctx, uint32_t direct, uint32_t iters);
/* The incremental node update of _ensureRadiosity (gpu/renderer.js:756-795: _movedSinceSolve, _submitNodeDirect
 * with a moved mask; nodeDirect's moved-body test, gpu/wgsl/kernels.wgsl.js:357-366). moved: bit b set for each
 * body whose pose changed since the last direct pass (bodies < 32). The direct pass then relights only the active
 * nodes whose segment to the emitter's centre passes within (bounding-capsule radius + emitter radius + 1 mm) of a
 * moved body's capsule, now or at the last direct pass; every other node keeps its B_dir, visibility code and
 * capsule visibility. A full pass (every node) runs instead when the light's own body moved (vt_scene.emit_posed),
 * no direct pass ran since vt_radiosity_build, or a material, light or profile change made the cache stale;
 * moved = 0 runs no direct pass. Then `iters` Jacobi
 * steps as vt_radiosity_update. ran (optional): 0 no direct pass, 1 incremental, 2 full. Enqueue only. */
enum { VT_DIRECT_NONE = 0, VT_DIRECT_INCREMENTAL = 1, VT_DIRECT_FULL = 2 };
int vt_radiosity_update_moved(vt_ctx *ctx, uint32_t moved, uint32_t iters, uint32_t *ran);
/* nodeSummary (gpu/renderer.js:824-852): the node state (16 floats per node: E_free.rgb + V, slot 1 source,
 * slot 2 source (w = 1 for a solved node), B_dir.rgb + visibility code) into node_state (NULL: skipped; else
 * nodes * 64 bytes) and the area-weighted mean irradiance eAvg over the active nodes, which the reference
 * estimator's bounce from a body to another body reads; eAvg is also stored for rendering. Blocks. */
int vt_radiosity_state(vt_ctx *ctx, float *node_state, size_t bytes, float e_avg[3]);
/* The built tables, for inspection (blocks): which = VT_RAD_GATHER (nodes x gather_k uint32), VT_RAD_MISSES
 * (nodes floats: escaping gather rays), VT_RAD_ACTIVE (active uint32 node ids), VT_RAD_NEIGHBOURS (active x nbr_k
 * uint32 active indices), VT_RAD_WEIGHTS (active x nbr_k floats), VT_RAD_ALBEDO (materials x 32 bins x 8 floats),
 * VT_RAD_HH (materials x 4 floats); bytes must equal the table's size. */
enum { VT_RAD_GATHER = 0, VT_RAD_MISSES = 1, VT_RAD_ACTIVE = 2, VT_RAD_NEIGHBOURS = 3, VT_RAD_WEIGHTS = 4,
       VT_RAD_ALBEDO = 5, VT_RAD_HH = 6 };
int vt_radiosity_tables(vt_ctx *ctx, uint32_t which, void *dst, size_t bytes);
/* The smoothing neighbours of vt_radiosity_build on caller-given nodes (host only, no device): pos and nrm n x 3
 * floats; idx and w n x k (k <= 32), rows sorted by weight, 0-weight entries unused. */
int vt_node_neighbours(const float *pos, const float *nrm, uint32_t n, uint32_t k, float radius, float sigma,
                       uint32_t *idx, float *w);

/* ---------------------------------------------------------------- the baked node table (the lite renderer)
 * The lite renderer's indirect light (lite/renderer.js:214-222, 486-517; lite/shaders.js:220-283): a baked table of
 * K cosine gather rays per node (assets/lite_nodes.bin: 12,976 nodes x K = 256 16-bit hit ids, 0xffff for a ray
 * that escapes), resident on the device. Per pose change: vt_baked_occlusion with the posed capsules (gatherOcc),
 * then vt_baked_jacobi with the nodes' reflected direct light (jacobi), warm started. */
typedef struct {
  uint32_t nodes, K;          /* nodes < 65535; K a multiple of 4, at most VT_BAKED_MAX_K */
  const uint16_t *gather;     /* nodes x K hit node ids (0xffff: the ray escapes) */
  const float *node_a;        /* nodes x 4: position xyz (m), area (m^2) (nodeA, lite/renderer.js:215-217) */
  const float *node_n;        /* nodes x 4: unit normal xyz, mesh id (uint32 bits) (nodeN) */
  const float *hh;            /* meshes x 4: hemispherical albedo rgb, - (Mat.hh) */
  uint32_t meshes;
} vt_baked_table;
#define VT_BAKED_MAX_K 1024
#define VT_BAKED_MAX_CAPS 16
/* Upload the table; every ray transmits, B, E_dir and E_ind start at 0. Blocks. */
int vt_baked_upload(vt_ctx *ctx, const vt_baked_table *table);
/* gatherOcc (lite/shaders.js:234-258): per gather ray the transmission byte 255, or min over the capsules whose
 * radius reaches the segment from the node (lifted 1 mm along its normal) to the hit node of round(255 rho).
 * caps: n_caps x 8 floats [a.xyz, radius, b.xyz, rho] (posed, world; the body's mean albedo rho). Enqueue only. */
int vt_baked_occlusion(vt_ctx *ctx, const float *caps, uint32_t n_caps);
/* `iters` steps of jacobi (lite/shaders.js:260-283): E_ind = pi/K sum_k t_k B(hit_k), t_k = byte / 255;
 * B' = (E_dir + hh E_ind) / pi, ping-ponged from the latest B. e_dir: nodes x 4 floats (E_dir A(mu_l) rgb, -) to
 * upload first, or NULL to keep the last. Enqueue only. */
int vt_baked_jacobi(vt_ctx *ctx, const float *e_dir, uint32_t iters);
/* Read back (blocks): VT_BAKED_B the latest B (nodes x 4 floats), VT_BAKED_EIND the last step's E_ind
 * (nodes x 4 floats), VT_BAKED_OCC the transmission bytes (nodes x K). bytes must equal the size. */
enum { VT_BAKED_B = 0, VT_BAKED_EIND = 1, VT_BAKED_OCC = 2 };
int vt_baked_read(vt_ctx *ctx, uint32_t which, void *dst, size_t bytes);

/* ---------------------------------------------------------------- node radiosity (host-given tables) */
/* One Jacobi step of the reference's sparse gather form (jacobiGather + jacobiUpdate,
 * gpu/wgsl/kernels.wgsl.js:416-461; _submitJacobi gpu/renderer.js:729-749) on host-given tables. node_state
 * (16 floats per node) is read and written in place. */
typedef struct {
  uint32_t n_static, n_active, gather_k, nbr_k, n_mat, albedo_bins;
  const uint32_t *active;       /* n_active node ids */
  const uint32_t *gtable;       /* n_static * gather_k hit node ids, 0xffffffff = miss */
  const float *normals;         /* n_static * 4: node normal xyz, - */
  const uint32_t *node_mat;     /* n_static material ids */
  const float *albedo;          /* n_mat * bins * 8: per bin [spec.rgb, -][diff.rgb, -] (gpu/renderer.js:294-297) */
  const float *hh;              /* n_mat * 4: hemispherical albedo rgb, - */
  const uint32_t *nbr;          /* n_active * nbr_k active indices */
  const float *nbr_w;           /* n_active * nbr_k weights (0 ends a row) */
} vt_gather_tables;
int vt_jacobi_sparse(vt_ctx *ctx, const vt_gather_tables *t, float *node_state, uint32_t iters,
                     float *aux_out);

/* Dense node radiosity x = e + F x (the precision configuration): F n x n row-major, e n x r. */
enum { VT_DENSE_FP32 = 0, VT_DENSE_BF16_REFINE = 1, VT_DENSE_FP64 = 2 };
typedef struct {
  uint32_t n, r;
  uint32_t method;        /* VT_DENSE_* */
  double tol;             /* stop when ||e + F x - x||_F <= tol ||e||_F (FP64 check) */
  uint32_t max_sweeps;
  uint32_t inner_sweeps;  /* BF16 sweeps per FP32 refinement step (0: until the correction changes by less
                             than 2^-9 of the residual, BF16's resolution) */
} vt_dense_opts;
typedef struct {
  double residual;        /* final relative residual, FP64 */
  uint32_t sweeps;        /* Jacobi sweeps done */
  uint32_t refinements;
  double ms;              /* GPU time of the solve */
  double ms_per_sweep;
} vt_dense_result;
/* Upload F (float32, n*n) and e (float32, n*r) once; the library keeps F in FP32, BF16 and FP64 copies as the
 * methods need. */
int vt_dense_upload(vt_ctx *ctx, uint32_t n, uint32_t r, const float *F, const float *e);
/* Device-side form-factor build for nodes on the uploaded scene: F_ij = V_ij cos_i cos_j A_j / (pi r^2),
 * rows scaled to sum <= row_max; V by shadow rays through the scene's BVH. nodes: n * 8 floats (p.xyz, area,
 * n.xyz, -). Emission e (n * r) is uploaded from the caller. */
int vt_dense_formfactors(vt_ctx *ctx, uint32_t n, const float *nodes, float row_max, uint32_t r, const float *e);
int vt_dense_solve(vt_ctx *ctx, const vt_dense_opts *opts, float *x_out, vt_dense_result *res);
/* F read back (float32 n*n), for tests */
int vt_dense_download_F(vt_ctx *ctx, float *F_out, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
