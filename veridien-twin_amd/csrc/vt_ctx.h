// The library context (struct vt_ctx behind the opaque handle of include/vt.h) and the host helpers the C-ABI
// translation units share: device buffers, error latching, uploads, host poses, the relative-illumination map.
#pragma once
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/vt.h"
#include "vt_internal.h"

namespace vt {

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

// what a camera's static first hits depend on: its rays (pose, lens, size, sample pattern) and the static scene;
// the occlusion cache adds the light samples per sub-sample and the shading inputs that decide which samples light
// a point (materials, lights, profiles)
struct CacheKey {
  float R[9], pos[3];
  uint32_t mode;
  float lens[9];
  uint32_t width, height, seed, cam_index, nl, nl_refine;
  float shutter[2];     // the sub-samples' exposure times (where a light that rides on a body stands)
  uint64_t scene_gen, shade_gen;
  bool operator==(const CacheKey &o) const { return std::memcmp(this, &o, sizeof *this) == 0; }
};
static_assert(sizeof(CacheKey) == 30 * 4 + 2 * 8, "CacheKey has no padding (compared bytewise)");

struct CamSlot {
  bool set = false;
  vt_camera c{};
  DevBuf acc, dn, ri, lin, adj;
  // glare levels; ISP: linear RGB, block partial sums, 8-bit RGBA, YUV planes, statistics, rgb8 staging
  DevBuf glare_a, glare_b, isp_rgb, isp_part, rgba, yuv, stats, rgb8;
  DevBuf keys;                // the reference estimator's first-pass keys (refinement)
  // the static first-hit and static-occlusion caches of a fixed camera, valid while their keys hold, for the first
  // `layers` sub-sample indices
  DevBuf hcache, scache;
  uint32_t hc_layers = 0, sc_layers = 0;
  CacheKey hc_key{}, sc_key{};
  vt_sensor_params sp{};      // the last vt_sensor's parameters (the ISP reads black, max DN, noise model, CFA)
  bool isp_done = false, isp420 = false;
  bool ri_valid = false;
  float ri_pupil = -1.0f;
  uint32_t ri_vn = 0;                     // the vignetting table the RI map was made with
  float ri_vt[VT_MAX_VIGNETTING] = {}, ri_vv[VT_MAX_VIGNETTING] = {};
  uint32_t spp_total = 0;     // samples accumulated in acc (for the mean)
  bool rendered = false, sensed = false;
};

}  // namespace

using namespace vt;

struct vt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool lost = false;
  std::string err;
  int cu_count = 0;
  // scene
  bool has_scene = false;
  uint64_t scene_gen = 0, shade_gen = 0;   // bumped by a scene upload / a material, light or profile change
                                           // (or a new pose record of the body the light rides on)
  float4 emit_rec[4];                      // that body's pose words (q0, q1, p0, p1) at the last vt_set_instances
  bool emit_rec_set = false;
  DevBuf nodes, tri, tshade, meshes, mats, lights, emit, inst, lprof, knots, tvid;
  DevBuf uv, tex;                          // per-vertex texture coordinates (textured scenes); the print mask
  bool has_textured = false;               // a mesh of the scene is VT_MESH_TEXTURED
  // host copies the node radiosity build reads: vertices (8 floats each), per slot vertex ids + mesh, meshes,
  // the static scene's slot count (its slots come first), the emitter records
  std::vector<float> h_verts;
  std::vector<uint4> h_tvid;
  std::vector<uint32_t> h_meshes;
  std::vector<float> h_emit;
  uint32_t static_slots = 0;
  // the node radiosity cache (vt_radiosity_build / update / state)
  bool rad_built = false;
  uint32_t rad_K = 0, rad_samples = 0, rad_nbr_k = 0, rad_active = 0, rad_slot = 1;
  DevBuf g_pos, g_miss, radu, g_area;
  std::vector<uint32_t> body_root;   // BLAS root node of each body
  std::vector<float> body_box;       // each body's box in its own frame (min xyz, max xyz)
  SceneDev S{};
  uint32_t tri_count = 0, node4_count = 0;
  // the interior tree (vt_interior_bvh): its root node and slot count past the scene's own nodes and slots
  uint32_t interior_root = 0, interior_tris = 0;
  // launch support
  DevBuf counters, overflow, spill, dcams, test_in, test_out, test_aux;
  DevBuf refine_list;           // the refinement pass's pixel list (camera, pixel) and its count
  DevBuf eavg_part;             // eAvg's block partials
  uint64_t spill_lanes = 0;
  DevCam *h_cams = nullptr;     // pinned staging for the camera table
  hipEvent_t ev_stage = nullptr;
  bool stage_pending = false;
  hipEvent_t ev_r0 = nullptr, ev_r1 = nullptr, ev_s0 = nullptr, ev_s1 = nullptr, ev_i0 = nullptr, ev_i1 = nullptr;
  std::vector<hipEvent_t> ev_launch;
  uint32_t last_launches = 0;
  std::vector<hipEvent_t> ev_ext;   // wavefront: around each extension-trace launch of the last frame
  uint32_t ext_launches = 0;
  bool render_timed = false, sensor_timed = false, backward_timed = false, isp_timed = false;
  hipEvent_t ev_b0 = nullptr, ev_b1 = nullptr;
  DevBuf grads, ltot;
  uint32_t n_mat = 0, n_light = 0;
  // launch shape used for vt_render_opts fields left 0 (vt_autotune replaces it)
  uint32_t def_block = DEFAULT_BLOCK, def_stack = STACK_LDS, def_chunk = 4, def_queues = 1;
  // the tuned frame form (VT_RENDER_AUTO) and the wavefront's batch, LDS stack and refill threshold
  uint32_t def_mode = VT_RENDER_BATCHED;
  // wavefront paths per batch when the options give none: larger batches sort more rays together (more coherent
  // waves) and launch less often (full-scale frame: 2,164 Mrays/s at 2^24, 2,359 at 2^26, 2,381 at 2^27, 2,401 with
  // the whole frame in one batch); the default batch shrinks until its buffers fit wf_budget bytes
  uint64_t def_wf_batch = 1ull << 27;
  uint64_t wf_budget = 0;      // an eighth of the device's memory (set at create)
  uint32_t def_wf_stack = 8, def_wf_refill = 16, def_wf_sort = 0, def_wf_queues = 1;
  bool def_wf_shadow_sort = true;
  int occ[2][2][3][STACK_LDS_MAX + 1] = {};   // resident render blocks per CU by (bodies, estimator, block, LDS depth)
  // scene box (for the wavefront sort key)
  float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
  // wavefront buffers: paths per batch and shadow slots per path they are sized for
  DevBuf wf, wf_temp;
  uint64_t wf_paths = 0;
  uint32_t wf_nl = 0;
  size_t wf_temp_bytes = 0;
  // the frame graph: the batched frame's launches, replayed while their parameters match `gkey`
  hipGraphExec_t gexec = nullptr;
  std::vector<uint8_t> gkey;
  uint32_t graph_builds = 0;
  bool graph_used = false;
  double rays_backward = 0;
  CamSlot cams[VT_MAX_CAMERAS];
  // radiosity
  DevBuf g_active, g_table, g_normals, g_mat, g_alb, g_hh, g_nbr, g_nbrw, g_state, g_aux;
  DevBuf g_w;                 // the gather entries' static weights (null use when rad_weights is false)
  bool rad_weights = false;
  // a material, light or profile change since the node cache's last direct pass: the next vt_radiosity_update
  // relights every active node whatever its `direct` argument says (B_dir holds albedo x radiance)
  bool rad_direct_stale = false;
  // the bodies' bounding capsules at the node cache's last direct pass (2 float4 per body), which the next
  // incremental pass compares with; rad_pcaps_valid once a direct pass has run since the build
  DevBuf rad_pcaps;
  bool rad_pcaps_valid = false;
  // vt_jacobi_sparse's own copies of the caller's tables (never the node cache's buffers)
  DevBuf s_active, s_table, s_normals, s_mat, s_alb, s_hh, s_nbr, s_nbrw, s_state, s_aux;
  // the baked node table (vt_baked_*): hits, node records, albedos, capsules, transmission bytes, E_dir, E_ind,
  // and the two B buffers of the ping-pong (bk_cur holds the latest)
  DevBuf bk_gather, bk_a, bk_n, bk_hh, bk_caps, bk_occ, bk_edir, bk_eind, bk_b[2];
  uint32_t bk_nodes = 0, bk_K = 0, bk_meshes = 0, bk_ncaps = 0, bk_cur = 0;
  // dense radiosity
  DevBuf dF32, dF16, dF64, dE, dX, dY, dXd, dR, dPart;
  uint32_t dn_n = 0, dn_r = 0, dn_rp = 0;
};

namespace vt {

int fail(vt_ctx *c, int code, const char *fmt, ...);
int hip_fail(vt_ctx *c, hipError_t e, const char *what);

#define VT_HIP(c, call)                                 \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return vt::hip_fail((c), e_, #call); \
  } while (0)

#define VT_LIVE(c)                                                                       \
  do {                                                                                   \
    if (!(c)) return VT_ERR_ARG;                                                         \
    if ((c)->lost) return vt::fail((c), VT_ERR_LOST, "context lost: %s", (c)->err.c_str()); \
  } while (0)

// device buffers: grown (never shrunk) to at least `bytes`; upload = ensure + synchronous copy
int ensure(vt_ctx *c, DevBuf &b, size_t bytes);
void release(DevBuf &b);
int upload(vt_ctx *c, DevBuf &b, const void *src, size_t bytes);
float bits_f(uint32_t u);

// host pose of an instance at exposure time t, the device's formulas in FP64 (slerp, quaternion columns)
struct HostPose {
  double c[3][3];   // columns
  double p[3];
};
HostPose host_pose(const vt_instance &I, double t);
void host_apply(const HostPose &P, const double v[3], double out[3]);

// grid for the trace kernels: every CU filled to its occupancy, never more lanes than spill slots
inline uint32_t max_blocks(vt_ctx *c, uint32_t block) { return (uint32_t)(c->spill_lanes / block); }

// relative illumination per pixel (gpu/sensor.js:103-124), FP64: the KB4 term times the vignetting table
// (vn knots, field angles in degrees)
void ri_map_host(const vt_camera &cam, double q, uint32_t vn, const float *vt_deg, const float *vv,
                 std::vector<float> &out);

// the node cache's material-dependent tables (vt_radiosity_api.cpp): the directional-albedo tables and the gather
// entries' static weights, recomputed on the stream after a material change
int refresh_node_materials(vt_ctx *c);

// the frame graph (vt_frame.cpp): dropped whenever what it captured changes
void drop_graph(vt_ctx *c);
// the estimator of a frame's cameras: all path tracer, or all the reference estimator
int frame_estimator(vt_ctx *c, const uint32_t *cams, uint32_t n, bool &ref);

// sensor uniforms (vt_sensor_api.cpp), shared with the test entry points
int sensor_params(vt_ctx *c, const vt_sensor_params *p, uint32_t w, uint32_t h, uint32_t cam_index, float scale,
                  SensorDev &P);
size_t glare_bytes(uint32_t w, uint32_t h, const SensorDev &P);
int isp_params(vt_ctx *c, const vt_sensor_params &sp, const vt_isp_params *q, uint32_t w, uint32_t h, IspDev &P);

}  // namespace vt
