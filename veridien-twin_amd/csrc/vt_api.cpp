// The C ABI (include/vt.h): context, scene upload and re-layout, cameras, render/sensor/readback, the
// reference's test entry points and node radiosity. Device memory is owned by the context; one HIP stream
// (default priority) orders every call's work.
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

using namespace vt;

namespace {

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
  uint64_t scene_gen, shade_gen;
  bool operator==(const CacheKey &o) const { return std::memcmp(this, &o, sizeof *this) == 0; }
};
static_assert(sizeof(CacheKey) == 28 * 4 + 2 * 8, "CacheKey has no padding (compared bytewise)");

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
  uint32_t spp_total = 0;     // samples accumulated in acc (for the mean)
  bool rendered = false, sensed = false;
};

}  // namespace

struct vt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool lost = false;
  std::string err;
  int cu_count = 0;
  // scene
  bool has_scene = false;
  uint64_t scene_gen = 0, shade_gen = 0;   // bumped by a scene upload / a material, light or profile change
  DevBuf nodes, tri, tshade, meshes, mats, lights, emit, inst, lprof, knots, tvid;
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
  uint64_t def_wf_batch = 1ull << 24;
  uint32_t def_wf_stack = 8, def_wf_refill = 16;
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
  // dense radiosity
  DevBuf dF32, dF16, dF64, dE, dX, dY, dXd, dR, dPart;
  uint32_t dn_n = 0, dn_r = 0, dn_rp = 0;
};

namespace {

int fail(vt_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

int hip_fail(vt_ctx *c, hipError_t e, const char *what) {
  c->lost = true;
  return fail(c, VT_ERR_HIP, "%s: %s (%d); the context is lost", what, hipGetErrorString(e), (int)e);
}

#define VT_HIP(c, call)                                 \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return hip_fail((c), e_, #call); \
  } while (0)

#define VT_LIVE(c)                                                                       \
  do {                                                                                   \
    if (!(c)) return VT_ERR_ARG;                                                         \
    if ((c)->lost) return fail((c), VT_ERR_LOST, "context lost: %s", (c)->err.c_str()); \
  } while (0)

int ensure(vt_ctx *c, DevBuf &b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return VT_OK;
  if (b.p) {
    hipError_t e = hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (e != hipSuccess) return hip_fail(c, e, "hipFree");
  }
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      return fail(c, VT_ERR_NOMEM, "device allocation of %zu bytes failed", bytes);
    }
    return hip_fail(c, e, "hipMalloc");
  }
  b.bytes = bytes;
  return VT_OK;
}
void release(DevBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

int upload(vt_ctx *c, DevBuf &b, const void *src, size_t bytes) {
  int rc = ensure(c, b, bytes);
  if (rc) return rc;
  if (bytes) VT_HIP(c, hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
  return VT_OK;
}

float bits_f(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// host pose of an instance at exposure time t, the device's formulas in FP64 (slerp, quaternion columns)
struct HostPose {
  double c[3][3];   // columns
  double p[3];
};
HostPose host_pose(const vt_instance &I, double t) {
  double a[4], b[4];
  for (int k = 0; k < 4; k++) { a[k] = I.q0[k]; b[k] = I.q1[k]; }
  double c = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
  if (c < 0) {
    for (double &x : b) x = -x;
    c = -c;
  }
  double q[4];
  if (c > 0.9995) {
    double l = 0;
    for (int k = 0; k < 4; k++) { q[k] = a[k] + (b[k] - a[k]) * t; l += q[k] * q[k]; }
    for (double &x : q) x /= std::sqrt(l);
  } else {
    const double th = std::acos(c), sn = std::sin(th);
    for (int k = 0; k < 4; k++) q[k] = (a[k] * std::sin((1 - t) * th) + b[k] * std::sin(t * th)) / sn;
  }
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  HostPose P;
  const double m[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w)},
                          {2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w)},
                          {2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)}};
  std::memcpy(P.c, m, sizeof m);
  for (int k = 0; k < 3; k++) P.p[k] = I.p0[k] + (I.p1[k] - I.p0[k]) * t;
  return P;
}
void host_apply(const HostPose &P, const double v[3], double out[3]) {
  for (int k = 0; k < 3; k++) out[k] = P.c[0][k] * v[0] + P.c[1][k] * v[1] + P.c[2][k] * v[2] + P.p[k];
}

// grid for the trace kernels: every CU filled to its occupancy, never more lanes than spill slots
uint32_t max_blocks(vt_ctx *c, uint32_t block) { return (uint32_t)(c->spill_lanes / block); }

// relative illumination per pixel (gpu/sensor.js:103-124), FP64, default vignetting (1 at every angle)
void ri_map_host(const vt_camera &cam, double q, std::vector<float> &out) {
  const double fx = cam.lens[0], fy = cam.lens[1], cx = cam.lens[2], cy = cam.lens[3], al = cam.lens[8];
  const double k[4] = {cam.lens[4], cam.lens[5], cam.lens[6], cam.lens[7]};
  auto dist = [&](double th) {
    double t2 = th * th;
    return th * (1 + t2 * (k[0] + t2 * (k[1] + t2 * (k[2] + t2 * k[3]))));
  };
  auto deriv = [&](double th) {
    double t2 = th * th;
    return 1 + t2 * (3 * k[0] + t2 * (5 * k[1] + t2 * (7 * k[2] + t2 * 9 * k[3])));
  };
  const uint32_t w = cam.width, h = cam.height;
  out.resize((size_t)w * h);
  for (uint32_t y = 0; y < h; y++)
    for (uint32_t x = 0; x < w; x++) {
      double ray_z;
      if (cam.mode == 0) {
        double yd = (y - cy) / fy, xd = (x - cx) / fx - al * yd;
        double thd = std::hypot(xd, yd);
        if (thd < 1e-300) {
          ray_z = 1;
        } else {
          double th = thd;
          for (int i = 0; i < 20; i++) {
            double step = (dist(th) - thd) / deriv(th);
            th -= step;
            if (std::fabs(step) < 1e-15) break;
          }
          ray_z = std::cos(th);
        }
      } else {
        double dx = (x - cx) / fx, dy = (y - cy) / fy;
        ray_z = 1 / std::sqrt(dx * dx + dy * dy + 1);
      }
      double th = std::acos(std::max(-1.0, std::min(1.0, ray_z)));
      double ri = 1;
      if (th > 1e-6) {
        if (cam.mode == 0) ri = std::pow(std::cos(th), q) * std::sin(th) / (dist(th) * deriv(th));
        else ri = std::pow(std::cos(th), q + 3);   // pinhole mapping: cos^3 * cos^q
      }
      out[(size_t)y * w + x] = (float)std::max(0.0, ri);
    }
}

}  // namespace

// ====================================================================== lifetime
extern "C" int vt_create(const vt_create_info *info, vt_ctx **out) {
  if (!out) return VT_ERR_ARG;
  *out = nullptr;
  vt_ctx *c = new vt_ctx();
  c->device = info ? info->device : 0;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= c->device || c->device < 0) {
    delete c;
    return VT_ERR_HIP;
  }
  e = hipSetDevice(c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  hipDeviceProp_t prop;
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, c->device);
  if (e != hipSuccess) {
    delete c;
    return VT_ERR_HIP;
  }
  c->cu_count = prop.multiProcessorCount;
  int occ = 0;
  // persistent grids: up to 8 resident 256-thread blocks per CU (LDS and VGPRs usually allow fewer)
  c->spill_lanes = (uint64_t)c->cu_count * 8 * 256;
  (void)occ;
  int rc = ensure(c, c->counters, 64);
  if (!rc) rc = ensure(c, c->overflow, 64);
  if (!rc) rc = ensure(c, c->spill, c->spill_lanes * SPILL_ENTRIES * sizeof(uint2));
  if (!rc) rc = ensure(c, c->dcams, sizeof(DevCam) * VT_MAX_CAMERAS);
  if (!rc && hipHostMalloc((void **)&c->h_cams, sizeof(DevCam) * VT_MAX_CAMERAS, hipHostMallocDefault) != hipSuccess) rc = VT_ERR_HIP;
  hipEvent_t *evs[] = {&c->ev_stage, &c->ev_r0, &c->ev_r1, &c->ev_s0, &c->ev_s1, &c->ev_b0, &c->ev_b1, &c->ev_i0,
                       &c->ev_i1};
  for (hipEvent_t *ev : evs)
    if (!rc && hipEventCreate(ev) != hipSuccess) rc = VT_ERR_HIP;
  if (!rc && hipMemsetAsync(c->overflow.p, 0, 64, c->stream) != hipSuccess) rc = VT_ERR_HIP;
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = VT_ERR_HIP;
  if (rc) {
    vt_destroy(c);
    return rc;
  }
  *out = c;
  return VT_OK;
}

extern "C" void vt_destroy(vt_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  DevBuf *bufs[] = {&c->nodes, &c->tri, &c->tshade, &c->meshes, &c->mats, &c->lights, &c->emit, &c->inst, &c->lprof, &c->knots, &c->tvid, &c->g_pos, &c->g_miss, &c->radu, &c->g_area,
                    &c->counters, &c->overflow, &c->spill, &c->dcams, &c->test_in, &c->test_out, &c->test_aux,
                    &c->refine_list, &c->eavg_part,
                    &c->g_active, &c->g_table, &c->g_normals, &c->g_mat, &c->g_alb, &c->g_hh, &c->g_nbr, &c->g_w,
                    &c->g_nbrw, &c->g_state, &c->g_aux, &c->dF32, &c->dF16, &c->dF64, &c->dE, &c->dX, &c->dY,
                    &c->dXd, &c->dR, &c->dPart, &c->grads, &c->wf, &c->wf_temp, &c->ltot};
  for (DevBuf *b : bufs) release(*b);
  for (auto &cs : c->cams) {
    release(cs.acc);
    release(cs.dn);
    release(cs.ri);
    release(cs.lin);
    release(cs.adj);
    DevBuf *isp[] = {&cs.glare_a, &cs.glare_b, &cs.isp_rgb, &cs.isp_part, &cs.rgba, &cs.yuv, &cs.stats, &cs.rgb8,
                     &cs.keys, &cs.hcache, &cs.scache};
    for (DevBuf *b : isp) release(*b);
  }
  for (hipEvent_t ev : c->ev_launch) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : c->ev_ext) (void)hipEventDestroy(ev);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  hipEvent_t evs[] = {c->ev_stage, c->ev_r0, c->ev_r1, c->ev_s0, c->ev_s1, c->ev_b0, c->ev_b1, c->ev_i0, c->ev_i1};
  for (hipEvent_t ev : evs)
    if (ev) (void)hipEventDestroy(ev);
  if (c->h_cams) (void)hipHostFree(c->h_cams);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char *vt_last_error(const vt_ctx *c) { return c ? c->err.c_str() : "null context"; }

// ====================================================================== BVH
extern "C" int vt_build_bvh(const float *pos, uint32_t nvert, const uint32_t *idx, uint32_t ntri, int leaf_size,
                            int bins, float *nodes_out, uint32_t cap_nodes, uint32_t *node_count,
                            uint32_t *order_out) {
  if ((!pos && nvert) || (!idx && ntri) || !nodes_out || !node_count || (!order_out && ntri)) return VT_ERR_ARG;
  if (leaf_size < 1 || bins < 2 || bins > 256) return VT_ERR_ARG;
  for (uint64_t k = 0; k < 3ull * ntri; k++)
    if (idx[k] >= nvert) return VT_ERR_ARG;
  int64_t n = build_bvh2(pos, idx, ntri, leaf_size, bins, nodes_out, cap_nodes, order_out);
  if (n < 0) return VT_ERR_SIZE;
  *node_count = (uint32_t)n;
  return VT_OK;
}

// ====================================================================== scene
extern "C" int vt_upload_scene(vt_ctx *c, const vt_scene *s) {
  VT_LIVE(c);
  if (!s || !s->tris || !s->verts || !s->meshes || !s->materials) return fail(c, VT_ERR_ARG, "scene: null array");
  if (s->tri_count == 0 || s->vert_count == 0 || s->mesh_count == 0 || s->material_count == 0)
    return fail(c, VT_ERR_ARG, "scene: empty triangle, vertex, mesh or material table");
  if (s->emit_count && (!s->emit || !s->lights || s->light_count == 0))
    return fail(c, VT_ERR_ARG, "scene: emitter triangles without lights");
  if (s->tri_count >= (1u << 26)) return fail(c, VT_ERR_ARG, "scene: more than 2^26 triangles");
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  const uint32_t T = s->tri_count, V = s->vert_count;
  const uint32_t *tw = reinterpret_cast<const uint32_t *>(s->tris);
  // validate the triangle records
  for (uint32_t k = 0; k < T; k++) {
    const uint32_t *r = tw + (size_t)k * 16;
    if (r[3] >= V || r[7] >= V || r[11] >= V) return fail(c, VT_ERR_ARG, "scene: triangle %u vertex id out of range", k);
    if (r[12] >= s->mesh_count) return fail(c, VT_ERR_ARG, "scene: triangle %u mesh id out of range", k);
  }
  for (uint32_t m = 0; m < s->mesh_count; m++) {
    if (s->meshes[4 * m] >= s->material_count) return fail(c, VT_ERR_ARG, "scene: mesh %u material out of range", m);
    if ((s->meshes[4 * m + 1] & VT_MESH_EMITTER) && s->meshes[4 * m + 2] >= std::max(1u, s->light_count))
      return fail(c, VT_ERR_ARG, "scene: mesh %u light out of range", m);
  }
  for (uint32_t e = 0; e < s->emit_count; e++) {
    uint32_t lid;
    std::memcpy(&lid, s->emit + 16 * (size_t)e + 11, 4);
    if (lid >= s->light_count) return fail(c, VT_ERR_ARG, "scene: emitter %u light out of range", e);
  }
  if (s->body_count > VT_MAX_BODIES) return fail(c, VT_ERR_ARG, "scene: more than %d bodies", VT_MAX_BODIES);
  if (s->body_count && (!s->tri_body || s->nodes))
    return fail(c, VT_ERR_ARG, "scene: bodies need tri_body and no caller BVH (nodes = NULL)");
  for (uint32_t k = 0; s->body_count && k < T; k++)
    if (s->tri_body[k] != 0xffffffffu) {
      if (s->tri_body[k] >= s->body_count) return fail(c, VT_ERR_ARG, "scene: triangle %u body out of range", k);
      // emitters are sampled from the world-space emitter table, so they stay in the static scene
      if (s->meshes[4 * (size_t)tw[16 * (size_t)k + 12] + 1] & VT_MESH_EMITTER)
        return fail(c, VT_ERR_ARG, "scene: triangle %u of a body is an emitter", k);
    }
  // BVH: the caller's (reference layout) or built here; slot order is the tris order
  std::vector<float> nodes2;
  std::vector<float> tris_sorted;
  const float *tr = s->tris;
  uint32_t n2 = s->node_count;
  const float *np = s->nodes;
  std::vector<Node4> n4;
  std::vector<uint32_t> body_root(s->body_count, 0);
  uint32_t body_slots = 0, static_count = 0;
  std::vector<float> body_box(6 * (size_t)s->body_count, 0.0f);
  if (s->body_count) {
    // one BVH per group (the static scene, then each body) over that group's triangles, built and collapsed
    // separately and concatenated: node refs offset by the group's first node, leaf slots by its first slot
    std::vector<float> pos((size_t)V * 3);
    for (uint32_t v = 0; v < V; v++)
      for (int k = 0; k < 3; k++) pos[3 * v + k] = s->verts[8 * (size_t)v + k];
    tris_sorted.resize((size_t)T * 16);
    uint32_t slot_base = 0;
    for (uint32_t g = 0; g <= s->body_count; g++) {
      const uint32_t want = g == 0 ? 0xffffffffu : g - 1;
      std::vector<uint32_t> members;
      for (uint32_t k = 0; k < T; k++)
        if (s->tri_body[k] == want) members.push_back(k);
      const uint32_t Tg = (uint32_t)members.size();
      if (g == 0) static_count = Tg;
      if (Tg == 0) return fail(c, VT_ERR_ARG, g ? "scene: body %u has no triangles" : "scene: no static triangles", g - 1);
      std::vector<uint32_t> idx((size_t)Tg * 3), order(Tg);
      for (uint32_t m = 0; m < Tg; m++)
        for (int j = 0; j < 3; j++) idx[3 * m + j] = tw[16 * (size_t)members[m] + 4 * j + 3];
      std::vector<float> ng2((size_t)(Tg + 2) * 16);
      const int64_t nn = build_bvh2(pos.data(), idx.data(), Tg, 4, 16, ng2.data(), Tg + 2, order.data());
      if (nn < 0) return fail(c, VT_ERR_ARG, "scene: BVH build failed");
      std::vector<Node4> g4;
      try {
        g4 = collapse_bvh4(ng2.data(), (uint32_t)nn, Tg);
      } catch (const std::exception &ex) {
        return fail(c, VT_ERR_ARG, "scene: %s", ex.what());
      }
      const uint32_t node_base = (uint32_t)n4.size();
      for (Node4 &nd : g4) {
        for (int ch = 0; ch < 4; ch++) {
          const uint32_t r = nd.ref[ch];
          if (r == EMPTY_REF) continue;
          if (r & LEAF_BIT) nd.ref[ch] = (r & ~0x03ffffffu) | ((r & 0x03ffffffu) + slot_base);
          else nd.ref[ch] = r + node_base;
        }
        n4.push_back(nd);
      }
      for (uint32_t m = 0; m < Tg; m++)
        std::memcpy(&tris_sorted[16 * (size_t)(slot_base + m)], s->tris + 16 * (size_t)members[order[m]], 64);
      if (g > 0) {
        body_root[g - 1] = node_base;
        float *bx = &body_box[6 * (size_t)(g - 1)];
        for (int k = 0; k < 3; k++) { bx[k] = INFINITY; bx[3 + k] = -INFINITY; }
        for (uint32_t m = 0; m < Tg; m++)
          for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
              const float x = pos[3 * (size_t)idx[3 * m + j] + k];
              bx[k] = std::min(bx[k], x);
              bx[3 + k] = std::max(bx[3 + k], x);
            }
      }
      slot_base += Tg;
    }
    tr = tris_sorted.data();
    body_slots = T - static_count;
  } else if (!np) {
    std::vector<float> pos((size_t)V * 3);
    for (uint32_t v = 0; v < V; v++)
      for (int k = 0; k < 3; k++) pos[3 * v + k] = s->verts[8 * (size_t)v + k];
    std::vector<uint32_t> idx((size_t)T * 3), order(T);
    for (uint32_t k = 0; k < T; k++)
      for (int j = 0; j < 3; j++) idx[3 * k + j] = tw[16 * (size_t)k + 4 * j + 3];
    nodes2.resize((size_t)(T + 2) * 16);
    int64_t n = build_bvh2(pos.data(), idx.data(), T, 4, 16, nodes2.data(), T + 2, order.data());
    if (n < 0) return fail(c, VT_ERR_ARG, "scene: BVH build failed");
    n2 = (uint32_t)n;
    np = nodes2.data();
    tris_sorted.resize((size_t)T * 16);
    for (uint32_t k = 0; k < T; k++) std::memcpy(&tris_sorted[16 * (size_t)k], s->tris + 16 * (size_t)order[k], 64);
    tr = tris_sorted.data();
  } else if (n2 == 0) {
    return fail(c, VT_ERR_ARG, "scene: node_count is 0");
  }
  if (!s->body_count) {
    try {
      n4 = collapse_bvh4(np, n2, T);
    } catch (const std::exception &ex) {
      return fail(c, VT_ERR_ARG, "scene: %s", ex.what());
    }
  }
  const uint32_t *trw = reinterpret_cast<const uint32_t *>(tr);
  std::vector<float4> tri((size_t)T * 3), vn(V);
  std::vector<uint4> attr(T);
  for (uint32_t k = 0; k < T; k++) {
    const uint32_t *r = trw + (size_t)k * 16;
    for (int j = 0; j < 3; j++) {
      const float *p = s->verts + 8 * (size_t)r[4 * j + 3];
      tri[3 * (size_t)k + j] = make_float4(p[0], p[1], p[2], 0.0f);
    }
    attr[k] = make_uint4(r[3], r[7], r[11], r[12]);
  }
  for (uint32_t v = 0; v < V; v++) {
    const float *p = s->verts + 8 * (size_t)v;
    vn[v] = make_float4(p[4], p[5], p[6], 0.0f);
  }
  int rc;
  if ((rc = upload(c, c->nodes, n4.data(), n4.size() * sizeof(Node4)))) return rc;
  if ((rc = upload(c, c->tri, tri.data(), tri.size() * sizeof(float4)))) return rc;
  {
    // per-slot vertex ids stay on the device (the node cache's lookups); vertex normals only feed the shading
    // records
    DevBuf d_vn;
    rc = upload(c, c->tvid, attr.data(), attr.size() * sizeof(uint4));
    if (!rc) rc = upload(c, d_vn, vn.data(), vn.size() * sizeof(float4));
    if (!rc) rc = ensure(c, c->tshade, (size_t)T * 64);
    if (!rc) {
      hipError_t e = launch_tri_shading((const float4 *)c->tri.p, (const uint4 *)c->tvid.p, (const float4 *)d_vn.p, T,
                                        (float4 *)c->tshade.p, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) rc = hip_fail(c, e, "triangle shading records");
    }
    release(d_vn);
    if (rc) return rc;
  }
  c->h_verts.assign(s->verts, s->verts + 8 * (size_t)V);
  c->h_tvid = attr;
  c->h_meshes.assign(s->meshes, s->meshes + 4 * (size_t)s->mesh_count);
  c->h_emit.assign(s->emit_count ? s->emit : nullptr, s->emit_count ? s->emit + 16 * (size_t)s->emit_count : nullptr);
  c->static_slots = s->body_count ? (uint32_t)(T - body_slots) : T;
  c->rad_built = false;
  c->S.nstate = nullptr;
  c->S.alb = nullptr;
  c->S.hh = nullptr;
  c->S.radu = nullptr;
  if ((rc = upload(c, c->meshes, s->meshes, (size_t)s->mesh_count * 16))) return rc;
  if ((rc = upload(c, c->mats, s->materials, (size_t)s->material_count * 64))) return rc;
  float zero4[4] = {0, 0, 0, 0};
  if ((rc = upload(c, c->lights, s->light_count ? (const void *)s->lights : zero4, std::max<size_t>(16, (size_t)s->light_count * 16)))) return rc;
  float zero16[16] = {0};
  if ((rc = upload(c, c->emit, s->emit_count ? (const void *)s->emit : zero16, std::max<size_t>(64, (size_t)s->emit_count * 64)))) return rc;
  c->S.nodes = (const Node4 *)c->nodes.p;
  c->S.tri = (const float4 *)c->tri.p;
  c->S.tshade = (const float4 *)c->tshade.p;
  c->S.meshes = (const uint4 *)c->meshes.p;
  c->S.tvid = (const uint4 *)c->tvid.p;
  c->S.mats = (const float4 *)c->mats.p;
  c->S.lights = (const float4 *)c->lights.p;
  {
    // every light Lambertian until vt_set_light_profile gives it a profile
    const size_t nl = std::max<uint32_t>(1, s->light_count);
    std::vector<float4> lp(nl, make_float4(0.f, 0.f, 1.f, 0.f));
    std::vector<float2> kn(nl * MAX_KNOTS, make_float2(0.f, 0.f));
    if ((rc = upload(c, c->lprof, lp.data(), lp.size() * sizeof(float4)))) return rc;
    if ((rc = upload(c, c->knots, kn.data(), kn.size() * sizeof(float2)))) return rc;
    c->S.lprof = (const float4 *)c->lprof.p;
    c->S.knots = (const float2 *)c->knots.p;
  }
  c->S.emit = (const float4 *)c->emit.p;
  c->S.emit_count = s->emit_count;
  c->S.root = 0;
  c->body_root = body_root;
  c->body_box = body_box;
  c->S.inst_count = s->body_count;
  c->S.shadow_skip = s->body_count ? s->shadow_skip : 0u;
  if (s->body_count) {
    // every body starts hidden until vt_set_instances gives it a pose
    std::vector<float4> rec((size_t)INST_WORDS * s->body_count, make_float4(0.f, 0.f, 0.f, 0.f));
    for (uint32_t b = 0; b < s->body_count; b++) {
      rec[INST_WORDS * b + 2].w = bits_f(body_root[b]);
      rec[INST_WORDS * b + 3].w = bits_f(INST_HIDDEN);
    }
    if ((rc = upload(c, c->inst, rec.data(), rec.size() * sizeof(float4)))) return rc;
    c->S.inst = (const float4 *)c->inst.p;
  } else {
    c->S.inst = nullptr;
  }
  c->S.emit_area = s->emit_area > 0 ? s->emit_area : 1e-12f;
  for (int k = 0; k < 3; k++) {
    c->bmin[k] = INFINITY;
    c->bmax[k] = -INFINITY;
  }
  for (uint32_t v = 0; v < V; v++)
    for (int k = 0; k < 3; k++) {
      c->bmin[k] = std::min(c->bmin[k], s->verts[8 * (size_t)v + k]);
      c->bmax[k] = std::max(c->bmax[k], s->verts[8 * (size_t)v + k]);
    }
  c->tri_count = T;
  c->node4_count = (uint32_t)n4.size();
  c->interior_root = 0;
  c->interior_tris = 0;
  c->n_mat = s->material_count;
  c->n_light = s->light_count;
  c->has_scene = true;
  c->scene_gen++;
  VT_HIP(c, hipMemsetAsync(c->overflow.p, 0, 64, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  return VT_OK;
}

extern "C" int vt_update_materials(vt_ctx *c, const float *materials, uint32_t count) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "update_materials: no scene uploaded");
  if (!materials || count != c->n_mat) return fail(c, VT_ERR_SIZE, "update_materials: %u materials expected", c->n_mat);
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(c->mats.p, materials, (size_t)count * 64, hipMemcpyHostToDevice));
  c->shade_gen++;
  return VT_OK;
}

extern "C" int vt_update_lights(vt_ctx *c, const float *lights, uint32_t count) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "update_lights: no scene uploaded");
  if (!lights || count != c->n_light || count == 0) return fail(c, VT_ERR_SIZE, "update_lights: %u lights expected", c->n_light);
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(c->lights.p, lights, (size_t)count * 16, hipMemcpyHostToDevice));
  c->shade_gen++;
  return VT_OK;
}

static_assert(VT_MAX_KNOTS == MAX_KNOTS, "knot table stride");

extern "C" int vt_set_light_profile(vt_ctx *c, uint32_t light, const float axis[3], const float *theta, const float *p,
                                    uint32_t count) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "light_profile: no scene uploaded");
  if (light >= c->n_light) return fail(c, VT_ERR_ARG, "light_profile: light %u of %u", light, c->n_light);
  if (count > VT_MAX_KNOTS) return fail(c, VT_ERR_SIZE, "light_profile: at most %d knots", VT_MAX_KNOTS);
  float4 rec = make_float4(0.f, 0.f, 1.f, 0.f);
  std::vector<float2> kn(MAX_KNOTS, make_float2(0.f, 0.f));
  if (count) {
    if (!axis || !theta || !p) return fail(c, VT_ERR_ARG, "light_profile: null axis or knots");
    const double l = std::sqrt((double)axis[0] * axis[0] + (double)axis[1] * axis[1] + (double)axis[2] * axis[2]);
    if (!(l > 1e-12) || !std::isfinite(l)) return fail(c, VT_ERR_ARG, "light_profile: zero axis");
    for (uint32_t k = 0; k < count; k++) {
      if (!std::isfinite(theta[k]) || !std::isfinite(p[k]) || p[k] < 0.0f || (k && theta[k] < theta[k - 1]))
        return fail(c, VT_ERR_ARG, "light_profile: knots need finite ascending theta and p >= 0 (knot %u)", k);
      kn[k] = make_float2(theta[k], p[k]);
    }
    rec = make_float4((float)(axis[0] / l), (float)(axis[1] / l), (float)(axis[2] / l), bits_f(count));
  }
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy((float4 *)c->lprof.p + light, &rec, sizeof rec, hipMemcpyHostToDevice));
  VT_HIP(c, hipMemcpy((float2 *)c->knots.p + (size_t)light * MAX_KNOTS, kn.data(), kn.size() * sizeof(float2),
                      hipMemcpyHostToDevice));
  c->shade_gen++;
  return VT_OK;
}

extern "C" int vt_set_camera(vt_ctx *c, uint32_t cam, const vt_camera *k) {
  VT_LIVE(c);
  if (cam >= VT_MAX_CAMERAS || !k) return fail(c, VT_ERR_ARG, "camera slot %u out of range", cam);
  if (k->width == 0 || k->height == 0 || k->width > 16384 || k->height > 16384)
    return fail(c, VT_ERR_ARG, "camera %u: bad size %ux%u", cam, k->width, k->height);
  if (k->spp == 0) return fail(c, VT_ERR_ARG, "camera %u: spp is 0", cam);
  if (k->mode > 1) return fail(c, VT_ERR_ARG, "camera %u: mode %u is not kb4 (0) or pinhole (1)", cam, k->mode);
  if (!(k->lens[0] > 0) || !(k->lens[1] > 0)) return fail(c, VT_ERR_ARG, "camera %u: focal lengths must be > 0", cam);
  if (!(k->shutter[0] >= 0.0f && k->shutter[0] <= 1.0f && k->shutter[1] >= 0.0f && k->shutter[1] <= 1.0f))
    return fail(c, VT_ERR_ARG, "camera %u: shutter times must lie in [0, 1]", cam);
  if (k->refine > 64 || k->refine_light_samples > 16)
    return fail(c, VT_ERR_ARG, "camera %u: refine %u (<= 64), refine light samples %u (<= 16)", cam, k->refine,
                k->refine_light_samples);
  VT_HIP(c, hipSetDevice(c->device));
  CamSlot &s = c->cams[cam];
  const size_t px = (size_t)k->width * k->height;
  const bool resized = !s.set || s.c.width != k->width || s.c.height != k->height;
  const bool lens_changed = resized || std::memcmp(s.c.lens, k->lens, sizeof k->lens) != 0 || s.c.mode != k->mode;
  int rc;
  if ((rc = ensure(c, s.acc, px * sizeof(float4)))) return rc;
  if ((rc = ensure(c, s.dn, px * sizeof(uint16_t)))) return rc;
  if ((rc = ensure(c, s.ri, px * sizeof(float)))) return rc;
  if ((rc = ensure(c, s.lin, px * sizeof(float4)))) return rc;
  if (lens_changed) s.ri_valid = false;
  s.c = *k;
  s.set = true;
  if (resized) {
    s.rendered = false;
    s.sensed = false;
    s.spp_total = 0;
  }
  return VT_OK;
}

extern "C" int vt_set_instances(vt_ctx *c, const vt_instance *in, uint32_t count) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "set_instances: no scene uploaded");
  const uint32_t B = (uint32_t)c->body_root.size();
  if (count != B || (B && !in)) return fail(c, VT_ERR_SIZE, "set_instances: %u bodies in the scene", B);
  if (!B) return VT_OK;
  std::vector<float4> rec((size_t)INST_WORDS * B);
  for (uint32_t b = 0; b < B; b++) {
    vt_instance I = in[b];
    for (float *q : {I.q0, I.q1}) {
      const double l = std::sqrt((double)q[0] * q[0] + (double)q[1] * q[1] + (double)q[2] * q[2] + (double)q[3] * q[3]);
      if (!(l > 1e-12) || !std::isfinite(l)) return fail(c, VT_ERR_ARG, "set_instances: body %u has no unit quaternion", b);
      for (int k = 0; k < 4; k++) q[k] = (float)(q[k] / l);
    }
    bool moving = false;
    for (int k = 0; k < 4; k++) moving |= I.q0[k] != I.q1[k];
    for (int k = 0; k < 3; k++) moving |= I.p0[k] != I.p1[k];
    const uint32_t flags = (moving ? INST_MOVING : 0u) | (I.hidden ? INST_HIDDEN : 0u);
    float4 *r = &rec[INST_WORDS * (size_t)b];
    r[0] = make_float4(I.q0[0], I.q0[1], I.q0[2], I.q0[3]);
    r[1] = make_float4(I.q1[0], I.q1[1], I.q1[2], I.q1[3]);
    r[2] = make_float4(I.p0[0], I.p0[1], I.p0[2], bits_f(c->body_root[b]));
    r[3] = make_float4(I.p1[0], I.p1[1], I.p1[2], bits_f(flags));
    // Culling bounds that hold at every exposure time. Between shutter open and close a body point y (body frame)
    // moves on R(t) y + p(t): R(t) turns about one axis by at most phi (the angle between q0 and q1), p(t) is
    // linear. The world box is the box of the body's corners posed at t = k/8, padded by the largest gap between
    // a circular arc of angle phi/8 and its chord, rad (1 - cos(phi/16)), rad = |y| bound over the box.
    const float *bx = &c->body_box[6 * (size_t)b];
    double ctr[3], half[3];
    for (int k = 0; k < 3; k++) { ctr[k] = 0.5 * (bx[k] + bx[3 + k]); half[k] = 0.5 * (bx[3 + k] - bx[k]); }
    const double halfdiag = std::sqrt(half[0] * half[0] + half[1] * half[1] + half[2] * half[2]);
    const double rad = std::sqrt(ctr[0] * ctr[0] + ctr[1] * ctr[1] + ctr[2] * ctr[2]) + halfdiag;
    double qd = 0, dp = 0;
    for (int k = 0; k < 4; k++) qd += (double)I.q0[k] * I.q1[k];
    for (int k = 0; k < 3; k++) dp += ((double)I.p1[k] - I.p0[k]) * ((double)I.p1[k] - I.p0[k]);
    const double phi = moving ? 2.0 * std::acos(std::min(1.0, std::fabs(qd))) : 0.0;
    dp = std::sqrt(dp);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const int nt = moving ? 9 : 1;
    for (int ti = 0; ti < nt; ti++) {
      const HostPose P = host_pose(I, moving ? ti / 8.0 : 0.0);
      for (int cn = 0; cn < 8; cn++) {
        const double v[3] = {cn & 1 ? bx[3] : bx[0], cn & 2 ? bx[4] : bx[1], cn & 4 ? bx[5] : bx[2]};
        double w[3];
        host_apply(P, v, w);
        for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], w[k]); hi[k] = std::max(hi[k], w[k]); }
      }
    }
    const double scale = std::max({std::fabs(lo[0]), std::fabs(lo[1]), std::fabs(lo[2]), std::fabs(hi[0]),
                                   std::fabs(hi[1]), std::fabs(hi[2]), 1e-3});
    const double pad = 1.01 * rad * (1.0 - std::cos(phi / 16.0)) + 4e-6 * scale + 1e-6;
    r[4] = make_float4((float)(lo[0] - pad), (float)(lo[1] - pad), (float)(lo[2] - pad), 0.f);
    r[5] = make_float4((float)(hi[0] + pad), (float)(hi[1] + pad), (float)(hi[2] + pad), 0.f);
    // Bounding capsule of the box along its longest axis, posed at mid-exposure; every point stays within
    // 2 sin(phi / 4) rad + |p1 - p0| / 2 of its mid-exposure place, which pads the radius.
    const int ax = half[0] >= half[1] ? (half[0] >= half[2] ? 0 : 2) : (half[1] >= half[2] ? 1 : 2);
    double ea[3], eb[3];
    for (int k = 0; k < 3; k++) { ea[k] = ctr[k]; eb[k] = ctr[k]; }
    ea[ax] -= half[ax];
    eb[ax] += half[ax];
    double rr = 0;
    for (int k = 0; k < 3; k++)
      if (k != ax) rr += half[k] * half[k];
    rr = std::sqrt(rr);
    const HostPose Pm = host_pose(I, moving ? 0.5 : 0.0);
    double wa[3], wb[3];
    host_apply(Pm, ea, wa);
    host_apply(Pm, eb, wb);
    const double motion = moving ? 1.01 * (2.0 * std::sin(phi / 4.0) * rad + 0.5 * dp) : 0.0;
    r[6] = make_float4((float)wa[0], (float)wa[1], (float)wa[2], (float)(rr * 1.0001 + motion + 4e-6 * scale + 1e-6));
    r[7] = make_float4((float)wb[0], (float)wb[1], (float)wb[2], 0.f);
    // occlusion capsule of the reference estimator's indirect light, posed at mid-exposure
    if (!(I.occ_r >= 0.0f) || !std::isfinite(I.occ_r) || !(I.rho_eff >= 0.0f && I.rho_eff <= 1.0f))
      return fail(c, VT_ERR_ARG, "set_instances: body %u needs occ_r >= 0 and rho_eff in [0, 1]", b);
    const double oa[3] = {I.occ_a[0], I.occ_a[1], I.occ_a[2]}, ob[3] = {I.occ_b[0], I.occ_b[1], I.occ_b[2]};
    double woa[3], wob[3];
    host_apply(Pm, oa, woa);
    host_apply(Pm, ob, wob);
    r[8] = make_float4((float)woa[0], (float)woa[1], (float)woa[2], I.occ_r);
    r[9] = make_float4((float)wob[0], (float)wob[1], (float)wob[2], I.rho_eff);
  }
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(c->inst.p, rec.data(), rec.size() * sizeof(float4), hipMemcpyHostToDevice));
  return VT_OK;
}

// ====================================================================== render
namespace {

struct Shape {
  uint32_t block, stack, chunk, queues;
};

int check_shape(vt_ctx *c, const Shape &sh) {
  if (sh.block != 64 && sh.block != 128 && sh.block != 256) return fail(c, VT_ERR_ARG, "render: block must be 64, 128 or 256");
  if (sh.stack < (uint32_t)STACK_LDS_MIN || sh.stack > (uint32_t)STACK_LDS_MAX)
    return fail(c, VT_ERR_ARG, "render: stack_lds must be %d..%d", STACK_LDS_MIN, STACK_LDS_MAX);
  if (sh.queues != 1 && sh.queues != (uint32_t)XCD_COUNT) return fail(c, VT_ERR_ARG, "render: queues must be 1 or %d", XCD_COUNT);
  if (sh.chunk == 0 || sh.chunk > 4096) return fail(c, VT_ERR_ARG, "render: chunk must be 1..4096");
  return VT_OK;
}

Shape shape_of(const vt_ctx *c, const vt_render_opts &o) {
  return Shape{o.block ? o.block : c->def_block, o.stack_lds ? o.stack_lds : c->def_stack,
               o.chunk ? o.chunk : c->def_chunk, o.queues ? o.queues : c->def_queues};
}

// blocks of a persistent grid: every CU filled to the kernel's occupancy, within the spill slots
int persistent_grid(vt_ctx *c, const Shape &sh, uint64_t waves, uint32_t &grid, bool ref = false) {
  const int bi = sh.block == 64 ? 0 : (sh.block == 128 ? 1 : 2);
  const bool inst = c->S.inst_count != 0;
  int &occ = c->occ[inst][ref][bi][sh.stack];
  if (occ == 0) {
    VT_HIP(c, render_occupancy(sh.block, sh.stack, inst, ref, &occ));
    if (occ < 1) occ = 1;
  }
  const uint64_t need = (waves * 64 + sh.block - 1) / sh.block;
  uint64_t g = std::min<uint64_t>((uint64_t)occ * c->cu_count, max_blocks(c, sh.block));
  grid = (uint32_t)std::max<uint64_t>(1, std::min(need, g));
  return VT_OK;
}

// a fixed camera's first-hit (VT_OPT_HIT_CACHE) and occlusion (VT_OPT_SHADOW_CACHE) caches for a whole frame of
// d.nspp + d.refine sub-samples: (re)allocated and cleared to empty by a fill kernel on the frame's stream when the
// key changes or more layers are needed, else kept from the last frame
int camera_caches(vt_ctx *c, CamSlot &s, DevCam &d) {
  const vt_camera &k = s.c;
  CacheKey key;
  std::memset(&key, 0, sizeof key);
  std::memcpy(key.R, k.R, sizeof key.R);
  std::memcpy(key.pos, k.pos, sizeof key.pos);
  key.mode = k.mode;
  std::memcpy(key.lens, k.lens, sizeof key.lens);
  key.width = k.width;
  key.height = k.height;
  key.seed = k.seed;
  key.cam_index = k.cam_index;
  key.scene_gen = c->scene_gen;
  const uint32_t layers = d.nspp + d.refine;
  const uint64_t px = (uint64_t)k.width * k.height;
  if (k.flags & VT_OPT_HIT_CACHE) {
    if (!(key == s.hc_key) || layers > s.hc_layers) {
      const uint64_t n = px * layers;
      if (int rc = ensure(c, s.hcache, n * sizeof(float4))) return rc;
      VT_HIP(c, launch_fill_u4((uint4 *)s.hcache.p, n, make_uint4(0u, 0u, 0u, 0u), c->stream));
      s.hc_key = key;
      s.hc_layers = layers;
    }
    d.hcache = (float4 *)s.hcache.p;
  }
  if (k.flags & VT_OPT_SHADOW_CACHE) {
    key.nl = d.nl;
    key.nl_refine = d.nl_refine;
    key.shade_gen = c->shade_gen;
    if (!(key == s.sc_key) || layers > s.sc_layers) {
      const uint64_t n = (px * layers + 3) / 4;     // 32-bit words, cleared 16 bytes at a time
      if (int rc = ensure(c, s.scache, n * sizeof(uint4))) return rc;
      VT_HIP(c, launch_fill_u4((uint4 *)s.scache.p, n, make_uint4(0u, 0u, 0u, 0u), c->stream));
      s.sc_key = key;
      s.sc_layers = layers;
    }
    d.scache = (uint32_t *)s.scache.p;
  }
  return VT_OK;
}

// the camera table of a frame into the pinned staging area; returns the frame's queue items
int stage_frame(vt_ctx *c, const uint32_t *cams, uint32_t n, const vt_render_opts &opt, uint64_t &items) {
  if (!cams || n == 0 || n > VT_MAX_CAMERAS) return fail(c, VT_ERR_ARG, "render: 1..%d cameras", VT_MAX_CAMERAS);
  for (uint32_t a = 0; a < n; a++) {
    if (cams[a] >= VT_MAX_CAMERAS || !c->cams[cams[a]].set) return fail(c, VT_ERR_STATE, "render: camera %u not set", cams[a]);
    for (uint32_t b = 0; b < a; b++)
      if (cams[a] == cams[b]) return fail(c, VT_ERR_ARG, "render: camera %u listed twice", cams[a]);
    const uint32_t s0 = opt.spp_begin, s1 = opt.spp_end ? opt.spp_end : c->cams[cams[a]].c.spp;
    if (s1 <= s0) return fail(c, VT_ERR_ARG, "render: empty sub-sample range [%u, %u)", s0, s1);
  }
  VT_HIP(c, hipSetDevice(c->device));
  // the staging table may still feed an earlier copy
  if (c->stage_pending) VT_HIP(c, hipEventSynchronize(c->ev_stage));
  items = 0;
  for (uint32_t a = 0; a < n; a++) {
    CamSlot &s = c->cams[cams[a]];
    const vt_camera &k = s.c;
    DevCam &d = c->h_cams[a];
    std::memset(&d, 0, sizeof d);
    std::memcpy(d.R, k.R, sizeof d.R);
    std::memcpy(d.pos, k.pos, sizeof d.pos);
    d.mode = k.mode;
    std::memcpy(d.lens, k.lens, sizeof d.lens);
    d.width = k.width; d.height = k.height; d.spp = k.spp; d.seed = k.seed;
    d.cam_index = k.cam_index;
    d.nl = k.light_samples ? k.light_samples : 1;
    d.depth = k.max_depth ? k.max_depth : 1;
    d.tiles_x = (k.width + 7) / 8;
    d.tiles_y = (k.height + 7) / 8;
    const uint32_t s0 = opt.spp_begin, s1 = opt.spp_end ? opt.spp_end : k.spp;
    d.nspp = s1 - s0;
    d.spp_begin = s0;
    d.accumulate = opt.accumulate;
    d.item_begin = items;
    d.item_count = (uint64_t)d.tiles_x * d.tiles_y * 64ull * d.nspp;
    d.acc = (float4 *)s.acc.p;
    d.adj = nullptr;
    d.scale = k.scale > 0 ? k.scale : 1.0f;
    d.shutter0 = k.shutter[0];
    d.shutter1 = k.shutter[1];
    d.skip = k.skip_bodies;
    d.opts = k.flags;
    d.refine = (k.flags & VT_OPT_REFERENCE) ? k.refine : 0u;
    d.nl_refine = k.refine_light_samples ? k.refine_light_samples : d.nl;
    d.keys = nullptr;
    if (d.refine) {
      if (s0 != 0 || opt.accumulate)
        return fail(c, VT_ERR_ARG, "render: the refinement pass needs a whole frame (spp_begin 0, no accumulate)");
      int rc = ensure(c, s.keys, (size_t)k.width * k.height * sizeof(uint4));
      if (rc) return rc;
      d.keys = (uint4 *)s.keys.p;
    }
    if ((k.flags & VT_OPT_REFERENCE) && (k.flags & (VT_OPT_HIT_CACHE | VT_OPT_SHADOW_CACHE)) && s0 == 0 &&
        !opt.accumulate) {
      int rc = camera_caches(c, s, d);
      if (rc) return rc;
    }
    items += d.item_count;
  }
  VT_HIP(c, hipMemcpyAsync(c->dcams.p, c->h_cams, sizeof(DevCam) * n, hipMemcpyHostToDevice, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_stage, c->stream));
  c->stage_pending = true;
  return VT_OK;
}

RenderLaunch base_launch(vt_ctx *c, const Shape &sh) {
  RenderLaunch L;
  std::memset(&L, 0, sizeof L);     // the launch record is also the frame graph's key: no stray padding bytes
  L.S = c->S;
  L.counters = (unsigned long long *)c->counters.p;
  L.overflow = (uint32_t *)c->overflow.p;
  L.spill = (uint2 *)c->spill.p;
  L.spill_stride = c->spill_lanes;
  L.chunk = sh.chunk;
  L.block = sh.block;
  L.stack_lds = sh.stack;
  L.queues = sh.queues;
  L.wave_stride = 1;
  return L;
}

// the batched frame's stream work after staging: clear the queue heads and counters, clear the sums, one
// persistent launch
int enqueue_batched(vt_ctx *c, uint32_t n, bool accumulate, const RenderLaunch &L) {
  VT_HIP(c, launch_zero_u64((unsigned long long *)c->counters.p, 16, c->stream));
  if (!accumulate)
    for (uint32_t a = 0; a < n; a++) {
      const DevCam &d = c->h_cams[a];
      VT_HIP(c, launch_zero_f4(d.acc, (uint64_t)d.width * d.height, c->stream));
    }
  VT_HIP(c, launch_render(L, c->stream));
  return VT_OK;
}

int ensure_launch_events(vt_ctx *c, uint32_t k) {
  while (c->ev_launch.size() < 2 * (size_t)k) {
    hipEvent_t ev;
    VT_HIP(c, hipEventCreate(&ev));
    c->ev_launch.push_back(ev);
  }
  return VT_OK;
}

void drop_graph(vt_ctx *c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
  c->gkey.clear();
}

// capture enqueue_batched as a graph (fill kernels only: no memset or copy nodes)
int build_graph(vt_ctx *c, uint32_t n, bool accumulate, const RenderLaunch &L) {
  drop_graph(c);
  VT_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_batched(c, n, accumulate, L);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
  const hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    c->gexec = nullptr;
    return hip_fail(c, ei, "hipGraphInstantiate");
  }
  c->graph_builds++;
  return VT_OK;
}

struct WfParams {
  uint64_t batch;     // paths per batch
  uint32_t slds;      // LDS stack entries of the trace kernels (4, 8, 16)
  uint32_t refill;    // idle lanes that make a wave fetch rays
  uint64_t stride;    // 1, or > 1: trace every stride-th wave of the frame (tuning probes)
};

WfParams wf_params(const vt_ctx *c, const vt_render_opts &opt) {
  return WfParams{opt.batch ? opt.batch : c->def_wf_batch, opt.stack_lds ? opt.stack_lds : c->def_wf_stack,
                  opt.chunk ? std::min(64u, opt.chunk) : c->def_wf_refill, 1};
}

// the wavefront frame: buffers for one batch, the batches in order, one event pair around all of them
int render_wavefront(vt_ctx *c, uint32_t n, const vt_render_opts &opt, const WfParams &wp, uint64_t frame_items) {
  const uint64_t items = (frame_items / 64 + wp.stride - 1) / wp.stride * 64;
  uint64_t P = wp.batch;
  P = std::min<uint64_t>((P + 63) / 64 * 64, std::max<uint64_t>(items, 64));
  if (P > (1ull << 31)) return fail(c, VT_ERR_ARG, "render: wavefront batch of %llu paths", (unsigned long long)P);
  uint32_t nl = 1, depth = 1;
  for (uint32_t a = 0; a < n; a++) {
    nl = std::max(nl, std::min(16u, std::max(1u, c->h_cams[a].nl)));
    depth = std::max(depth, c->h_cams[a].depth);
  }
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  // per path 84 B of state, per queue slot 36 B, per shadow slot 48 B
  const size_t bytes = al(P * 4) + 5 * al(P * 16) + 4 * al(P * 4) + al(P * 16) + al(P * 4) + 3 * al(P * nl * 16) + 256;
  int rc;
  if (P > c->wf_paths || nl > c->wf_nl) {
    if ((rc = ensure(c, c->wf, bytes))) return rc;
    const size_t tb = wf_sort_temp_bytes((uint32_t)P);
    if ((rc = ensure(c, c->wf_temp, tb))) return rc;
    c->wf_temp_bytes = tb;
    c->wf_paths = P;
    c->wf_nl = nl;
  }
  char *base = (char *)c->wf.p;
  size_t off = 0;
  auto take = [&](size_t b) { void *p = base + off; off += al(b); return p; };
  WfLaunch W;
  std::memset(&W, 0, sizeof W);
  W.S = c->S;
  W.cams = (const DevCam *)c->dcams.p;
  W.ncam = n;
  W.nl_max = nl;
  W.cam = (uint32_t *)take(P * 4);
  W.o = (float4 *)take(P * 16);
  W.d = (float4 *)take(P * 16);
  W.thr = (float4 *)take(P * 16);
  W.thr_slot = (float4 *)take(P * 16);
  W.L = (float4 *)take(P * 16);
  W.q_next = (uint32_t *)take(P * 4);
  W.key_next = (uint32_t *)take(P * 4);
  uint32_t *q_sorted = (uint32_t *)take(P * 4);
  uint32_t *keys_sorted = (uint32_t *)take(P * 4);
  W.hit = (float4 *)take(P * 16);
  W.nee = (uint32_t *)take(P * 4);
  W.sh_o = (float4 *)take(P * nl * 16);
  W.sh_d = (float4 *)take(P * nl * 16);
  W.sh_c = (float4 *)take(P * nl * 16);
  W.count = (uint32_t *)take(16);
  float ext[3];
  for (int k = 0; k < 3; k++) ext[k] = std::max(c->bmax[k] - c->bmin[k], 1e-6f);
  W.bmin = make_float3(c->bmin[0], c->bmin[1], c->bmin[2]);
  W.bscale = make_float3(128.0f / ext[0], 128.0f / ext[1], 128.0f / ext[2]);
  W.counters = (unsigned long long *)c->counters.p;
  W.overflow = (uint32_t *)c->overflow.p;
  W.spill = (uint2 *)c->spill.p;
  W.spill_stride = c->spill_lanes;
  VT_HIP(c, hipEventRecord(c->ev_r0, c->stream));
  VT_HIP(c, launch_zero_u64((unsigned long long *)c->counters.p, 16, c->stream));
  if (!opt.accumulate)
    for (uint32_t a = 0; a < n; a++) {
      const DevCam &d = c->h_cams[a];
      VT_HIP(c, launch_zero_f4(d.acc, (uint64_t)d.width * d.height, c->stream));
    }
  VT_HIP(c, hipEventRecord(c->ev_launch[0], c->stream));
  const uint32_t slds = wp.slds;
  if (slds != 4 && slds != 8 && slds != 16) return fail(c, VT_ERR_ARG, "render: wavefront stack_lds must be 4, 8 or 16");
  W.wave_stride = wp.stride;
  int occ = 0;
  VT_HIP(c, wf_trace_occupancy(slds, c->S.inst_count != 0, &occ));
  const uint32_t grid_cap = (uint32_t)c->cu_count * 8;
  const uint32_t trace_cap = std::min<uint32_t>(max_blocks(c, 256), (uint32_t)std::max(1, occ) * c->cu_count);
  const uint32_t refill = wp.refill;
  const uint64_t batches = (items + P - 1) / P;
  while (c->ev_ext.size() < 2 * batches * depth) {
    hipEvent_t ev;
    VT_HIP(c, hipEventCreate(&ev));
    c->ev_ext.push_back(ev);
  }
  uint64_t bi = 0;
  for (uint64_t bb = 0; bb < items; bb += P, bi++) {
    W.batch_begin = bb;
    W.n = (uint32_t)std::min<uint64_t>(P, items - bb);
    VT_HIP(c, launch_wavefront_batch(W, depth, grid_cap, trace_cap, refill, slds, keys_sorted, q_sorted,
                                     c->wf_temp.p, c->wf_temp_bytes, c->ev_ext.data() + 2 * depth * bi, c->stream));
  }
  c->ext_launches = (uint32_t)(batches * depth);
  VT_HIP(c, hipEventRecord(c->ev_launch[1], c->stream));
  return VT_OK;
}

}  // namespace

// the estimator of a frame's cameras: all path tracer, or all the reference estimator (which needs the node cache)
int frame_estimator(vt_ctx *c, const uint32_t *cams, uint32_t n, bool &ref) {
  ref = false;
  if (!cams || n == 0) return VT_OK;
  for (uint32_t a = 0; a < n; a++) {
    if (cams[a] >= VT_MAX_CAMERAS || !c->cams[cams[a]].set) return VT_OK;   // stage_frame reports it
    const bool r = (c->cams[cams[a]].c.flags & VT_OPT_REFERENCE) != 0;
    if (a && r != ref) return fail(c, VT_ERR_ARG, "render: cameras of one frame use one estimator");
    ref = r;
  }
  if (ref && !c->rad_built) return fail(c, VT_ERR_STATE, "render: the reference estimator needs vt_radiosity_build");
  return VT_OK;
}

extern "C" int vt_render(vt_ctx *c, const uint32_t *cams, uint32_t n, const vt_render_opts *o) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "render: no scene uploaded");
  vt_render_opts opt{};
  if (o) opt = *o;
  bool ref = false;
  int rc = frame_estimator(c, cams, n, ref);
  if (rc) return rc;
  if (ref && opt.mode == VT_RENDER_WAVEFRONT)
    return fail(c, VT_ERR_ARG, "render: the wavefront form runs the path tracer; the reference estimator runs persistent");
  if (opt.mode == VT_RENDER_AUTO) opt.mode = ref ? (uint32_t)VT_RENDER_BATCHED : c->def_mode;
  const Shape sh = shape_of(c, opt);
  rc = opt.mode == VT_RENDER_WAVEFRONT ? VT_OK : check_shape(c, sh);   // the wavefront checks its own
  if (rc) return rc;
  if (opt.graph && opt.mode == VT_RENDER_WAVEFRONT) opt.graph = 0;   // the wavefront form is not captured
  if (opt.graph && opt.mode != VT_RENDER_BATCHED) return fail(c, VT_ERR_ARG, "render: graph needs the batched mode");
  if (opt.mode > VT_RENDER_AUTO) return fail(c, VT_ERR_ARG, "render: mode %u", opt.mode);
  uint64_t items = 0;
  if ((rc = stage_frame(c, cams, n, opt, items))) return rc;
  RenderLaunch L = base_launch(c, sh);
  L.ref = ref ? 1u : 0u;
  uint32_t launches = 0;
  c->graph_used = false;
  c->ext_launches = 0;
  // the refinement pass's first-pass keys start cleared (a fill kernel, before the frame's launches or graph)
  for (uint32_t a = 0; a < n; a++) {
    const DevCam &d = c->h_cams[a];
    if (d.keys) VT_HIP(c, launch_fill_keys(d.keys, (uint64_t)d.width * d.height, c->stream));
  }
  if (opt.mode == VT_RENDER_WAVEFRONT) {
    if ((rc = ensure_launch_events(c, 1))) return rc;
    if ((rc = render_wavefront(c, n, opt, wf_params(c, opt), items))) return rc;
    launches = 1;
  } else if (opt.mode == VT_RENDER_BATCHED) {
    if ((rc = ensure_launch_events(c, 1))) return rc;
    L.cams = (const DevCam *)c->dcams.p;
    L.ncam = n;
    L.item_base = 0;
    L.total_waves = items / 64;
    L.persistent = 1;
    if ((rc = persistent_grid(c, sh, L.total_waves, L.grid, ref))) return rc;
    if (opt.graph) {
      // key: the launch record, and the sums the graph clears
      std::vector<uint8_t> key(sizeof L + 1 + n * sizeof(DevCam));
      std::memcpy(key.data(), &L, sizeof L);
      key[sizeof L] = opt.accumulate ? 1 : 0;
      for (uint32_t a = 0; a < n; a++) {
        const DevCam &d = c->h_cams[a];
        uint64_t v[2] = {(uint64_t)(uintptr_t)d.acc, (uint64_t)d.width * d.height};
        std::memcpy(key.data() + sizeof L + 1 + a * sizeof(DevCam), v, sizeof v);
      }
      if (!c->gexec || key != c->gkey) {
        if ((rc = build_graph(c, n, opt.accumulate != 0, L))) return rc;
        c->gkey = key;
      }
      VT_HIP(c, hipEventRecord(c->ev_r0, c->stream));
      VT_HIP(c, hipEventRecord(c->ev_launch[0], c->stream));
      VT_HIP(c, hipGraphLaunch(c->gexec, c->stream));
      VT_HIP(c, hipEventRecord(c->ev_launch[1], c->stream));
      c->graph_used = true;
    } else {
      VT_HIP(c, hipEventRecord(c->ev_r0, c->stream));
      VT_HIP(c, launch_zero_u64((unsigned long long *)c->counters.p, 16, c->stream));
      if (!opt.accumulate)
        for (uint32_t a = 0; a < n; a++) {
          const DevCam &d = c->h_cams[a];
          VT_HIP(c, launch_zero_f4(d.acc, (uint64_t)d.width * d.height, c->stream));
        }
      VT_HIP(c, hipEventRecord(c->ev_launch[0], c->stream));
      VT_HIP(c, launch_render(L, c->stream));
      VT_HIP(c, hipEventRecord(c->ev_launch[1], c->stream));
    }
    launches = 1;
  } else {
    if ((rc = ensure_launch_events(c, n))) return rc;
    VT_HIP(c, hipEventRecord(c->ev_r0, c->stream));
    VT_HIP(c, launch_zero_u64((unsigned long long *)c->counters.p, 16, c->stream));
    if (!opt.accumulate)
      for (uint32_t a = 0; a < n; a++) {
        const DevCam &d = c->h_cams[a];
        VT_HIP(c, launch_zero_f4(d.acc, (uint64_t)d.width * d.height, c->stream));
      }
    const uint32_t gmax = max_blocks(c, sh.block);
    for (uint32_t a = 0; a < n; a++) {
      const DevCam &d = c->h_cams[a];
      L.cams = (const DevCam *)c->dcams.p + a;
      L.ncam = 1;
      L.item_base = d.item_begin;
      L.total_waves = d.item_count / 64;
      L.persistent = 0;
      L.queues = 1;
      const uint64_t need = (L.total_waves * 64 + sh.block - 1) / sh.block;
      L.grid = (uint32_t)std::min<uint64_t>(need, gmax);
      VT_HIP(c, hipEventRecord(c->ev_launch[2 * a], c->stream));
      VT_HIP(c, launch_render(L, c->stream));
      VT_HIP(c, hipEventRecord(c->ev_launch[2 * a + 1], c->stream));
      launches++;
    }
  }
  // the reference estimator's refinement pass (camRefine): the cameras that ask for it list their pixels, then one
  // launch refines them all
  {
    uint32_t R = 0;
    uint64_t px = 0;
    for (uint32_t a = 0; a < n; a++) {
      const DevCam &d = c->h_cams[a];
      if (d.refine && d.keys) {
        R = std::max(R, d.refine);
        px += (uint64_t)d.width * d.height;
      }
    }
    if (R) {
      if ((rc = ensure(c, c->refine_list, px * sizeof(uint2) + 256))) return rc;
      uint32_t *count = (uint32_t *)((char *)c->refine_list.p + px * sizeof(uint2));
      VT_HIP(c, launch_zero_u64((unsigned long long *)count, 1, c->stream));
      for (uint32_t a = 0; a < n; a++) {
        const DevCam &d = c->h_cams[a];
        if (d.refine && d.keys)
          VT_HIP(c, launch_refine_mark((const DevCam *)c->dcams.p + a, a, d.width, d.height,
                                       (uint2 *)c->refine_list.p, count, c->stream));
      }
      RenderLaunch Lr = L;
      Lr.cams = (const DevCam *)c->dcams.p;
      Lr.ncam = n;
      VT_HIP(c, launch_refine(Lr, (const uint2 *)c->refine_list.p, count, px, R, c->stream));
    }
  }
  VT_HIP(c, hipEventRecord(c->ev_r1, c->stream));
  c->last_launches = launches;
  c->render_timed = true;
  for (uint32_t a = 0; a < n; a++) {
    CamSlot &s = c->cams[cams[a]];
    if (!opt.accumulate) s.spp_total = 0;
    s.spp_total += c->h_cams[a].nspp;
    s.rendered = true;
  }
  return VT_OK;
}

extern "C" int vt_autotune(vt_ctx *c, const uint32_t *cams, uint32_t n, const vt_render_opts *o, uint32_t probe_stride,
                           vt_tune_result *out) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "autotune: no scene uploaded");
  if (!out) return fail(c, VT_ERR_ARG, "autotune: null result");
  vt_render_opts opt{};
  if (o) opt = *o;
  opt.accumulate = 0;
  const uint64_t stride = probe_stride ? probe_stride : 32;
  bool ref = false;
  int rc = frame_estimator(c, cams, n, ref);
  if (rc) return rc;
  uint64_t items = 0;
  if ((rc = stage_frame(c, cams, n, opt, items))) return rc;
  const uint64_t frame_waves = items / 64;
  const uint64_t probe_waves = (frame_waves + stride - 1) / stride;
  if ((rc = ensure_launch_events(c, 1))) return rc;
  auto probe = [&](const Shape &sh, double &mrays) -> int {
    RenderLaunch L = base_launch(c, sh);
    L.cams = (const DevCam *)c->dcams.p;
    L.ncam = n;
    L.total_waves = probe_waves;
    L.wave_stride = stride;
    L.persistent = 1;
    L.ref = ref ? 1u : 0u;
    int r = persistent_grid(c, sh, probe_waves, L.grid, ref);
    if (r) return r;
    VT_HIP(c, launch_zero_u64((unsigned long long *)c->counters.p, 16, c->stream));
    VT_HIP(c, hipEventRecord(c->ev_launch[0], c->stream));
    VT_HIP(c, launch_render(L, c->stream));
    VT_HIP(c, hipEventRecord(c->ev_launch[1], c->stream));
    VT_HIP(c, hipEventSynchronize(c->ev_launch[1]));
    float ms = 0;
    VT_HIP(c, hipEventElapsedTime(&ms, c->ev_launch[0], c->ev_launch[1]));
    unsigned long long cnt[2];
    VT_HIP(c, hipMemcpy(cnt, c->counters.p, sizeof cnt, hipMemcpyDeviceToHost));
    mrays = ms > 0 ? (double)cnt[1] / (ms * 1e-3) / 1e6 : 0.0;
    return VT_OK;
  };
  const Shape def{(uint32_t)DEFAULT_BLOCK, (uint32_t)STACK_LDS, 4u, 1u};
  double m_def = 0, warm = 0;
  if ((rc = probe(def, warm))) return rc;          // first touch of the scene and the caches
  if ((rc = probe(def, m_def))) return rc;
  Shape best = def;
  double m_best = m_def;
  uint32_t tried = 1;
  const uint32_t blocks[] = {64, 128, 256}, stacks[] = {8, 16, 24}, chunks[] = {1, 4, 16}, queues[] = {1, (uint32_t)XCD_COUNT};
  for (uint32_t b : blocks)
    for (uint32_t st : stacks)
      for (uint32_t ch : chunks)
        for (uint32_t q : queues) {
          const Shape sh{b, st, ch, q};
          if (sh.block == def.block && sh.stack == def.stack && sh.chunk == def.chunk && sh.queues == def.queues) continue;
          double m = 0;
          if ((rc = probe(sh, m))) return rc;
          tried++;
          if (m > m_best) {
            m_best = m;
            best = sh;
          }
        }
  // the winner and the default once more, so a single noisy probe does not decide
  double m2 = 0, d2 = 0;
  if ((rc = probe(best, m2))) return rc;
  if ((rc = probe(def, d2))) return rc;
  m_best = m2;
  m_def = d2;
  if (m_best < m_def) {
    best = def;
    m_best = m_def;
  }
  c->def_block = best.block;
  c->def_stack = best.stack;
  c->def_chunk = best.chunk;
  c->def_queues = best.queues;
  // the wavefront form on the same probe: batch {4M, 16M} paths x LDS stack {4, 8, 16} x refill {8, 16, 32};
  // the reference estimator runs the persistent form only
  if (ref) {
    drop_graph(c);
    for (uint32_t a = 0; a < n; a++) {
      CamSlot &s = c->cams[cams[a]];
      VT_HIP(c, launch_zero_f4((float4 *)s.acc.p, (uint64_t)s.c.width * s.c.height, c->stream));
      s.spp_total = 0;
      s.rendered = false;
    }
    VT_HIP(c, hipStreamSynchronize(c->stream));
    out->block = best.block; out->stack_lds = best.stack; out->chunk = best.chunk; out->queues = best.queues;
    out->candidates = tried; out->probe_mrays_best = m_best; out->probe_mrays_default = m_def;
    out->probe_fraction = (double)probe_waves / (double)std::max<uint64_t>(1, frame_waves);
    out->mode = VT_RENDER_BATCHED;
    out->wf_batch = (uint32_t)c->def_wf_batch; out->wf_stack_lds = c->def_wf_stack; out->wf_refill = c->def_wf_refill;
    out->probe_mrays_wavefront = 0.0;
    return VT_OK;
  }
  WfParams wbest{c->def_wf_batch, c->def_wf_stack, c->def_wf_refill, stride};
  double m_wf = 0;
  auto probe_wf = [&](const WfParams &wp, double &mrays) -> int {
    vt_render_opts po = opt;
    po.mode = VT_RENDER_WAVEFRONT;
    po.accumulate = 0;
    int r = render_wavefront(c, n, po, wp, items);
    if (r) return r;
    VT_HIP(c, hipEventSynchronize(c->ev_launch[1]));
    float ms = 0;
    VT_HIP(c, hipEventElapsedTime(&ms, c->ev_launch[0], c->ev_launch[1]));
    unsigned long long cnt[2];
    VT_HIP(c, hipMemcpy(cnt, c->counters.p, sizeof cnt, hipMemcpyDeviceToHost));
    mrays = ms > 0 ? (double)cnt[1] / (ms * 1e-3) / 1e6 : 0.0;
    return VT_OK;
  };
  {
    double warm = 0;
    if ((rc = probe_wf(wbest, warm))) return rc;
  }
  const uint64_t batches[] = {1ull << 22, 1ull << 24};
  const uint32_t wstacks[] = {4, 8, 16}, refills[] = {8, 16, 32};
  for (uint64_t bsz : batches)
    for (uint32_t st : wstacks)
      for (uint32_t rf : refills) {
        const WfParams wp{bsz, st, rf, stride};
        double m = 0;
        if ((rc = probe_wf(wp, m))) return rc;
        tried++;
        if (m > m_wf) {
          m_wf = m;
          wbest = wp;
        }
      }
  {
    double m2w = 0;
    if ((rc = probe_wf(wbest, m2w))) return rc;
    m_wf = m2w;
  }
  c->def_wf_batch = wbest.batch;
  c->def_wf_stack = wbest.slds;
  c->def_wf_refill = wbest.refill;
  c->def_mode = m_wf > m_best ? (uint32_t)VT_RENDER_WAVEFRONT : (uint32_t)VT_RENDER_BATCHED;
  drop_graph(c);
  // the probes left partial sums behind: clear the listed cameras
  for (uint32_t a = 0; a < n; a++) {
    CamSlot &s = c->cams[cams[a]];
    VT_HIP(c, launch_zero_f4((float4 *)s.acc.p, (uint64_t)s.c.width * s.c.height, c->stream));
    s.spp_total = 0;
    s.rendered = false;
  }
  VT_HIP(c, hipStreamSynchronize(c->stream));
  out->block = best.block;
  out->stack_lds = best.stack;
  out->chunk = best.chunk;
  out->queues = best.queues;
  out->candidates = tried;
  out->probe_mrays_best = m_best;
  out->probe_mrays_default = m_def;
  out->mode = c->def_mode;
  out->wf_batch = (uint32_t)wbest.batch;
  out->wf_stack_lds = wbest.slds;
  out->wf_refill = wbest.refill;
  out->probe_mrays_wavefront = m_wf;
  out->probe_fraction = (double)probe_waves / (double)std::max<uint64_t>(1, frame_waves);
  return VT_OK;
}

// ====================================================================== sensor
static int sensor_params(vt_ctx *c, const vt_sensor_params *p, uint32_t w, uint32_t h, uint32_t cam_index,
                         float scale, SensorDev &P) {
  if (!p) return fail(c, VT_ERR_ARG, "sensor: null parameters");
  if (p->cfa > 3) return fail(c, VT_ERR_ARG, "sensor: cfa code %u", p->cfa);
  if (!(p->esf_sigma >= 0) || p->esf_sigma > 3.8f) return fail(c, VT_ERR_ARG, "sensor: esf_sigma outside [0, 3.8] px");
  P.w = w; P.h = h; P.cfa = p->cfa; P.frame = p->frame; P.seed = p->seed; P.cam_index = cam_index;
  P.optics_k = p->optics_k; P.esf_sigma = p->esf_sigma; P.exposure = p->exposure; P.gain = p->gain;
  std::memcpy(P.m, p->m, sizeof P.m);
  P.read_noise = p->read_noise; P.full_well = p->full_well; P.dark = p->dark; P.conv_gain = p->conv_gain;
  P.max_dn = p->max_dn; P.black = p->black; P.noise_on = p->noise_on ? 1u : 0u;
  P.scale = scale;
  P.glare_k = p->glare_fraction;
  P.glare_sigma = p->glare_sigma;
  P.glare_f = p->glare_level;
  if (!(p->glare_fraction >= 0.0f && p->glare_fraction <= 1.0f))
    return fail(c, VT_ERR_ARG, "sensor: glare fraction outside [0, 1]");
  if (p->glare_fraction > 0.0f && (p->glare_level < 1 || p->glare_level > 64))
    return fail(c, VT_ERR_ARG, "sensor: glare level factor %u outside 1..64", p->glare_level);
  return VT_OK;
}
static size_t glare_bytes(uint32_t w, uint32_t h, const SensorDev &P) {
  if (!(P.glare_k > 0.0f)) return 0;
  const size_t gw = (w + P.glare_f - 1) / P.glare_f, gh = (h + P.glare_f - 1) / P.glare_f;
  return gw * gh * sizeof(float4);
}

// the ISP's uniform words from the raw stage's parameters and the ISP parameters (gpu/sensor.js:169-176)
static int isp_params(vt_ctx *c, const vt_sensor_params &sp, const vt_isp_params *q, uint32_t w, uint32_t h,
                      IspDev &P) {
  if (!q) return fail(c, VT_ERR_ARG, "isp: null parameters");
  if (q->demosaic > 2) return fail(c, VT_ERR_ARG, "isp: demosaic mode %u", q->demosaic);
  if (q->tone > 3) return fail(c, VT_ERR_ARG, "isp: tone mode %u", q->tone);
  if (!(q->sharpen_amount >= 0.0f)) return fail(c, VT_ERR_ARG, "isp: negative sharpening amount");
  P.w = w; P.h = h; P.cfa = sp.cfa; P.demosaic = q->demosaic;
  P.nr = (q->nr && q->nr_sigma_px > 0.0f) ? 1u : 0u;
  P.tone = q->tone; P.chroma420 = q->chroma420 ? 1u : 0u;
  P.max_dn = sp.max_dn; P.black = sp.black; P.read_noise = sp.read_noise; P.conv_gain = sp.conv_gain;
  P.gain = sp.gain;
  std::memcpy(P.wb, q->wb, sizeof P.wb);
  std::memcpy(P.ccm, q->ccm, sizeof P.ccm);
  P.nr_sigma = q->nr_sigma_px; P.nr_scale = q->nr_range_scale; P.gamma = q->gamma; P.s_c = q->s_curve_c;
  P.black_off = q->black_offset; P.sharpen_amount = q->sharpen_amount;
  const float rr = std::nearbyint(q->sharpen_radius);
  P.sharpen_rad = rr != rr ? 1 : (int)std::min(4.0f, std::max(1.0f, rr));
  return VT_OK;
}
static int ensure_isp(vt_ctx *c, CamSlot &s, uint32_t w, uint32_t h) {
  const size_t px = (size_t)w * h, cpx = (size_t)((w + 1) / 2) * ((h + 1) / 2);
  int rc = ensure(c, s.isp_rgb, px * 16);
  if (!rc) rc = ensure(c, s.isp_part, (size_t)isp_partials(w, h) * 16);
  if (!rc) rc = ensure(c, s.rgba, px * 4);
  if (!rc) rc = ensure(c, s.yuv, px + 2 * cpx);
  if (!rc) rc = ensure(c, s.stats, 16);
  return rc;
}

extern "C" int vt_sensor(vt_ctx *c, uint32_t cam, const vt_sensor_params *p) {
  VT_LIVE(c);
  if (cam >= VT_MAX_CAMERAS || !c->cams[cam].set) return fail(c, VT_ERR_STATE, "sensor: camera %u not set", cam);
  CamSlot &s = c->cams[cam];
  if (!s.rendered || s.spp_total == 0) return fail(c, VT_ERR_STATE, "sensor: camera %u has no radiance", cam);
  VT_HIP(c, hipSetDevice(c->device));
  SensorDev P;
  const float inv = s.c.scale > 0 ? s.c.scale : 1.0f;     // over each pixel's own sample count
  int rc = sensor_params(c, p, s.c.width, s.c.height, s.c.cam_index, inv, P);
  if (rc) return rc;
  if (!s.ri_valid || s.ri_pupil != p->pupil_exp) {
    std::vector<float> ri;
    ri_map_host(s.c, p->pupil_exp, ri);
    VT_HIP(c, hipMemcpy(s.ri.p, ri.data(), ri.size() * 4, hipMemcpyHostToDevice));
    s.ri_valid = true;
    s.ri_pupil = p->pupil_exp;
  }
  if (const size_t gb = glare_bytes(s.c.width, s.c.height, P)) {
    if ((rc = ensure(c, s.glare_a, gb)) || (rc = ensure(c, s.glare_b, gb))) return rc;
  }
  VT_HIP(c, hipEventRecord(c->ev_s0, c->stream));
  VT_HIP(c, launch_sensor(P, (const float4 *)s.acc.p, (const float *)s.ri.p, (float4 *)s.glare_a.p,
                          (float4 *)s.glare_b.p, (uint16_t *)s.dn.p, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_s1, c->stream));
  c->sensor_timed = true;
  s.sensed = true;
  s.isp_done = false;
  s.sp = *p;
  return VT_OK;
}

extern "C" int vt_isp(vt_ctx *c, uint32_t cam, const vt_isp_params *q) {
  VT_LIVE(c);
  if (cam >= VT_MAX_CAMERAS || !c->cams[cam].set) return fail(c, VT_ERR_STATE, "isp: camera %u not set", cam);
  CamSlot &s = c->cams[cam];
  if (!s.sensed) return fail(c, VT_ERR_STATE, "isp: camera %u has no raw DN (vt_sensor first)", cam);
  VT_HIP(c, hipSetDevice(c->device));
  IspDev P;
  int rc = isp_params(c, s.sp, q, s.c.width, s.c.height, P);
  if (!rc) rc = ensure_isp(c, s, s.c.width, s.c.height);
  if (rc) return rc;
  VT_HIP(c, hipEventRecord(c->ev_i0, c->stream));
  VT_HIP(c, launch_isp(P, (const uint16_t *)s.dn.p, (float4 *)s.isp_rgb.p, (uint4 *)s.isp_part.p,
                       (uchar4 *)s.rgba.p, (uint8_t *)s.yuv.p, (unsigned int *)s.stats.p, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_i1, c->stream));
  c->isp_timed = true;
  s.isp_done = true;
  s.isp420 = P.chroma420 != 0;
  return VT_OK;
}

// ====================================================================== readback
extern "C" int vt_readback(vt_ctx *c, uint32_t cam, int kind, void *dst, size_t bytes) {
  VT_LIVE(c);
  if (cam >= VT_MAX_CAMERAS || !c->cams[cam].set) return fail(c, VT_ERR_STATE, "readback: camera %u not set", cam);
  if (!dst) return fail(c, VT_ERR_ARG, "readback: null destination");
  CamSlot &s = c->cams[cam];
  const size_t px = (size_t)s.c.width * s.c.height;
  VT_HIP(c, hipSetDevice(c->device));
  if (kind == VT_KIND_LINEAR) {
    if (bytes != px * 16) return fail(c, VT_ERR_SIZE, "readback linear: %zu bytes, want %zu", bytes, px * 16);
    if (!s.rendered || s.spp_total == 0) return fail(c, VT_ERR_STATE, "readback: camera %u not rendered", cam);
    const float sc = s.c.scale > 0 ? s.c.scale : 1.0f;     // over each pixel's own sample count
    VT_HIP(c, launch_finalize_linear((const float4 *)s.acc.p, sc, px, (float4 *)s.lin.p, c->stream));
    VT_HIP(c, hipMemcpyAsync(dst, s.lin.p, bytes, hipMemcpyDeviceToHost, c->stream));
  } else if (kind == VT_KIND_RAW) {
    if (bytes != px * 2) return fail(c, VT_ERR_SIZE, "readback raw: %zu bytes, want %zu", bytes, px * 2);
    if (!s.sensed) return fail(c, VT_ERR_STATE, "readback raw: camera %u has no sensor output", cam);
    VT_HIP(c, hipMemcpyAsync(dst, s.dn.p, bytes, hipMemcpyDeviceToHost, c->stream));
  } else if (kind == VT_KIND_RI) {
    if (bytes != px * 4) return fail(c, VT_ERR_SIZE, "readback ri: %zu bytes, want %zu", bytes, px * 4);
    if (!s.ri_valid) return fail(c, VT_ERR_STATE, "readback ri: no sensor run yet");
    VT_HIP(c, hipMemcpyAsync(dst, s.ri.p, bytes, hipMemcpyDeviceToHost, c->stream));
  } else if (kind == VT_KIND_RGB8 || kind == VT_KIND_RGBA8 || kind == VT_KIND_YUV420 || kind == VT_KIND_STATS) {
    if (!s.isp_done) return fail(c, VT_ERR_STATE, "readback: camera %u has no ISP output (vt_isp first)", cam);
    const size_t cpx = (size_t)((s.c.width + 1) / 2) * ((s.c.height + 1) / 2);
    const size_t want = kind == VT_KIND_RGB8 ? px * 3 : kind == VT_KIND_RGBA8 ? px * 4
                        : kind == VT_KIND_YUV420 ? px + 2 * cpx : 16;
    if (bytes != want) return fail(c, VT_ERR_SIZE, "readback kind %d: %zu bytes, want %zu", kind, bytes, want);
    if (kind == VT_KIND_RGB8) {
      int rc = ensure(c, s.rgb8, px * 3);
      if (rc) return rc;
      VT_HIP(c, launch_pack_rgb8((const uchar4 *)s.rgba.p, px, (uint8_t *)s.rgb8.p, c->stream));
      VT_HIP(c, hipMemcpyAsync(dst, s.rgb8.p, bytes, hipMemcpyDeviceToHost, c->stream));
    } else if (kind == VT_KIND_YUV420) {
      if (!s.isp420) return fail(c, VT_ERR_STATE, "readback yuv420: the last ISP pass ran 4:4:4");
      VT_HIP(c, hipMemcpyAsync(dst, s.yuv.p, bytes, hipMemcpyDeviceToHost, c->stream));
    } else {
      VT_HIP(c, hipMemcpyAsync(dst, kind == VT_KIND_RGBA8 ? s.rgba.p : s.stats.p, bytes, hipMemcpyDeviceToHost,
                               c->stream));
    }
  } else {
    return fail(c, VT_ERR_ARG, "readback: unknown kind %d", kind);
  }
  VT_HIP(c, hipStreamSynchronize(c->stream));
  return VT_OK;
}

extern "C" int vt_sync(vt_ctx *c) {
  VT_LIVE(c);
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  return VT_OK;
}

extern "C" int vt_frame_stats_get(vt_ctx *c, vt_frame_stats *out) {
  VT_LIVE(c);
  if (!out) return VT_ERR_ARG;
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  std::memset(out, 0, sizeof *out);
  unsigned long long cnt[8];
  VT_HIP(c, hipMemcpy(cnt, c->counters.p, sizeof cnt, hipMemcpyDeviceToHost));
  out->graph_used = c->graph_used ? 1u : 0u;
  out->graph_builds = c->graph_builds;
  out->ext_launches = c->ext_launches;
  if (c->ext_launches) {
    out->rays_extension = (double)cnt[3];
    out->rays_shadow = (double)cnt[4];
    double sum = 0;
    for (uint32_t k = 0; k < c->ext_launches; k++) {
      float m = 0;
      VT_HIP(c, hipEventElapsedTime(&m, c->ev_ext[2 * k], c->ev_ext[2 * k + 1]));
      sum += m;
    }
    out->ms_ext_avg = sum / c->ext_launches;
  }
  uint32_t ov[4];
  VT_HIP(c, hipMemcpy(ov, c->overflow.p, sizeof ov, hipMemcpyDeviceToHost));
  out->rays = (double)cnt[1];
  out->samples = (double)cnt[2];
  out->stack_overflows = ov[0];
  out->launches = c->last_launches;
  if (c->render_timed) {
    float ms = 0;
    VT_HIP(c, hipEventElapsedTime(&ms, c->ev_r0, c->ev_r1));
    out->ms_render = ms;
    double sum = 0;
    for (uint32_t k = 0; k < c->last_launches; k++) {
      float m = 0;
      VT_HIP(c, hipEventElapsedTime(&m, c->ev_launch[2 * k], c->ev_launch[2 * k + 1]));
      sum += m;
    }
    out->ms_kernel_avg = c->last_launches ? sum / c->last_launches : 0;
  }
  if (c->sensor_timed) {
    float ms = 0;
    VT_HIP(c, hipEventElapsedTime(&ms, c->ev_s0, c->ev_s1));
    out->ms_sensor = ms;
  }
  if (c->isp_timed) {
    float ms = 0;
    VT_HIP(c, hipEventElapsedTime(&ms, c->ev_i0, c->ev_i1));
    out->ms_isp = ms;
  }
  if (c->backward_timed) {
    float ms = 0;
    VT_HIP(c, hipEventElapsedTime(&ms, c->ev_b0, c->ev_b1));
    out->ms_backward = ms;
    out->rays_backward = (double)cnt[5] + (double)cnt[7];
  }
  return VT_OK;
}

// ====================================================================== adjoint
extern "C" int vt_backward(vt_ctx *c, const uint32_t *cams, uint32_t n, const float *const *adj,
                           const vt_render_opts *o, vt_grads *out) {
  VT_LIVE(c);
  {
    bool ref = false;
    if (int rc0 = frame_estimator(c, cams, n, ref)) return rc0;
    if (ref) return fail(c, VT_ERR_ARG, "backward: the adjoint differentiates the path tracer, not the reference estimator");
  }
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "backward: no scene uploaded");
  if (!cams || !adj || !out || n == 0 || n > VT_MAX_CAMERAS) return fail(c, VT_ERR_ARG, "backward: bad arguments");
  if (out->cameras && out->camera_count != n)
    return fail(c, VT_ERR_SIZE, "backward: cameras needs %u x %d floats (one row per listed camera)", n, VT_CAMERA_PARAMS);
  if (!out->materials || out->material_count != c->n_mat) return fail(c, VT_ERR_SIZE, "backward: materials needs %u x 3 floats", c->n_mat);
  if (c->n_light && (!out->lights || out->light_count != c->n_light)) return fail(c, VT_ERR_SIZE, "backward: lights needs %u x 3 floats", c->n_light);
  vt_render_opts opt{};
  if (o) opt = *o;
  const uint32_t chunk = opt.chunk ? opt.chunk : 4;
  for (uint32_t a = 0; a < n; a++) {
    if (cams[a] >= VT_MAX_CAMERAS || !c->cams[cams[a]].set) return fail(c, VT_ERR_STATE, "backward: camera %u not set", cams[a]);
    if (!adj[a]) return fail(c, VT_ERR_ARG, "backward: null adjoint image for camera %u", cams[a]);
    for (uint32_t b = 0; b < a; b++)
      if (cams[a] == cams[b]) return fail(c, VT_ERR_ARG, "backward: camera %u listed twice", cams[a]);
  }
  VT_HIP(c, hipSetDevice(c->device));
  if (c->stage_pending) VT_HIP(c, hipEventSynchronize(c->ev_stage));
  uint64_t items = 0;
  int rc;
  for (uint32_t a = 0; a < n; a++) {
    CamSlot &s = c->cams[cams[a]];
    const vt_camera &k = s.c;
    const size_t px = (size_t)k.width * k.height;
    if ((rc = ensure(c, s.adj, px * 16))) return rc;
    VT_HIP(c, hipMemcpyAsync(s.adj.p, adj[a], px * 16, hipMemcpyHostToDevice, c->stream));
    DevCam &d = c->h_cams[a];
    std::memcpy(d.R, k.R, sizeof d.R);
    std::memcpy(d.pos, k.pos, sizeof d.pos);
    d.mode = k.mode;
    std::memcpy(d.lens, k.lens, sizeof d.lens);
    d.width = k.width; d.height = k.height; d.spp = k.spp; d.seed = k.seed;
    d.cam_index = k.cam_index;
    d.nl = k.light_samples ? k.light_samples : 1;
    d.depth = k.max_depth ? k.max_depth : 1;
    d.tiles_x = (k.width + 7) / 8;
    d.tiles_y = (k.height + 7) / 8;
    const uint32_t s0 = opt.spp_begin, s1 = opt.spp_end ? opt.spp_end : k.spp;
    if (s1 <= s0) return fail(c, VT_ERR_ARG, "backward: empty sub-sample range");
    d.nspp = s1 - s0;
    d.spp_begin = s0;
    d.accumulate = 0;
    d.item_begin = items;
    d.item_count = (uint64_t)d.tiles_x * d.tiles_y * 64ull * d.nspp;
    d.acc = (float4 *)s.acc.p;
    d.adj = (const float4 *)s.adj.p;
    d.scale = k.scale > 0 ? k.scale : 1.0f;
    d.shutter0 = k.shutter[0];
    d.shutter1 = k.shutter[1];
    d.skip = k.skip_bodies;
    d.opts = k.flags;
    items += d.item_count;
  }
  VT_HIP(c, hipMemcpyAsync(c->dcams.p, c->h_cams, sizeof(DevCam) * n, hipMemcpyHostToDevice, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_stage, c->stream));
  c->stage_pending = true;
  const size_t ng = 3 * (size_t)(c->n_mat + c->n_light);
  const size_t ngc = out->cameras ? (size_t)VT_CAMERA_PARAMS * n : 0;
  const size_t ng_pad = (ng + 3) / 4 * 4;          // camera gradients start on a float4 boundary
  if ((rc = ensure(c, c->grads, (ng_pad + ngc) * 4))) return rc;
  VT_HIP(c, launch_zero_f4((float4 *)c->grads.p, (ng_pad + ngc + 3) / 4, c->stream));
  VT_HIP(c, launch_zero_u64((unsigned long long *)c->counters.p, 8, c->stream));
  RenderLaunch L{};
  L.S = c->S;
  L.cams = (const DevCam *)c->dcams.p;
  L.ncam = n;
  L.item_base = 0;
  L.total_waves = items / 64;
  L.counters = (unsigned long long *)c->counters.p + 4;   // [4] queue head (counters[0] of the launch), [5] rays
  L.overflow = (uint32_t *)c->overflow.p;
  L.spill = (uint2 *)c->spill.p;
  L.spill_stride = c->spill_lanes;
  L.chunk = chunk;
  L.persistent = 1;
  L.block = 256;
  L.grid = (uint32_t)std::min<uint64_t>((L.total_waves * 64 + 255) / 256, max_blocks(c, 256));
  float *gm = (float *)c->grads.p;
  VT_HIP(c, hipEventRecord(c->ev_b0, c->stream));
  if (ngc) {
    // the camera walk first, on its own queue head and ray counter ([6], [7]); it leaves each sample's
    // radiance for the gradient walk, which then replays each path once instead of twice
    if ((rc = ensure(c, c->ltot, (size_t)L.total_waves * 64 * sizeof(float4)))) return rc;
    L.ltot = (float4 *)c->ltot.p;
    RenderLaunch Lcam = L;
    Lcam.counters = (unsigned long long *)c->counters.p + 6;
    VT_HIP(c, launch_camgrad(Lcam, gm + ng_pad, c->stream));
  }
  VT_HIP(c, launch_backward(L, gm, gm + 3 * (size_t)c->n_mat, c->n_mat, c->n_light, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_b1, c->stream));
  c->backward_timed = true;
  VT_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<float> g(ng_pad + ngc);
  VT_HIP(c, hipMemcpy(g.data(), c->grads.p, g.size() * 4, hipMemcpyDeviceToHost));
  std::memcpy(out->materials, g.data(), 12 * (size_t)c->n_mat);
  if (c->n_light) std::memcpy(out->lights, g.data() + 3 * (size_t)c->n_mat, 12 * (size_t)c->n_light);
  if (ngc) std::memcpy(out->cameras, g.data() + ng_pad, 4 * ngc);
  return VT_OK;
}

// ====================================================================== bandwidth probe
extern "C" int vt_copy_bandwidth(vt_ctx *c, uint64_t bytes, uint32_t iters, double *gbs) {
  VT_LIVE(c);
  if (!gbs || bytes < 4096 || iters == 0) return fail(c, VT_ERR_ARG, "copy_bandwidth: bytes >= 4096, iters > 0");
  VT_HIP(c, hipSetDevice(c->device));
  DevBuf a, b;
  int rc = ensure(c, a, bytes);
  if (!rc) rc = ensure(c, b, bytes);
  struct Free {
    DevBuf &x, &y;
    ~Free() { release(x); release(y); }
  } fr{a, b};
  if (rc) return rc;
  const uint64_t n = bytes / 16;
  const uint32_t grid = (uint32_t)c->cu_count * 8;
  VT_HIP(c, launch_zero_f4((float4 *)a.p, n, c->stream));
  for (int w = 0; w < 3; w++) VT_HIP(c, launch_copy_f4((const float4 *)a.p, (float4 *)b.p, n, grid, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_s0, c->stream));
  for (uint32_t k = 0; k < iters; k++) VT_HIP(c, launch_copy_f4((const float4 *)a.p, (float4 *)b.p, n, grid, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_s1, c->stream));
  VT_HIP(c, hipEventSynchronize(c->ev_s1));
  float ms = 0;
  VT_HIP(c, hipEventElapsedTime(&ms, c->ev_s0, c->ev_s1));
  *gbs = 2.0 * 16.0 * (double)n * iters / (ms * 1e-3) / 1e9;
  c->sensor_timed = false;
  return VT_OK;
}

// ====================================================================== test entry points
extern "C" int vt_trace_test(vt_ctx *c, const float *rays, uint32_t n, int any_hit, float *hits) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "trace_test: no scene");
  if (n && (!rays || !hits)) return VT_ERR_ARG;
  if (n == 0) return VT_OK;
  VT_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = ensure(c, c->test_in, (size_t)n * 32))) return rc;
  if ((rc = ensure(c, c->test_out, (size_t)n * 16))) return rc;
  VT_HIP(c, hipMemcpy(c->test_in.p, rays, (size_t)n * 32, hipMemcpyHostToDevice));
  TestLaunch T{c->S, (uint32_t *)c->overflow.p, (uint2 *)c->spill.p, c->spill_lanes,
               (uint32_t)(c->spill_lanes / DEFAULT_BLOCK)};
  VT_HIP(c, launch_trace_test(T, (const float *)c->test_in.p, n, any_hit, (float *)c->test_out.p, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(hits, c->test_out.p, (size_t)n * 16, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_visible(vt_ctx *c, uint32_t cam, const float *points, uint32_t n, float tol, float time,
                          uint8_t *out) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "visible: no scene");
  if (cam >= VT_MAX_CAMERAS || !c->cams[cam].set) return fail(c, VT_ERR_STATE, "visible: camera %u not set", cam);
  if (n && (!points || !out)) return fail(c, VT_ERR_ARG, "visible: null points or output");
  if (!(tol >= 0.0f) || !(time >= 0.0f && time <= 1.0f)) return fail(c, VT_ERR_ARG, "visible: tol >= 0, time in [0, 1]");
  if (n == 0) return VT_OK;
  VT_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = ensure(c, c->test_in, (size_t)n * 12))) return rc;
  if ((rc = ensure(c, c->test_out, (size_t)n))) return rc;
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(c->test_in.p, points, (size_t)n * 12, hipMemcpyHostToDevice));
  TestLaunch T{c->S, (uint32_t *)c->overflow.p, (uint2 *)c->spill.p, c->spill_lanes,
               (uint32_t)(c->spill_lanes / DEFAULT_BLOCK)};
  T.S.root = 0;                                 // the whole static scene, as the reference's CPU walk
  const vt_camera &k = c->cams[cam].c;
  VT_HIP(c, launch_visible(T, k.pos, k.skip_bodies, tol, time, (const float *)c->test_in.p, n, (uint8_t *)c->test_out.p,
                           c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(out, c->test_out.p, n, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_lens_test(vt_ctx *c, const float lens[9], const float *uv, uint32_t n, float *rays) {
  VT_LIVE(c);
  if (!lens || (n && (!uv || !rays))) return VT_ERR_ARG;
  if (n == 0) return VT_OK;
  VT_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = ensure(c, c->test_in, (size_t)n * 8 + 64))) return rc;
  if ((rc = ensure(c, c->test_out, (size_t)n * 12))) return rc;
  VT_HIP(c, hipMemcpy(c->test_in.p, lens, 36, hipMemcpyHostToDevice));
  VT_HIP(c, hipMemcpy((char *)c->test_in.p + 64, uv, (size_t)n * 8, hipMemcpyHostToDevice));
  VT_HIP(c, launch_lens_test((const float *)c->test_in.p, (const float *)((char *)c->test_in.p + 64), n,
                             (float *)c->test_out.p, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(rays, c->test_out.p, (size_t)n * 12, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_camera_ray_test(vt_ctx *c, uint32_t cam, const uint32_t *xyi, uint32_t n, float *out) {
  VT_LIVE(c);
  if (cam >= VT_MAX_CAMERAS || !c->cams[cam].set) return fail(c, VT_ERR_STATE, "camera %u not set", cam);
  if (n && (!xyi || !out)) return VT_ERR_ARG;
  if (n == 0) return VT_OK;
  VT_HIP(c, hipSetDevice(c->device));
  const vt_camera &k = c->cams[cam].c;
  DevCam d{};
  std::memcpy(d.R, k.R, sizeof d.R);
  std::memcpy(d.pos, k.pos, sizeof d.pos);
  d.mode = k.mode;
  std::memcpy(d.lens, k.lens, sizeof d.lens);
  d.width = k.width; d.height = k.height; d.spp = k.spp; d.seed = k.seed; d.cam_index = k.cam_index;
  int rc;
  if ((rc = ensure(c, c->test_in, (size_t)n * 12 + 256))) return rc;
  if ((rc = ensure(c, c->test_out, (size_t)n * 24))) return rc;
  VT_HIP(c, hipMemcpy(c->test_in.p, &d, sizeof d, hipMemcpyHostToDevice));
  VT_HIP(c, hipMemcpy((char *)c->test_in.p + 256, xyi, (size_t)n * 12, hipMemcpyHostToDevice));
  VT_HIP(c, launch_camray_test((const DevCam *)c->test_in.p, (const uint32_t *)((char *)c->test_in.p + 256), n,
                               (float *)c->test_out.p, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(out, c->test_out.p, (size_t)n * 24, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_bsdf_test(vt_ctx *c, const float *cases, uint32_t n, float *out) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "bsdf_test: no scene (materials)");
  if (n && (!cases || !out)) return VT_ERR_ARG;
  if (n == 0) return VT_OK;
  VT_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = ensure(c, c->test_in, (size_t)n * 64))) return rc;
  if ((rc = ensure(c, c->test_out, (size_t)n * 32))) return rc;
  VT_HIP(c, hipMemcpy(c->test_in.p, cases, (size_t)n * 64, hipMemcpyHostToDevice));
  VT_HIP(c, launch_bsdf_test(c->S.mats, (const float *)c->test_in.p, n, (float *)c->test_out.p, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(out, c->test_out.p, (size_t)n * 32, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_capsule_test(vt_ctx *c, const float *cases, uint32_t n, float *out) {
  VT_LIVE(c);
  if (n && (!cases || !out)) return VT_ERR_ARG;
  if (n == 0) return VT_OK;
  VT_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = ensure(c, c->test_in, (size_t)n * 64))) return rc;
  if ((rc = ensure(c, c->test_out, (size_t)n * 4))) return rc;
  VT_HIP(c, hipMemcpy(c->test_in.p, cases, (size_t)n * 64, hipMemcpyHostToDevice));
  VT_HIP(c, launch_capsule_test((const float *)c->test_in.p, n, (float *)c->test_out.p, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(out, c->test_out.p, (size_t)n * 4, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_sensor_test(vt_ctx *c, const vt_sensor_params *p, uint32_t w, uint32_t h, const float *radiance,
                              const float *ri, uint16_t *dn) {
  VT_LIVE(c);
  if (!radiance || !ri || !dn || w == 0 || h == 0) return VT_ERR_ARG;
  VT_HIP(c, hipSetDevice(c->device));
  SensorDev P;
  int rc = sensor_params(c, p, w, h, 0, 1.0f, P);
  if (rc) return rc;
  P.cam_index = 0;
  const size_t px = (size_t)w * h;
  // test_aux: the RI map, then the two glare levels (float4-aligned)
  const size_t gb = glare_bytes(w, h, P), ga = (px * 4 + 255) / 256 * 256;
  if ((rc = ensure(c, c->test_in, px * 16))) return rc;
  if ((rc = ensure(c, c->test_aux, ga + 2 * gb))) return rc;
  if ((rc = ensure(c, c->test_out, px * 2))) return rc;
  VT_HIP(c, hipMemcpy(c->test_in.p, radiance, px * 16, hipMemcpyHostToDevice));
  VT_HIP(c, hipMemcpy(c->test_aux.p, ri, px * 4, hipMemcpyHostToDevice));
  char *g = (char *)c->test_aux.p + ga;
  VT_HIP(c, launch_sensor(P, (const float4 *)c->test_in.p, (const float *)c->test_aux.p, (float4 *)g,
                          (float4 *)(g + gb), (uint16_t *)c->test_out.p, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(dn, c->test_out.p, px * 2, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_isp_test(vt_ctx *c, const vt_sensor_params *sp, const vt_isp_params *q, uint32_t w, uint32_t h,
                           const uint16_t *dn, uint8_t *rgba, uint8_t *yuv, uint32_t *stats) {
  VT_LIVE(c);
  if (!sp || !dn || !rgba || !stats || w == 0 || h == 0) return fail(c, VT_ERR_ARG, "isp_test: null argument");
  if (sp->cfa > 3) return fail(c, VT_ERR_ARG, "isp_test: cfa code %u", sp->cfa);
  VT_HIP(c, hipSetDevice(c->device));
  IspDev P;
  int rc = isp_params(c, *sp, q, w, h, P);
  if (rc) return rc;
  const size_t px = (size_t)w * h, cpx = (size_t)((w + 1) / 2) * ((h + 1) / 2);
  // test_in: the RGB plane, then the partial sums; test_aux: rgba, yuv, stats; test_out: the DN
  if ((rc = ensure(c, c->test_in, px * 16 + (size_t)isp_partials(w, h) * 16))) return rc;
  const size_t o_rgba = 0, o_yuv = o_rgba + (px * 4 + 255) / 256 * 256,
               o_st = o_yuv + (px + 2 * cpx + 255) / 256 * 256;
  if ((rc = ensure(c, c->test_aux, o_st + 16))) return rc;
  if ((rc = ensure(c, c->test_out, px * 2))) return rc;
  char *aux = (char *)c->test_aux.p;
  VT_HIP(c, hipMemcpy(c->test_out.p, dn, px * 2, hipMemcpyHostToDevice));
  VT_HIP(c, launch_isp(P, (const uint16_t *)c->test_out.p, (float4 *)c->test_in.p,
                       (uint4 *)((float4 *)c->test_in.p + px), (uchar4 *)(aux + o_rgba), (uint8_t *)(aux + o_yuv),
                       (unsigned int *)(aux + o_st), c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(rgba, aux + o_rgba, px * 4, hipMemcpyDeviceToHost));
  if (yuv && P.chroma420) VT_HIP(c, hipMemcpy(yuv, aux + o_yuv, px + 2 * cpx, hipMemcpyDeviceToHost));
  VT_HIP(c, hipMemcpy(stats, aux + o_st, 16, hipMemcpyDeviceToHost));
  return VT_OK;
}

// ====================================================================== node radiosity
extern "C" int vt_jacobi_sparse(vt_ctx *c, const vt_gather_tables *t, float *node_state, uint32_t iters,
                                float *aux_out) {
  VT_LIVE(c);
  if (!t || !node_state || !t->active || !t->gtable || !t->normals || !t->node_mat || !t->albedo || !t->hh)
    return fail(c, VT_ERR_ARG, "jacobi: null table");
  if (t->nbr_k && (!t->nbr || !t->nbr_w)) return fail(c, VT_ERR_ARG, "jacobi: null neighbour table");
  if (t->nbr_k > 32) return fail(c, VT_ERR_ARG, "jacobi: at most 32 smoothing neighbours");
  if (t->gather_k == 0 || t->albedo_bins == 0 || t->n_mat == 0) return fail(c, VT_ERR_ARG, "jacobi: empty sizes");
  const size_t N = t->n_static, A = t->n_active, K = t->gather_k;
  for (size_t k = 0; k < A; k++)
    if (t->active[k] >= N) return fail(c, VT_ERR_ARG, "jacobi: active node out of range");
  for (size_t k = 0; k < N * K; k++)
    if (t->gtable[k] != 0xffffffffu && t->gtable[k] >= N) return fail(c, VT_ERR_ARG, "jacobi: gather entry out of range");
  for (size_t k = 0; k < N; k++)
    if (t->node_mat[k] >= t->n_mat) return fail(c, VT_ERR_ARG, "jacobi: node material out of range");
  for (size_t k = 0; k < A * t->nbr_k; k++)
    if (t->nbr_w[k] > 0 && t->nbr[k] >= A) return fail(c, VT_ERR_ARG, "jacobi: neighbour out of range");
  VT_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = upload(c, c->g_active, t->active, A * 4))) return rc;
  if ((rc = upload(c, c->g_table, t->gtable, N * K * 4))) return rc;
  if ((rc = upload(c, c->g_normals, t->normals, N * 16))) return rc;
  if ((rc = upload(c, c->g_mat, t->node_mat, N * 4))) return rc;
  if ((rc = upload(c, c->g_alb, t->albedo, (size_t)t->n_mat * t->albedo_bins * 32))) return rc;
  if ((rc = upload(c, c->g_hh, t->hh, (size_t)t->n_mat * 16))) return rc;
  if (t->nbr_k) {
    if ((rc = upload(c, c->g_nbr, t->nbr, A * t->nbr_k * 4))) return rc;
    if ((rc = upload(c, c->g_nbrw, t->nbr_w, A * t->nbr_k * 4))) return rc;
  }
  if ((rc = upload(c, c->g_state, node_state, N * 64))) return rc;
  if ((rc = ensure(c, c->g_aux, std::max<size_t>(A, 1) * 16))) return rc;
  GatherDev G{};
  G.n_active = (uint32_t)A; G.gather_k = (uint32_t)K; G.nbr_k = t->nbr_k; G.bins = t->albedo_bins;
  G.active = (const uint32_t *)c->g_active.p; G.gtable = (const uint32_t *)c->g_table.p;
  G.node_mat = (const uint32_t *)c->g_mat.p; G.nbr = (const uint32_t *)c->g_nbr.p;
  G.normals = (const float4 *)c->g_normals.p; G.albedo = (const float4 *)c->g_alb.p; G.hh = (const float4 *)c->g_hh.p;
  G.nbr_w = (const float *)c->g_nbrw.p; G.node_state = (float4 *)c->g_state.p; G.aux = (float4 *)c->g_aux.p;
  // ping-pong slots 1 and 2 (gpu/renderer.js:729-749): step k reads slot_in and writes the other
  uint32_t slot_in = 1;
  for (uint32_t it = 0; it < iters; it++) {
    G.slot_in = slot_in;
    G.slot_out = 3 - slot_in;
    VT_HIP(c, launch_jacobi_gather(G, c->stream));
    VT_HIP(c, launch_jacobi_update(G, c->stream));
    slot_in = 3 - slot_in;
  }
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(node_state, c->g_state.p, N * 64, hipMemcpyDeviceToHost));
  if (aux_out) VT_HIP(c, hipMemcpy(aux_out, c->g_aux.p, A * 16, hipMemcpyDeviceToHost));
  return VT_OK;
}

// ====================================================================== node radiosity on the uploaded scene
namespace {

// smoothing neighbours of every node in a list (_neighbours, gpu/renderer.js:574-616): the same-facing
// (n.n' >= 0.9) nodes of the list within R, Gaussian weights exp(-d^2 / (2 sigma^2)), the K largest (ties: lower
// index first), sorted by weight; a uniform grid of cell R finds the candidates. sigma <= 0: each node alone.
void node_neighbours(const float *pos, const float *nrm, size_t n, uint32_t K, double R, double sg,
                     std::vector<uint32_t> &nb, std::vector<float> &nw) {
  nb.assign(n * K, 0u);
  nw.assign(n * K, 0.0f);
  if (K == 0) return;
  if (!(sg > 0)) {
    for (size_t k = 0; k < n; k++) {
      nb[k * K] = (uint32_t)k;
      nw[k * K] = 1.0f;
    }
    return;
  }
  const int64_t OFF = 1 << 20;
  auto cell = [&](double x) { return (int64_t)std::floor(x / R) + OFF; };
  auto key = [](int64_t x, int64_t y, int64_t z) {
    return ((uint64_t)(x & 0x1fffff) << 42) | ((uint64_t)(y & 0x1fffff) << 21) | (uint64_t)(z & 0x1fffff);
  };
  std::vector<std::pair<uint64_t, uint32_t>> order(n);
  for (size_t k = 0; k < n; k++)
    order[k] = {key(cell(pos[3 * k]), cell(pos[3 * k + 1]), cell(pos[3 * k + 2])), (uint32_t)k};
  std::sort(order.begin(), order.end());
  std::vector<double> cw(K);
  std::vector<uint32_t> cm(K);
  for (size_t k = 0; k < n; k++) {
    const double px = pos[3 * k], py = pos[3 * k + 1], pz = pos[3 * k + 2];
    const double nx = nrm[3 * k], ny = nrm[3 * k + 1], nz = nrm[3 * k + 2];
    const int64_t cx = cell(px), cy = cell(py), cz = cell(pz);
    uint32_t cnt = 0;
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -1; dz <= 1; dz++) {
          const uint64_t kk = key(cx + dx, cy + dy, cz + dz);
          auto lo = std::lower_bound(order.begin(), order.end(), std::make_pair(kk, 0u));
          for (auto it = lo; it != order.end() && it->first == kk; ++it) {
            const uint32_t m = it->second;
            const double ex = pos[3 * m] - px, ey = pos[3 * m + 1] - py, ez = pos[3 * m + 2] - pz;
            const double d2 = ex * ex + ey * ey + ez * ez;
            if (d2 > R * R) continue;
            if (nx * nrm[3 * m] + ny * nrm[3 * m + 1] + nz * nrm[3 * m + 2] < 0.9) continue;
            const double w = std::exp(-0.5 * d2 / (sg * sg));
            uint32_t at = cnt < K ? cnt : K;
            while (at > 0 && (cw[at - 1] < w || (cw[at - 1] == w && cm[at - 1] > m))) at--;
            if (at >= K) continue;
            for (uint32_t q = std::min(cnt, K - 1); q > at; q--) {
              cw[q] = cw[q - 1];
              cm[q] = cm[q - 1];
            }
            cw[at] = w;
            cm[at] = m;
            if (cnt < K) cnt++;
          }
        }
    for (uint32_t q = 0; q < cnt; q++) {
      nb[k * K + q] = cm[q];
      nw[k * K + q] = (float)cw[q];
    }
  }
}

NodeLaunch node_launch(vt_ctx *c) {
  NodeLaunch L;
  std::memset(&L, 0, sizeof L);
  L.S = c->S;
  L.npos = (const float4 *)c->g_pos.p;
  L.nnorm = (const float4 *)c->g_normals.p;
  L.nmat = (const uint32_t *)c->g_mat.p;
  L.n_nodes = (uint32_t)(c->h_verts.size() / 8);
  L.K = c->rad_K;
  L.node_samples = c->rad_samples;
  L.n_active = c->rad_active;
  L.n_mat = c->n_mat;
  L.gtable = (uint32_t *)c->g_table.p;
  L.miss = (float *)c->g_miss.p;
  L.active = (const uint32_t *)c->g_active.p;
  L.state = (float4 *)c->g_state.p;
  L.alb = (float4 *)c->g_alb.p;
  L.hh = (float4 *)c->g_hh.p;
  L.overflow = (uint32_t *)c->overflow.p;
  L.spill = (uint2 *)c->spill.p;
  L.spill_stride = c->spill_lanes;
  L.max_blocks = (uint32_t)(c->spill_lanes / 256);
  return L;
}


}  // namespace

extern "C" int vt_node_neighbours(const float *pos, const float *nrm, uint32_t n, uint32_t k, float radius, float sigma,
                                  uint32_t *idx, float *w) {
  if (n && (!pos || !nrm || !idx || !w)) return VT_ERR_ARG;
  if (k == 0 || k > 32 || !(radius > 0)) return VT_ERR_ARG;
  std::vector<uint32_t> nb;
  std::vector<float> nw;
  node_neighbours(pos, nrm, n, k, radius, sigma, nb, nw);
  std::memcpy(idx, nb.data(), nb.size() * 4);
  std::memcpy(w, nw.data(), nw.size() * 4);
  return VT_OK;
}

extern "C" int vt_radiosity_build(vt_ctx *c, const vt_radiosity_opts *o, vt_radiosity_info *info) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "radiosity: no scene uploaded");
  vt_radiosity_opts op{64, 4, 16, 0.006f, 0.003f};
  if (o) op = *o;
  if (op.gather_k < 8 || op.gather_k > 1024 || op.node_samples < 1 || op.node_samples > 1024 || op.nbr_k < 1 ||
      op.nbr_k > 32 || !(op.smooth_radius > 0) || !(op.smooth_sigma >= 0))
    return fail(c, VT_ERR_ARG, "radiosity: gather_k 8..1024, node_samples 1..1024, nbr_k 1..32, radius > 0, sigma >= 0");
  const auto t0 = std::chrono::steady_clock::now();
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  // the gather rays walk the whole static scene; an interior tree of an earlier build is dropped (its slots stay
  // unused until vt_interior_bvh builds the next one from the new miss counts)
  if (c->S.root) c->scene_gen++;
  c->S.root = 0;
  c->interior_root = 0;
  c->interior_tris = 0;
  const size_t V = c->h_verts.size() / 8;
  // static nodes: the vertices of static-flagged triangles of the static scene, with a third of every such
  // triangle's area; a node's material is that of its own mesh (the vertex record's word 3, as the reference's
  // prepared vertices carry it, gpu/scene_gpu.js:384-387)
  std::vector<float4> npos(V), nnorm(V);
  std::vector<uint32_t> nmat(V, 0u);
  std::vector<float> area(V, 0.0f);
  std::vector<uint8_t> stat(V, 0);
  for (uint32_t sl = 0; sl < c->static_slots; sl++) {
    const uint4 t = c->h_tvid[sl];
    if (!(c->h_meshes[4 * (size_t)t.w + 1] & VT_MESH_STATIC)) continue;
    const uint32_t vv[3] = {t.x, t.y, t.z};
    const float *a = &c->h_verts[8 * (size_t)t.x], *b = &c->h_verts[8 * (size_t)t.y], *q = &c->h_verts[8 * (size_t)t.z];
    const double e1[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
    const double e2[3] = {(double)q[0] - a[0], (double)q[1] - a[1], (double)q[2] - a[2]};
    const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
    const double ar = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
    for (uint32_t v : vv) {
      stat[v] = 1;
      area[v] += (float)(ar / 3.0);
    }
  }
  const uint32_t n_mesh = (uint32_t)(c->h_meshes.size() / 4);
  for (size_t v = 0; v < V; v++) {
    if (!stat[v]) continue;
    uint32_t mesh;
    std::memcpy(&mesh, &c->h_verts[8 * v + 3], 4);
    if (mesh >= n_mesh) return fail(c, VT_ERR_ARG, "radiosity: vertex %zu mesh id %u out of range", v, mesh);
    nmat[v] = c->h_meshes[4 * (size_t)mesh];
  }
  uint32_t n_static = 0;
  for (size_t v = 0; v < V; v++) {
    const float *p = &c->h_verts[8 * v];
    double nx = p[4], ny = p[5], nz = p[6];
    const double l = std::sqrt(nx * nx + ny * ny + nz * nz);
    if (l > 0) { nx /= l; ny /= l; nz /= l; } else { nx = 0; ny = 0; nz = 1; }
    npos[v] = make_float4(p[0], p[1], p[2], stat[v] ? 1.0f : 0.0f);
    nnorm[v] = make_float4((float)nx, (float)ny, (float)nz, 0.0f);
    n_static += stat[v];
  }
  int rc;
  if ((rc = upload(c, c->g_pos, npos.data(), V * sizeof(float4)))) return rc;
  if ((rc = upload(c, c->g_normals, nnorm.data(), V * sizeof(float4)))) return rc;
  if ((rc = upload(c, c->g_mat, nmat.data(), V * 4))) return rc;
  if ((rc = upload(c, c->g_area, area.data(), V * 4))) return rc;
  if ((rc = ensure(c, c->g_table, V * op.gather_k * 4))) return rc;
  if ((rc = ensure(c, c->g_miss, std::max<size_t>(V, 1) * 4))) return rc;
  if ((rc = ensure(c, c->g_alb, (size_t)c->n_mat * ALB_BINS * 32))) return rc;
  if ((rc = ensure(c, c->g_hh, (size_t)c->n_mat * 16))) return rc;
  if ((rc = ensure(c, c->g_state, std::max<size_t>(V, 1) * 64))) return rc;
  if ((rc = ensure(c, c->radu, 2 * sizeof(float4)))) return rc;
  VT_HIP(c, launch_zero_f4((float4 *)c->g_state.p, (uint64_t)V * 4, c->stream));
  c->rad_K = op.gather_k;
  c->rad_samples = op.node_samples;
  c->rad_nbr_k = op.nbr_k;
  c->rad_active = 0;
  NodeLaunch L = node_launch(c);
  VT_HIP(c, launch_albedo_tables(L, c->stream));
  VT_HIP(c, launch_gather_build(L, c->stream));
  VT_HIP(c, launch_gather_count(L, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<float> miss(V);
  if (V) VT_HIP(c, hipMemcpy(miss.data(), c->g_miss.p, V * 4, hipMemcpyDeviceToHost));
  std::vector<uint32_t> act;
  for (size_t v = 0; v < V; v++)
    if (stat[v] && miss[v] <= 0.5f * (float)op.gather_k) act.push_back((uint32_t)v);
  const size_t A = act.size();
  if ((rc = upload(c, c->g_active, act.empty() ? nullptr : act.data(), std::max<size_t>(A, 1) * 4))) return rc;
  c->rad_active = (uint32_t)A;
  L = node_launch(c);
  VT_HIP(c, launch_gather_remap(L, c->stream));
  // smoothing neighbours among the active nodes, by active index
  std::vector<float> ap(3 * A), an(3 * A);
  for (size_t k = 0; k < A; k++)
    for (int j = 0; j < 3; j++) {
      ap[3 * k + j] = (&npos[act[k]].x)[j];
      an[3 * k + j] = (&nnorm[act[k]].x)[j];
    }
  std::vector<uint32_t> nb;
  std::vector<float> nw;
  node_neighbours(ap.data(), an.data(), A, op.nbr_k, op.smooth_radius, op.smooth_sigma, nb, nw);
  if ((rc = upload(c, c->g_nbr, nb.empty() ? nullptr : nb.data(), std::max<size_t>(nb.size(), 1) * 4))) return rc;
  if ((rc = upload(c, c->g_nbrw, nw.empty() ? nullptr : nw.data(), std::max<size_t>(nw.size(), 1) * 4))) return rc;
  if ((rc = ensure(c, c->g_aux, std::max<size_t>(A, 1) * 16))) return rc;
  // the emitters' centroid and radius for the shadow culling (gpu/scene_gpu.js:418-422)
  double cen[3] = {0, 0, 0}, tot = 0, rad = 0;
  const size_t E = c->h_emit.size() / 16;
  for (size_t e = 0; e < E; e++) {
    const float *r = &c->h_emit[16 * e];
    for (int k = 0; k < 3; k++) cen[k] += ((double)r[k] + ((double)r[4 + k] + r[8 + k]) / 3.0) * r[7];
    tot += r[7];
  }
  for (int k = 0; k < 3; k++) cen[k] /= tot > 0 ? tot : 1.0;
  for (size_t e = 0; e < E; e++) {
    const float *r = &c->h_emit[16 * e];
    for (int q = 0; q < 3; q++) {
      double d2 = 0;
      for (int k = 0; k < 3; k++) {
        const double x = (double)r[k] + (q == 1 ? r[4 + k] : 0.0) + (q == 2 ? r[8 + k] : 0.0) - cen[k];
        d2 += x * x;
      }
      rad = std::max(rad, std::sqrt(d2));
    }
  }
  // the gather entries' static weights, once per build, when the table stays within 4 GiB (else each step
  // computes them)
  c->rad_weights = false;
  if (A && (uint64_t)A * op.gather_k * 16 <= (4ull << 30)) {
    if ((rc = ensure(c, c->g_w, (size_t)A * op.gather_k * 16))) return rc;
    GatherDev G{};
    G.n_active = (uint32_t)A; G.gather_k = op.gather_k; G.bins = ALB_BINS;
    G.active = (const uint32_t *)c->g_active.p; G.gtable = (const uint32_t *)c->g_table.p;
    G.node_mat = (const uint32_t *)c->g_mat.p; G.normals = (const float4 *)c->g_normals.p;
    G.albedo = (const float4 *)c->g_alb.p; G.gw = (float4 *)c->g_w.p;
    VT_HIP(c, launch_jacobi_weights(G, c->stream));
    c->rad_weights = true;
  }
  c->rad_slot = 1;
  VT_HIP(c, launch_fill_f4((float4 *)c->radu.p, 1, make_float4(0.f, 0.f, 0.f, bits_f(1u)), c->stream));
  VT_HIP(c, launch_fill_f4((float4 *)c->radu.p + 1, 1, make_float4((float)cen[0], (float)cen[1], (float)cen[2], (float)rad),
                           c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  c->S.nstate = (const float4 *)c->g_state.p;
  c->S.alb = (const float4 *)c->g_alb.p;
  c->S.hh = (const float4 *)c->g_hh.p;
  c->S.radu = (const float4 *)c->radu.p;
  c->rad_built = true;
  drop_graph(c);
  if (info) {
    info->nodes = (uint32_t)V;
    info->static_nodes = n_static;
    info->active = (uint32_t)A;
    info->ms_build = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return VT_OK;
}

// _buildInteriorBVH (gpu/renderer.js:533-571): a second static tree without the triangles whose three nodes see
// the sky in >= 90 % of their gather rays (the shell's outer skin; bench triangles always stay). Its triangles take
// new slots past the scene's (copies of their records, so a hit's slot still names its vertices and mesh) and its
// nodes follow the scene's; bodies keep their trees.
static int build_interior(vt_ctx *c) {
  const size_t V = c->h_verts.size() / 8;
  const uint32_t T = c->tri_count, K = c->rad_K;
  std::vector<float> miss(V);
  if (V) VT_HIP(c, hipMemcpy(miss.data(), c->g_miss.p, V * 4, hipMemcpyDeviceToHost));
  std::vector<uint32_t> keep;
  for (uint32_t k = 0; k < c->static_slots; k++) {
    const uint4 t = c->h_tvid[k];
    const bool bench = (c->h_meshes[4 * (size_t)t.w + 1] & 64u) != 0;    // MESH_FLAG.BENCH
    const float lim = 0.9f * (float)K;
    if (!bench && miss[t.x] >= lim && miss[t.y] >= lim && miss[t.z] >= lim) continue;
    keep.push_back(k);
  }
  const uint32_t n = (uint32_t)keep.size();
  c->interior_root = 0;
  c->interior_tris = 0;
  if (n == 0 || n == c->static_slots) return VT_OK;     // nothing to drop
  if ((uint64_t)T + n >= (1u << 26)) return fail(c, VT_ERR_SIZE, "interior_bvh: more than 2^26 slots");
  std::vector<float> pos(V * 3);
  for (size_t v = 0; v < V; v++)
    for (int k = 0; k < 3; k++) pos[3 * v + k] = c->h_verts[8 * v + k];
  std::vector<uint32_t> idx((size_t)n * 3), order(n);
  for (uint32_t m = 0; m < n; m++) {
    const uint4 t = c->h_tvid[keep[m]];
    idx[3 * m] = t.x;
    idx[3 * m + 1] = t.y;
    idx[3 * m + 2] = t.z;
  }
  std::vector<float> ng2((size_t)(n + 2) * 16);
  const int64_t nn = build_bvh2(pos.data(), idx.data(), n, 4, 16, ng2.data(), n + 2, order.data());
  if (nn < 0) return fail(c, VT_ERR_ARG, "interior_bvh: BVH build failed");
  std::vector<Node4> g4;
  try {
    g4 = collapse_bvh4(ng2.data(), (uint32_t)nn, n);
  } catch (const std::exception &ex) {
    return fail(c, VT_ERR_ARG, "interior_bvh: %s", ex.what());
  }
  // the scene's nodes, then the interior tree's, relocated
  std::vector<Node4> nodes(c->node4_count + g4.size());
  VT_HIP(c, hipMemcpy(nodes.data(), c->nodes.p, (size_t)c->node4_count * sizeof(Node4), hipMemcpyDeviceToHost));
  for (size_t q = 0; q < g4.size(); q++) {
    Node4 nd = g4[q];
    for (int ch = 0; ch < 4; ch++) {
      const uint32_t r = nd.ref[ch];
      if (r == EMPTY_REF) continue;
      if (r & LEAF_BIT) nd.ref[ch] = (r & ~0x03ffffffu) | ((r & 0x03ffffffu) + T);
      else nd.ref[ch] = r + c->node4_count;
    }
    nodes[c->node4_count + q] = nd;
  }
  // slot records for T + n slots: the scene's, then copies of the kept triangles in the interior tree's order
  const size_t S2 = (size_t)T + n;
  std::vector<uint4> attr(S2);
  for (uint32_t k = 0; k < T; k++) attr[k] = c->h_tvid[k];
  for (uint32_t m = 0; m < n; m++) attr[T + m] = c->h_tvid[keep[order[m]]];
  std::vector<float4> tri(S2 * 3), vn(V);
  for (size_t k = 0; k < S2; k++) {
    const uint32_t vid[3] = {attr[k].x, attr[k].y, attr[k].z};
    for (int j = 0; j < 3; j++) {
      const float *p = &c->h_verts[8 * (size_t)vid[j]];
      tri[3 * k + j] = make_float4(p[0], p[1], p[2], 0.0f);
    }
  }
  for (size_t v = 0; v < V; v++) {
    const float *p = &c->h_verts[8 * v];
    vn[v] = make_float4(p[4], p[5], p[6], 0.0f);
  }
  int rc;
  if ((rc = upload(c, c->nodes, nodes.data(), nodes.size() * sizeof(Node4)))) return rc;
  if ((rc = upload(c, c->tri, tri.data(), tri.size() * sizeof(float4)))) return rc;
  if ((rc = upload(c, c->tvid, attr.data(), attr.size() * sizeof(uint4)))) return rc;
  DevBuf d_vn;
  rc = upload(c, d_vn, vn.data(), vn.size() * sizeof(float4));
  if (!rc) rc = ensure(c, c->tshade, S2 * 64);
  if (!rc) {
    hipError_t e = launch_tri_shading((const float4 *)c->tri.p, (const uint4 *)c->tvid.p, (const float4 *)d_vn.p,
                                      (uint32_t)S2, (float4 *)c->tshade.p, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_fail(c, e, "interior triangle shading records");
  }
  release(d_vn);
  if (rc) return rc;
  c->S.nodes = (const Node4 *)c->nodes.p;
  c->S.tri = (const float4 *)c->tri.p;
  c->S.tvid = (const uint4 *)c->tvid.p;
  c->S.tshade = (const float4 *)c->tshade.p;
  c->interior_root = c->node4_count;
  c->interior_tris = n;
  return VT_OK;
}

extern "C" int vt_interior_bvh(vt_ctx *c, uint32_t enable, uint32_t *interior_tris) {
  VT_LIVE(c);
  if (!c->rad_built) return fail(c, VT_ERR_STATE, "interior_bvh: the node cache is not built");
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  if (enable && !c->interior_tris) {
    if (int rc = build_interior(c)) return rc;
  }
  c->S.root = enable ? c->interior_root : 0u;
  c->scene_gen++;                 // the fixed cameras' caches hold slots and hits of the other tree
  drop_graph(c);
  if (interior_tris) *interior_tris = c->S.root ? c->interior_tris : c->static_slots;
  return VT_OK;
}

extern "C" int vt_radiosity_update(vt_ctx *c, uint32_t direct, uint32_t iters) {
  VT_LIVE(c);
  if (!c->rad_built) return fail(c, VT_ERR_STATE, "radiosity: not built");
  VT_HIP(c, hipSetDevice(c->device));
  c->S.nstate = (const float4 *)c->g_state.p;     // a scene's instances may have changed S since the build
  NodeLaunch L = node_launch(c);
  if (direct) VT_HIP(c, launch_node_direct(L, c->stream));
  GatherDev G{};
  G.n_active = c->rad_active; G.gather_k = c->rad_K; G.nbr_k = c->rad_nbr_k; G.bins = ALB_BINS;
  G.active = (const uint32_t *)c->g_active.p; G.gtable = (const uint32_t *)c->g_table.p;
  G.node_mat = (const uint32_t *)c->g_mat.p; G.nbr = (const uint32_t *)c->g_nbr.p;
  G.normals = (const float4 *)c->g_normals.p; G.albedo = (const float4 *)c->g_alb.p; G.hh = (const float4 *)c->g_hh.p;
  G.nbr_w = (const float *)c->g_nbrw.p; G.node_state = (float4 *)c->g_state.p; G.aux = (float4 *)c->g_aux.p;
  G.gw = c->rad_weights ? (float4 *)c->g_w.p : nullptr;
  // ping-pong source slots 1 and 2 (gpu/renderer.js:729-749): a step reads slot_in and writes the other
  for (uint32_t it = 0; it < iters; it++) {
    G.slot_in = c->rad_slot;
    G.slot_out = 3 - c->rad_slot;
    VT_HIP(c, launch_jacobi_gather(G, c->stream));
    VT_HIP(c, launch_jacobi_update(G, c->stream));
    c->rad_slot = 3 - c->rad_slot;
  }
  // eAvg (the mean irradiance a body-to-body bounce reads) and the source slot, on the device
  if (int erc = ensure(c, c->eavg_part, 256 * 4 * sizeof(double))) return erc;
  VT_HIP(c, launch_eavg(L, (const float *)c->g_area.p, (float4 *)c->radu.p, c->rad_slot, (double *)c->eavg_part.p,
                        c->stream));
  return VT_OK;
}

extern "C" int vt_radiosity_tables(vt_ctx *c, uint32_t which, void *dst, size_t bytes) {
  VT_LIVE(c);
  if (!c->rad_built) return fail(c, VT_ERR_STATE, "radiosity: not built");
  const size_t V = c->h_verts.size() / 8, A = c->rad_active;
  const DevBuf *src = nullptr;
  size_t need = 0;
  switch (which) {
    case VT_RAD_GATHER: src = &c->g_table; need = V * c->rad_K * 4; break;
    case VT_RAD_MISSES: src = &c->g_miss; need = V * 4; break;
    case VT_RAD_ACTIVE: src = &c->g_active; need = A * 4; break;
    case VT_RAD_NEIGHBOURS: src = &c->g_nbr; need = A * c->rad_nbr_k * 4; break;
    case VT_RAD_WEIGHTS: src = &c->g_nbrw; need = A * c->rad_nbr_k * 4; break;
    case VT_RAD_ALBEDO: src = &c->g_alb; need = (size_t)c->n_mat * ALB_BINS * 32; break;
    case VT_RAD_HH: src = &c->g_hh; need = (size_t)c->n_mat * 16; break;
    default: return fail(c, VT_ERR_ARG, "radiosity_tables: table %u", which);
  }
  if (bytes != need || (need && !dst)) return fail(c, VT_ERR_SIZE, "radiosity_tables: table %u has %zu bytes", which, need);
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  if (need) VT_HIP(c, hipMemcpy(dst, src->p, need, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_radiosity_state(vt_ctx *c, float *node_state, size_t bytes, float e_avg[3]) {
  VT_LIVE(c);
  if (!c->rad_built) return fail(c, VT_ERR_STATE, "radiosity: not built");
  const size_t V = c->h_verts.size() / 8;
  if (node_state && bytes != V * 64) return fail(c, VT_ERR_SIZE, "radiosity_state: %zu bytes for %zu nodes", V * 64, V);
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  if (node_state && V) VT_HIP(c, hipMemcpy(node_state, c->g_state.p, V * 64, hipMemcpyDeviceToHost));
  float4 ru;
  VT_HIP(c, hipMemcpy(&ru, c->radu.p, sizeof ru, hipMemcpyDeviceToHost));
  if (e_avg) {
    e_avg[0] = ru.x;
    e_avg[1] = ru.y;
    e_avg[2] = ru.z;
  }
  return VT_OK;
}

static int dense_alloc(vt_ctx *c, uint32_t n, uint32_t r) {
  const size_t nn = (size_t)n * n, m = (size_t)n * r;
  int rc;
  if ((rc = ensure(c, c->dF32, nn * 4))) return rc;
  if ((rc = ensure(c, c->dE, m * 4))) return rc;
  if ((rc = ensure(c, c->dX, m * 4))) return rc;
  if ((rc = ensure(c, c->dY, m * 4))) return rc;
  if ((rc = ensure(c, c->dR, m * 4))) return rc;
  if ((rc = ensure(c, c->dXd, m * 16))) return rc;      // x64, y64
  if ((rc = ensure(c, c->dPart, RES_BLOCKS_H * 4 * 8))) return rc;
  c->dn_n = n;
  c->dn_r = r;
  c->dn_rp = r;
  release(c->dF16);
  return VT_OK;
}

extern "C" int vt_dense_upload(vt_ctx *c, uint32_t n, uint32_t r, const float *F, const float *e) {
  VT_LIVE(c);
  if (!F || !e || n == 0 || r == 0) return fail(c, VT_ERR_ARG, "dense: null or empty input");
  if (r > 48) return fail(c, VT_ERR_ARG, "dense: at most 48 right-hand sides");
  VT_HIP(c, hipSetDevice(c->device));
  int rc = dense_alloc(c, n, r);
  if (rc) return rc;
  VT_HIP(c, hipMemcpy(c->dF32.p, F, (size_t)n * n * 4, hipMemcpyHostToDevice));
  VT_HIP(c, hipMemcpy(c->dE.p, e, (size_t)n * r * 4, hipMemcpyHostToDevice));
  return VT_OK;
}

extern "C" int vt_dense_formfactors(vt_ctx *c, uint32_t n, const float *nodes, float row_max, uint32_t r,
                                    const float *e) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "dense: form factors need an uploaded scene");
  if (!nodes || !e || n == 0 || r == 0 || r > 48) return fail(c, VT_ERR_ARG, "dense: bad form-factor input");
  VT_HIP(c, hipSetDevice(c->device));
  int rc = dense_alloc(c, n, r);
  if (rc) return rc;
  DevBuf nb;
  if ((rc = upload(c, nb, nodes, (size_t)n * 32))) return rc;
  VT_HIP(c, hipMemcpy(c->dE.p, e, (size_t)n * r * 4, hipMemcpyHostToDevice));
  hipError_t er = launch_formfactors(c->S, (uint32_t *)c->overflow.p, (uint2 *)c->spill.p, c->spill_lanes, n,
                                     (const float4 *)nb.p, (float *)c->dF32.p, c->stream);
  if (er == hipSuccess) er = launch_row_scale((float *)c->dF32.p, n, row_max, c->stream);
  if (er == hipSuccess) er = hipStreamSynchronize(c->stream);
  release(nb);
  if (er != hipSuccess) return hip_fail(c, er, "form factors");
  return VT_OK;
}

extern "C" int vt_dense_download_F(vt_ctx *c, float *F_out, size_t bytes) {
  VT_LIVE(c);
  if (!c->dn_n) return fail(c, VT_ERR_STATE, "dense: nothing uploaded");
  if (!F_out || bytes != (size_t)c->dn_n * c->dn_n * 4) return fail(c, VT_ERR_SIZE, "dense: F is n*n*4 bytes");
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipMemcpy(F_out, c->dF32.p, bytes, hipMemcpyDeviceToHost));
  return VT_OK;
}

namespace {
struct Norms {
  double res2, e2, err2, ref2;
};
int read_norms(vt_ctx *c, Norms &o) {
  double p[RES_BLOCKS_H * 4];
  VT_HIP(c, hipMemcpyAsync(p, c->dPart.p, sizeof p, hipMemcpyDeviceToHost, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  o = {0, 0, 0, 0};
  for (int b = 0; b < RES_BLOCKS_H; b++) {
    o.res2 += p[4 * b];
    o.e2 += p[4 * b + 1];
    o.err2 += p[4 * b + 2];
    o.ref2 += p[4 * b + 3];
  }
  return VT_OK;
}
}  // namespace

extern "C" int vt_dense_solve(vt_ctx *c, const vt_dense_opts *o, float *x_out, vt_dense_result *res) {
  VT_LIVE(c);
  if (!o || !res) return VT_ERR_ARG;
  if (!c->dn_n) return fail(c, VT_ERR_STATE, "dense: nothing uploaded");
  if (o->n != c->dn_n || o->r != c->dn_r) return fail(c, VT_ERR_ARG, "dense: n, r differ from the upload");
  const uint32_t n = c->dn_n, rp = c->dn_rp;
  const uint64_t m = (uint64_t)n * rp;
  const uint32_t maxs = o->max_sweeps ? o->max_sweeps : 1000;
  VT_HIP(c, hipSetDevice(c->device));
  std::memset(res, 0, sizeof *res);
  float *F = (float *)c->dF32.p, *E = (float *)c->dE.p, *X = (float *)c->dX.p, *Y = (float *)c->dY.p,
        *R = (float *)c->dR.p;
  double *X64 = (double *)c->dXd.p, *Y64 = X64 + m, *part = (double *)c->dPart.p;
  int rc;
  Norms nm;
  VT_HIP(c, hipEventRecord(c->ev_s0, c->stream));
  // x transposed (RHS-major) for the MFMA sweeps
  DevBuf xT32;
  if ((rc = ensure(c, xT32, (size_t)rp * n * 4))) return rc;
  float *XT = (float *)xT32.p;
  struct Free {
    DevBuf &b;
    ~Free() { release(b); }
  } free_xT{xT32};
  if (o->method == VT_DENSE_FP32) {
    VT_HIP(c, hipMemcpyAsync(X, E, m * 4, hipMemcpyDeviceToDevice, c->stream));
    for (uint32_t s = 0; s < maxs; s++) {
      VT_HIP(c, launch_sweep_fp32_mfma(F, n, rp, X, XT, E, Y, c->stream));
      VT_HIP(c, launch_resid_f32(Y, X, E, nullptr, m, part, c->stream));
      std::swap(X, Y);
      res->sweeps++;
      if ((rc = read_norms(c, nm))) return rc;
      if (std::sqrt(nm.res2) <= o->tol * std::sqrt(nm.e2)) break;
    }
  } else if (o->method == VT_DENSE_BF16_REFINE) {
    if (n % 128) return fail(c, VT_ERR_ARG, "dense bf16: n must be a multiple of 128");
    if (!c->dF16.p) {
      if ((rc = ensure(c, c->dF16, (size_t)n * n * 2))) return rc;
      VT_HIP(c, launch_convert_bf16(F, (uint16_t *)c->dF16.p, (uint64_t)n * n, c->stream));
    }
    DevBuf xT, D0, D1;
    if ((rc = ensure(c, xT, m * 2))) return rc;
    if ((rc = ensure(c, D0, m * 4))) return rc;
    if ((rc = ensure(c, D1, m * 4))) return rc;
    float *Da = (float *)D0.p, *Db = (float *)D1.p;
    // inner sweeps: a fixed count, or (0) until the correction changes by less than BF16's resolution
    const uint32_t inner = o->inner_sweeps ? o->inner_sweeps : 64;
    const bool adaptive = o->inner_sweeps == 0;
    VT_HIP(c, hipMemsetAsync(X, 0, m * 4, c->stream));
    bool ok = true;
    hipError_t er = hipSuccess;
    while (res->sweeps < maxs) {
      // FP32 residual r = e + F x - x
      if ((er = launch_sweep_fp32_mfma(F, n, rp, X, XT, E, Y, c->stream)) != hipSuccess) break;
      if ((er = launch_sub(Y, X, R, m, c->stream)) != hipSuccess) break;
      if ((er = launch_resid_f32(Y, X, E, nullptr, m, part, c->stream)) != hipSuccess) break;
      res->sweeps++;
      if ((rc = read_norms(c, nm))) { ok = false; break; }
      if (std::sqrt(nm.res2) <= o->tol * std::sqrt(nm.e2)) break;
      // correction d = r + F d by BF16 sweeps from d = r, then x += d
      if ((er = hipMemcpyAsync(Da, R, m * 4, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess) break;
      for (uint32_t k = 0; k < inner && er == hipSuccess; k++) {
        er = launch_x_to_bf16T(Da, n, rp, (uint16_t *)xT.p, c->stream);
        if (er == hipSuccess) er = launch_sweep_bf16_T((const uint16_t *)c->dF16.p, (const uint16_t *)xT.p, n, rp, R, Db, c->stream);
        res->sweeps++;
        std::swap(Da, Db);
        if (adaptive && er == hipSuccess && k >= 1) {
          // |d_new - d_old| against |r|: below 2^-9 (BF16's unit roundoff) the sweeps cannot improve d
          if ((er = launch_resid_f32(Da, Db, R, nullptr, m, part, c->stream)) != hipSuccess) break;
          Norms dn;
          if ((rc = read_norms(c, dn))) { ok = false; break; }
          if (std::sqrt(dn.res2) <= 0.002 * std::sqrt(dn.e2)) break;
        }
      }
      if (!ok || er != hipSuccess) break;
      if ((er = launch_add(X, Da, m, c->stream)) != hipSuccess) break;
      res->refinements++;
    }
    release(xT);
    release(D0);
    release(D1);
    if (er != hipSuccess) return hip_fail(c, er, "bf16 refinement");
    if (!ok) return rc;
  } else if (o->method == VT_DENSE_FP64) {
    VT_HIP(c, launch_widen(E, X64, m, c->stream));
    for (uint32_t s = 0; s < maxs; s++) {
      VT_HIP(c, launch_sweep_fp64(F, n, rp, X64, E, Y64, c->stream));
      VT_HIP(c, launch_resid_f64(Y64, X64, E, nullptr, m, part, c->stream));
      std::swap(X64, Y64);
      res->sweeps++;
      if ((rc = read_norms(c, nm))) return rc;
      if (std::sqrt(nm.res2) <= o->tol * std::sqrt(nm.e2)) break;
    }
  } else {
    return fail(c, VT_ERR_ARG, "dense: unknown method %u", o->method);
  }
  VT_HIP(c, hipEventRecord(c->ev_s1, c->stream));
  VT_HIP(c, hipEventSynchronize(c->ev_s1));
  float ms = 0;
  VT_HIP(c, hipEventElapsedTime(&ms, c->ev_s0, c->ev_s1));
  res->ms = ms;
  res->ms_per_sweep = res->sweeps ? ms / res->sweeps : 0;
  // final residual in FP64: r = e + F x - x with F widened exactly
  double *Xc = X64, *Yc = Y64;
  if (o->method != VT_DENSE_FP64) VT_HIP(c, launch_widen(X, Xc, m, c->stream));
  VT_HIP(c, launch_sweep_fp64(F, n, rp, Xc, E, Yc, c->stream));
  VT_HIP(c, launch_resid_f64(Yc, Xc, E, nullptr, m, part, c->stream));
  if ((rc = read_norms(c, nm))) return rc;
  res->residual = std::sqrt(nm.res2) / std::max(1e-300, std::sqrt(nm.e2));
  if (x_out) {
    if (o->method == VT_DENSE_FP64) {
      std::vector<double> tmp(m);
      VT_HIP(c, hipMemcpy(tmp.data(), Xc, m * 8, hipMemcpyDeviceToHost));
      for (uint64_t i = 0; i < m; i++) x_out[i] = (float)tmp[i];
    } else {
      VT_HIP(c, hipMemcpy(x_out, X, m * 4, hipMemcpyDeviceToHost));
    }
  }
  // keep the solution in dX for later calls
  if (X != (float *)c->dX.p && o->method != VT_DENSE_FP64) VT_HIP(c, hipMemcpy(c->dX.p, X, m * 4, hipMemcpyDeviceToDevice));
  return VT_OK;
}
