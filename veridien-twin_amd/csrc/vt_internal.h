// Host-side declarations shared by the library's C++ and HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>

#include "vt_device.h"

namespace vt {

// launch shapes for the trace kernels (block threads; LDS stack entries per lane). The render kernel takes
// its LDS depth at run time (STACK_LDS_MIN..STACK_LDS_MAX) and spills to a global area of depth
// STACK_TOTAL - LDS depth, so every configuration holds STACK_TOTAL entries per lane.
constexpr int DEFAULT_BLOCK = 256;
constexpr int STACK_LDS = 16;
constexpr int STACK_SPILL = 32;
constexpr int STACK_TOTAL = 48;
constexpr int STACK_LDS_MIN = 8;
constexpr int STACK_LDS_MAX = 24;
constexpr int SPILL_ENTRIES = STACK_TOTAL - STACK_LDS_MIN;   // depth of the global spill area
constexpr int XCD_COUNT = 8;
constexpr int MAX_CAMS = 64;                                 // cameras per launch (VT_MAX_CAMERAS)
constexpr int CAM_PARAMS = 14;                               // VT_CAMERA_PARAMS                                 // blocks are dispatched round-robin over them

struct RenderLaunch {
  SceneDev S;
  const DevCam *cams;          // device array
  uint32_t ncam;
  uint64_t item_base;          // queue item of wave 0 (a camera's offset when it is launched alone)
  uint64_t total_waves;        // 64-item waves in the queue
  unsigned long long *counters;// [0] queue head, [1] rays traced, [2] samples, [8..15] per-XCD queue heads
  uint32_t *overflow;          // stack overflow counter
  uint2 *spill;                // per-lane spill stacks
  uint64_t spill_stride;       // lanes with a spill slot (persistent grid size)
  uint32_t chunk;              // waves per queue grab
  uint32_t persistent;         // 1: waves take chunks from the atomic queue head; 0: static grid stride
  uint32_t block;              // threads per block (64, 128, 256)
  uint32_t grid;               // blocks
  uint32_t stack_lds;          // LDS stack entries per lane (STACK_LDS_MIN..STACK_LDS_MAX)
  uint32_t queues;             // persistent: 1 global queue, or XCD_COUNT queues over contiguous eighths of
                               // the waves (a block drains its XCD's eighth, then the others)
  uint64_t wave_stride;        // queue wave w runs frame wave w * wave_stride (1; > 1 for tuning probes)
  uint32_t ref;                // 1: the cameras use the reference estimator (VT_OPT_REFERENCE)
  float4 *ltot;                // adjoint: each sample's radiance by queue item, written by the camera walk and
                               // read by the gradient walk in place of its own first walk (null: walk it)
};
hipError_t launch_render(const RenderLaunch &L, hipStream_t s);
// the reference estimator's refinement pass: each camera marks its pixels into one list (cam: the camera's device
// record, c its index in L.cams), then one launch refines every listed pixel (R: the frame's largest refine); and
// the first pass's key reset
hipError_t launch_refine_mark(const DevCam *cam, uint32_t c, uint32_t width, uint32_t height, uint2 *list,
                              uint32_t *count, hipStream_t s);
hipError_t launch_refine(const RenderLaunch &L, const uint2 *list, const uint32_t *count, uint64_t max_items,
                         uint32_t R, hipStream_t s);
hipError_t launch_fill_keys(uint4 *keys, uint64_t n, hipStream_t s);
hipError_t launch_fill_u4(uint4 *p, uint64_t n, uint4 v, hipStream_t s);
hipError_t render_occupancy(uint32_t block, uint32_t stack_lds, bool inst, bool ref, int *blocks_per_cu);
// path-replay adjoint: gradients of sum_p adj_p . I_p with respect to material colours and light radiances
// (block 256, persistent queue; L.cams[*].adj set)
hipError_t launch_backward(const RenderLaunch &L, float *gmat, float *glight, uint32_t n_mat, uint32_t n_light,
                           hipStream_t s);
// camera adjoint: CAM_PARAMS gradients per camera of the launch (fx fy cx cy k1..k4, rotation, centre), by
// forward-mode tangents of each sample's camera ray (block 256, persistent queue; L.cams[*].adj set)
hipError_t launch_camgrad(const RenderLaunch &L, float *gcam, hipStream_t s);
// the wavefront form: one batch of paths (items [batch_begin, batch_begin + n) of the frame queue, n a multiple
// of 64), one kernel per stage per vertex, live paths radix-sorted by ray_key between vertices
constexpr uint32_t WF_COUNT_WORDS = 32;   // WfLaunch::count
struct WfLaunch {
  SceneDev S;
  const DevCam *cams;
  uint32_t ncam;
  uint64_t batch_begin;        // in items of the (strided) queue
  uint32_t n;
  uint32_t bounce;
  uint64_t wave_stride;        // queue wave w is frame wave w * wave_stride (1; > 1 for tuning probes)
  uint32_t nl_max;             // shadow slots per queue slot (the largest light_samples of the frame)
  // path state, by path id (0..n): one 64-B record per path, so a path's state is one line wherever the sorted
  // queue sends it: [0] next ray origin, w: pdf of the BSDF sample that chose it; [1] next ray direction, w: the
  // sample's time in the exposure; [2] throughput, w: alive (the shade kernel moves it past the vertex's
  // continuation); [3] x: camera of the path (index in cams, bits)
  float4 *ps;
  float4 *L;                   // radiance so far
  // queue of live paths, by queue slot (0..count[0])
  const uint32_t *q;           // path ids (null: identity, vertex 1)
  uint32_t *q_next, *key_next; // next vertex: path id and sort key per slot (DEAD key: ended)
  uint32_t *count;             // [0] live slots of this vertex, [1] of the next, [2] [3] trace queue heads,
                               // [8..15] [16..23] the per-XCD heads of the extension and shadow traces
  uint32_t xcd_queues;         // trace queues: 1 (one head) or 8 (one per XCD over contiguous eighths)
  float4 *hit;                 // t, u, v, triangle (bits) per slot
  float4 *thr_slot;            // the throughput at this vertex per slot (weight of its NEE in the gather)
  uint32_t *nee;               // shadow slots used per queue slot
  float4 *sh_o, *sh_d, *sh_c;  // shadow slots [slot * nl_max + j]: origin + tmax, direction, contribution + visible
  // the shadow sort (sh_q non-null): per shadow slot its ray's key (the dead key for an unused slot) and its own
  // index, and the slot indices sorted by key; count[4] = the live shadow slots, traced in that order
  uint32_t *sh_key, *sh_idx, *sh_keys_sorted, *sh_q;
  float3 bmin, bscale;         // scene box and 128 / extent, for the origin's Morton code
  unsigned long long *counters;// [1] rays, [2] samples, [3] extension rays, [4] shadow rays
  uint32_t *overflow;
  uint2 *spill;
  uint64_t spill_stride;
};
// ev_ext (optional): 2 * max_depth events, recorded around each extension-trace launch of the batch
size_t wf_sort_temp_bytes(uint32_t n);
hipError_t wf_trace_occupancy(uint32_t slds, bool inst, int *blocks_per_cu);
// trace_grid_cap: resident blocks of the persistent trace kernels; refill: idle lanes of a wave that trigger a
// fetch of new rays (1..64); slds: LDS stack entries per lane of the trace kernels (4, 8 or 16)
// sort_bits: 0 sorts the slots between vertices by their whole 23-bit key (rocPRIM's radix sort, whose memset
// nodes keep it out of graphs); 4..WF_SORT_BITS_MAX bins them by the key's top bits (kernels only)
constexpr uint32_t WF_SORT_BITS_MAX = 13;
hipError_t launch_wavefront_batch(WfLaunch W, uint32_t max_depth, uint32_t grid_cap, uint32_t trace_grid_cap,
                                  uint32_t refill, uint32_t slds, uint32_t sort_bits, uint32_t *keys_sorted,
                                  uint32_t *q_sorted, void *sort_temp, size_t sort_temp_bytes, bool sort_shadows,
                                  hipEvent_t *ev_ext, hipStream_t s);
hipError_t launch_zero_u64(unsigned long long *p, uint32_t n, hipStream_t s);
hipError_t launch_copy_f4(const float4 *x, float4 *y, uint64_t n, uint32_t grid, hipStream_t s);
hipError_t launch_read_u4(const uint4 *x, uint64_t n, uint32_t grid, uint32_t *sink, hipStream_t s);
// shading records of every triangle slot (SceneDev::tshade) from the slot's vertices, vertex ids and normals
hipError_t launch_tri_shading(const float4 *tri, const uint4 *attr, const float4 *vnorm, uint32_t T, float4 *out,
                              hipStream_t s);
hipError_t launch_zero_f4(float4 *p, uint64_t n, hipStream_t s);

struct TestLaunch {
  SceneDev S;
  uint32_t *overflow;
  uint2 *spill;
  uint64_t spill_stride;   // lanes with a spill slot; the grid never exceeds it
  uint32_t max_blocks;     // spill_stride / DEFAULT_BLOCK
};
hipError_t launch_trace_test(const TestLaunch &L, const float *rays, uint32_t n, int any_hit, float *hits,
                             hipStream_t s);
hipError_t launch_lens_test(const float *lens9, const float *uv, uint32_t n, float *rays, hipStream_t s);
hipError_t launch_visible(const TestLaunch &L, const float c[3], uint32_t skip, float tol, float time, const float *pts,
                          uint32_t n, uint8_t *out, hipStream_t s);
hipError_t launch_camray_test(const DevCam *cam, const uint32_t *xyi, uint32_t n, float *out, hipStream_t s);
hipError_t launch_bsdf_test(const float4 *mats, const float *cases, uint32_t n, float *out, hipStream_t s);
hipError_t launch_capsule_test(const float *cases, uint32_t n, float *out, hipStream_t s);

// the fused sensor epilogue: radiance = acc.rgb * scale / acc.a (the pixel's sample count), RI map, -> DN
struct SensorDev {
  uint32_t w, h, cfa, frame, seed, cam_index;
  float optics_k, esf_sigma, exposure, gain;
  float m[9];
  float read_noise, full_well, dark, conv_gain, max_dn, black;
  uint32_t noise_on;
  float scale;
  float glare_k, glare_sigma;   // veiling glare: fraction (0 = off), sigma (level px)
  uint32_t glare_f;             // level factor F
};
// glare: two level buffers of ceil(w/F) x ceil(h/F) float4 (unused when glare_k == 0)
hipError_t launch_sensor(const SensorDev &P, const float4 *acc, const float *ri, float4 *glare_a, float4 *glare_b,
                         uint16_t *dn, hipStream_t s);
// the raw stage on caller device buffers with its glare levels in the library's static device array (vt_sensor.hip):
// sensor_scratch_init resolves the array's address for the current device (once per device, at vt_create);
// launch_sensor_noalloc refuses a glare level above GLARE_LEVEL_MAX float4 or a device not initialised
constexpr uint32_t GLARE_LEVEL_MAX = 262144;
constexpr int MAX_DEVICES = 64;
hipError_t sensor_scratch_init(int device);
hipError_t launch_sensor_noalloc(int device, const SensorDev &P, const float4 *acc, const float *ri, uint16_t *dn,
                                 hipStream_t s);
hipError_t launch_pack_rgb8(const uchar4 *rgba, uint64_t n, uint8_t *rgb, hipStream_t s);
hipError_t launch_finalize_linear(const float4 *acc, float scale, uint64_t n, float4 *out, hipStream_t s);
// the ISP after the raw DN
struct IspDev {
  uint32_t w, h, cfa, demosaic, nr, tone, chroma420;
  float max_dn, black, read_noise, conv_gain, gain;
  float wb[3], ccm[9];
  float nr_sigma, nr_scale, gamma, s_c, black_off, sharpen_amount;
  int sharpen_rad;
};
// stats: [0] luma sum, [1..3] grey-world sums (x 1023); yuv: Y plane w*h, then Cb and Cr ((w+1)/2 x (h+1)/2);
// rgb: w*h linear RGB; part: isp_partials(w, h) block partial sums
uint32_t isp_partials(uint32_t w, uint32_t h);
hipError_t launch_isp(const IspDev &P, const uint16_t *dn, float4 *rgb, uint4 *part, uchar4 *rgba, uint8_t *yuv,
                      unsigned int *stats, hipStream_t s);

// node radiosity
struct GatherDev {
  uint32_t n_active, gather_k, nbr_k, bins, slot_in, slot_out;
  const uint32_t *active, *gtable, *node_mat, *nbr;
  const float4 *normals, *albedo, *hh;
  const float *nbr_w;
  float4 *node_state;
  float4 *aux;
  // per (active node, gather entry): the static weight A_i(mu_in) A_hit(mu_out) (rgb, 0 for a miss) and the hit
  // node id (w, as bits); null: each step reads the table and computes the weight
  float4 *gw;
};
// the node radiosity cache on the uploaded scene (vt_nodes.hip)
struct NodeLaunch {
  SceneDev S;
  const float4 *npos;        // per vertex: position, w = 1 for a static node
  const float4 *nnorm;       // per vertex: unit normal
  const uint32_t *nmat;      // per vertex: material
  uint32_t n_nodes, K, node_samples, n_active, n_mat;
  uint32_t *gtable;          // n_nodes * K
  float *miss;               // per vertex: escaping gather rays (K for a vertex that is no static node)
  const uint32_t *active;
  float4 *state;             // 4 per vertex
  float4 *alb, *hh;          // albedo tables (ALB_BINS x 2 per material), hemispherical albedo
  uint32_t *overflow;
  uint2 *spill;
  uint64_t spill_stride;
  uint32_t max_blocks;       // spill_stride / 256
  // k_node_direct's incremental pass: bit b = body b moved (0: every active node); pcaps: per body, the
  // bounding capsule (a.xyz + radius, b.xyz) at the last direct pass
  uint32_t moved;
  const float4 *pcaps;
};
hipError_t launch_albedo_tables(const NodeLaunch &L, hipStream_t s);
hipError_t launch_fill_f4(float4 *p, uint32_t n, float4 v, hipStream_t s);    // a fill kernel (no memset nodes)
// part: 256 x 4 doubles of scratch for the block partials
hipError_t launch_eavg(const NodeLaunch &L, const float *area, float4 *radu, uint32_t slot, double *part,
                       hipStream_t s);
hipError_t launch_gather_build(const NodeLaunch &L, hipStream_t s);
hipError_t launch_gather_count(const NodeLaunch &L, hipStream_t s);
hipError_t launch_gather_remap(const NodeLaunch &L, hipStream_t s);
hipError_t launch_node_direct(const NodeLaunch &L, hipStream_t s);

hipError_t launch_jacobi_gather(const GatherDev &G, hipStream_t s);
hipError_t launch_jacobi_weights(const GatherDev &G, hipStream_t s);
hipError_t launch_jacobi_update(const GatherDev &G, hipStream_t s);

// the lite renderer's baked node table on the device (vt_baked.hip)
struct BakedDev {
  uint32_t nodes, K, n_caps;
  const uint16_t *gather;    // nodes x K hit ids, 0xffff: escaped
  const float4 *node_a;      // position, area
  const float4 *node_n;      // normal, mesh (bits)
  const float4 *hh;          // per mesh: hemispherical albedo
  const float4 *caps;        // 2 per capsule: a.xyz + radius, b.xyz + rho
  uint32_t *occ;             // nodes x K transmission bytes, 4 per word
  const float4 *e_dir;       // per node: the reflected direct light E_dir A(mu_l)
  float4 *e_ind;             // per node: the last step's indirect irradiance
};
hipError_t launch_baked_occ(const BakedDev &B, hipStream_t s);
hipError_t launch_baked_jacobi(const BakedDev &B, const float4 *b_in, float4 *b_out, hipStream_t s);

// dense node radiosity: x, y, e are n x rp (row-major, RHS fastest); K split four ways, sums added into y
// FP32 on MFMA (v_mfma_f32_16x16x4_f32): xT is scratch of rp * n floats; n a multiple of 16, rp <= 48
hipError_t launch_sweep_fp32_mfma(const float *F32, uint32_t n, uint32_t rp, const float *x, float *xT,
                                  const float *e, float *y, hipStream_t s);
hipError_t launch_sweep_fp64(const float *F32, uint32_t n, uint32_t rp, const double *x, const float *e, double *y,
                             hipStream_t s);
// BF16 F on MFMA with FP32 accumulation; xT is x rounded to BF16 and transposed (rp x n); n % 128 == 0
hipError_t launch_sweep_bf16_T(const uint16_t *F16, const uint16_t *xT, uint32_t n, uint32_t rp, const float *e,
                               float *y, hipStream_t s);
hipError_t launch_x_to_bf16T(const float *x, uint32_t n, uint32_t rp, uint16_t *xT, hipStream_t s);
hipError_t launch_convert_bf16(const float *src, uint16_t *dst, uint64_t n, hipStream_t s);
// per-block partial sums (RES_BLOCKS x 4): sum (y - x)^2, sum e^2, sum (x - ref)^2, sum ref^2
constexpr int RES_BLOCKS_H = 256;
hipError_t launch_resid_f32(const float *y, const float *x, const float *e, const double *ref, uint64_t m,
                            double *part, hipStream_t s);
hipError_t launch_resid_f64(const double *y, const double *x, const float *e, const double *ref, uint64_t m,
                            double *part, hipStream_t s);
hipError_t launch_widen(const float *x, double *y, uint64_t m, hipStream_t s);
hipError_t launch_add(float *a, const float *b, uint64_t m, hipStream_t s);
hipError_t launch_sub(const float *y, const float *x, float *r, uint64_t m, hipStream_t s);
hipError_t launch_formfactors(const SceneDev &S, uint32_t *overflow, uint2 *spill, uint64_t spill_stride,
                              uint32_t n, const float4 *nodes, float *F, hipStream_t s);
hipError_t launch_row_scale(float *F, uint32_t n, float row_max, hipStream_t s);

// ------------------------------------------------------------------ host BVH utilities (vt_bvh.cpp)
// the reference's binned-SAH BVH2 (gpu/bvh.js:18-161); returns node count, or -1 if cap_nodes is too small
int64_t build_bvh2(const float *pos, const uint32_t *idx, uint32_t ntri, int leaf, int bins, float *nodes,
                   uint32_t cap_nodes, uint32_t *order);
// BVH2 (reference layout) -> 128-B four-wide nodes; throws std::runtime_error on a malformed tree
std::vector<Node4> collapse_bvh4(const float *nodes2, uint32_t node_count, uint32_t tri_count);

}  // namespace vt
