// The C ABI, part 2: the frame (renderCamera / renderFrame, gpu/renderer.js:935-996): staging of the camera
// table, the persistent batched and per-camera launches, the frame hipGraph, the wavefront form, the launch-shape
// autotuner; frame statistics and the path-replay adjoint.
#include "vt_ctx.h"

using namespace vt;

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
    key.shutter[0] = k.shutter[0];
    key.shutter[1] = k.shutter[1];
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

}  // namespace

void vt::drop_graph(vt_ctx *c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
  c->gkey.clear();
}

namespace {

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
  uint32_t sort_bits; // 0: the 23-bit radix sort between vertices; 4..13: the bin sort by the key's top bits
  bool graph;         // run the frame's batches as a hipGraph
  uint32_t queues;    // trace queues: 1, or one per XCD
  bool shadow_sort;   // trace the shadow rays in the order of their keys
  bool fit_batch;     // the batch is the default: shrink it to the memory budget instead of failing
};

constexpr uint32_t GRAPH_SORT_BITS = 12;   // a captured wavefront frame bins by this many key bits unless told

WfParams wf_params(const vt_ctx *c, const vt_render_opts &opt) {
  uint32_t sb = opt.sort_bits ? (opt.sort_bits == 1 ? 0u : opt.sort_bits) : c->def_wf_sort;
  if (opt.graph && sb == 0) sb = GRAPH_SORT_BITS;
  return WfParams{opt.batch ? opt.batch : c->def_wf_batch, opt.stack_lds ? opt.stack_lds : c->def_wf_stack,
                  opt.chunk ? std::min(64u, opt.chunk) : c->def_wf_refill, 1, sb, opt.graph != 0,
                  opt.queues ? opt.queues : c->def_wf_queues,
                  opt.shadow_sort ? opt.shadow_sort == 1 : c->def_wf_shadow_sort, opt.batch == 0};
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
  // per path 80 B of state, per queue slot 36 B, per shadow slot 48 B + 16 B of the shadow sort
  auto batch_bytes = [&](uint64_t p) {
    return al(p * 64) + 2 * al(p * 16) + 4 * al(p * 4) + al(p * 16) + al(p * 4) + 3 * al(p * nl * 16) +
           4 * al(p * nl * 4) + al(WF_COUNT_WORDS * 4);
  };
  // the default batch: halved (in multiples of 64) until the buffers fit the budget and the shadow slots 32 bits
  if (wp.fit_batch)
    while (P > (1u << 20) && ((c->wf_budget && batch_bytes(P) > c->wf_budget) || P * nl > (1ull << 32) - 1))
      P = (P / 2 + 63) / 64 * 64;
  const size_t bytes = batch_bytes(P);
  int rc;
  if (P * nl > (1ull << 32) - 1) return fail(c, VT_ERR_ARG, "render: wavefront batch of %llu shadow slots", (unsigned long long)(P * nl));
  if (P > c->wf_paths || nl > c->wf_nl) {
    if ((rc = ensure(c, c->wf, bytes))) return rc;
    const size_t tb = wf_sort_temp_bytes((uint32_t)(P * nl));
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
  W.ps = (float4 *)take(P * 64);
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
  W.sh_key = (uint32_t *)take(P * nl * 4);
  W.sh_idx = (uint32_t *)take(P * nl * 4);
  W.sh_keys_sorted = (uint32_t *)take(P * nl * 4);
  W.sh_q = wp.shadow_sort ? (uint32_t *)take(P * nl * 4) : nullptr;   // null: the shadow slots in slot order
  W.count = (uint32_t *)take(WF_COUNT_WORDS * 4);
  W.xcd_queues = wp.queues;
  float ext[3];
  for (int k = 0; k < 3; k++) ext[k] = std::max(c->bmax[k] - c->bmin[k], 1e-6f);
  W.bmin = make_float3(c->bmin[0], c->bmin[1], c->bmin[2]);
  W.bscale = make_float3(128.0f / ext[0], 128.0f / ext[1], 128.0f / ext[2]);
  W.counters = (unsigned long long *)c->counters.p;
  W.overflow = (uint32_t *)c->overflow.p;
  W.spill = (uint2 *)c->spill.p;
  W.spill_stride = c->spill_lanes;
  const uint32_t slds = wp.slds;
  if (slds != 4 && slds != 8 && slds != 16) return fail(c, VT_ERR_ARG, "render: wavefront stack_lds must be 4, 8 or 16");
  if (wp.queues != 1 && wp.queues != (uint32_t)XCD_COUNT)
    return fail(c, VT_ERR_ARG, "render: queues must be 1 or %d", XCD_COUNT);
  if (wp.sort_bits && (wp.sort_bits < 4 || wp.sort_bits > WF_SORT_BITS_MAX))
    return fail(c, VT_ERR_ARG, "render: wavefront sort_bits must be 1 or 4..%u", WF_SORT_BITS_MAX);
  if (wp.graph && (wp.stride != 1 || wp.sort_bits == 0)) return fail(c, VT_ERR_ARG, "render: a wavefront graph bins");
  W.wave_stride = wp.stride;
  int occ = 0;
  VT_HIP(c, wf_trace_occupancy(slds, c->S.inst_count != 0, &occ));
  const uint32_t grid_cap = (uint32_t)c->cu_count * 8;
  const uint32_t trace_cap = std::min<uint32_t>(max_blocks(c, 256), (uint32_t)std::max(1, occ) * c->cu_count);
  const uint32_t refill = wp.refill;
  const uint64_t batches = (items + P - 1) / P;
  // the frame's work: clear the sums, then every batch (with the extension traces timed when events are given)
  auto enqueue = [&](bool timed) -> int {
    VT_HIP(c, launch_zero_u64((unsigned long long *)c->counters.p, 16, c->stream));
    if (!opt.accumulate)
      for (uint32_t a = 0; a < n; a++) {
        const DevCam &d = c->h_cams[a];
        VT_HIP(c, launch_zero_f4(d.acc, (uint64_t)d.width * d.height, c->stream));
      }
    if (timed) VT_HIP(c, hipEventRecord(c->ev_launch[0], c->stream));
    uint64_t bi = 0;
    for (uint64_t bb = 0; bb < items; bb += P, bi++) {
      W.batch_begin = bb;
      W.n = (uint32_t)std::min<uint64_t>(P, items - bb);
      VT_HIP(c, launch_wavefront_batch(W, depth, grid_cap, trace_cap, refill, slds, wp.sort_bits, keys_sorted, q_sorted,
                                       c->wf_temp.p, c->wf_temp_bytes, W.sh_q != nullptr,
                                       timed ? c->ev_ext.data() + 2 * depth * bi : nullptr, c->stream));
    }
    return VT_OK;
  };
  if (!wp.graph) {
    while (c->ev_ext.size() < 2 * batches * depth) {
      hipEvent_t ev;
      VT_HIP(c, hipEventCreate(&ev));
      c->ev_ext.push_back(ev);
    }
    VT_HIP(c, hipEventRecord(c->ev_r0, c->stream));
    if ((rc = enqueue(true))) return rc;
    c->ext_launches = (uint32_t)(batches * depth);
    VT_HIP(c, hipEventRecord(c->ev_launch[1], c->stream));
    return VT_OK;
  }
  // a graph of the whole frame (kernels only: the fill kernels and the bin sort), captured once and replayed while
  // its launch records, shape and cameras' sums stay the same; its extension traces are not timed one by one
  std::vector<uint8_t> key(1 + sizeof W + 8 * sizeof(uint64_t) + n * 2 * sizeof(uint64_t));
  key[0] = 0x57;    // the wavefront form's graphs, apart from the batched form's
  W.batch_begin = 0;
  W.n = (uint32_t)std::min<uint64_t>(P, items);
  std::memcpy(key.data() + 1, &W, sizeof W);
  const uint64_t sc[8] = {items, P, depth, grid_cap, trace_cap, refill, slds | ((uint64_t)wp.sort_bits << 8),
                          opt.accumulate ? 1u : 0u};
  std::memcpy(key.data() + 1 + sizeof W, sc, sizeof sc);
  for (uint32_t a = 0; a < n; a++) {
    const DevCam &d = c->h_cams[a];
    const uint64_t v[2] = {(uint64_t)(uintptr_t)d.acc, (uint64_t)d.width * d.height};
    std::memcpy(key.data() + 1 + sizeof W + sizeof sc + a * sizeof v, v, sizeof v);
  }
  if (!c->gexec || key != c->gkey) {
    drop_graph(c);
    VT_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int erc = enqueue(false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (erc) {
      if (g) (void)hipGraphDestroy(g);
      return erc;
    }
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    const hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      c->gexec = nullptr;
      return hip_fail(c, ei, "hipGraphInstantiate");
    }
    c->graph_builds++;
    c->gkey = key;
  }
  VT_HIP(c, hipEventRecord(c->ev_r0, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_launch[0], c->stream));
  VT_HIP(c, hipGraphLaunch(c->gexec, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_launch[1], c->stream));
  c->graph_used = true;
  c->ext_launches = 0;
  return VT_OK;
}

}  // namespace

// the estimator of a frame's cameras: all path tracer, or all the reference estimator (which needs the node cache)
int vt::frame_estimator(vt_ctx *c, const uint32_t *cams, uint32_t n, bool &ref) {
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
  if (opt.graph && opt.mode != VT_RENDER_BATCHED && opt.mode != VT_RENDER_WAVEFRONT)
    return fail(c, VT_ERR_ARG, "render: graph needs the batched or the wavefront form");
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
  auto probe_at = [&](const Shape &sh, uint64_t strd, double &mrays) -> int {
    const uint64_t pw = (frame_waves + strd - 1) / strd;
    RenderLaunch L = base_launch(c, sh);
    L.cams = (const DevCam *)c->dcams.p;
    L.ncam = n;
    L.total_waves = pw;
    L.wave_stride = strd;
    L.persistent = 1;
    L.ref = ref ? 1u : 0u;
    int r = persistent_grid(c, sh, pw, L.grid, ref);
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
  auto probe = [&](const Shape &sh, double &mrays) { return probe_at(sh, stride, mrays); };
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
  // the wavefront form on the same probe: LDS stack {4, 8, 16} x refill {8, 16, 32} at the default batch (the
  // probe covers a few percent of the frame, too few paths to show what a larger batch gains, and the gain only
  // grows with the batch); the reference estimator runs the persistent form only
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
    out->wf_sort_bits = c->def_wf_sort;
    out->probe_mrays_wavefront = 0.0;
    return VT_OK;
  }
  // the wavefront form's probes cover eight times more of the frame than the persistent form's (a quarter at the
  // default stride): its rate grows with the paths it sorts together, and on small probes it favours shapes
  // (a 16-entry LDS stack) that lose on the whole frame
  const uint64_t dstride = std::max<uint64_t>(1, stride / 8);
  WfParams wbest{c->def_wf_batch, c->def_wf_stack, c->def_wf_refill, dstride, c->def_wf_sort, false, c->def_wf_queues,
                 c->def_wf_shadow_sort, true};
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
  const uint32_t wstacks[] = {4, 8, 16}, refills[] = {8, 16, 32};
  for (uint32_t st : wstacks)
    for (uint32_t rf : refills) {
      const WfParams wp{wbest.batch, st, rf, dstride, wbest.sort_bits, false, wbest.queues, wbest.shadow_sort, true};
      double m = 0;
      if ((rc = probe_wf(wp, m))) return rc;
      tried++;
      if (m > m_wf) {
        m_wf = m;
        wbest = wp;
      }
    }
  // the shadow rays in slot order or sorted by key
  for (bool ss : {false, true}) {
    WfParams wp = wbest;
    wp.shadow_sort = ss;
    double m = 0;
    if ((rc = probe_wf(wp, m))) return rc;
    tried++;
    if (m > m_wf) {
      m_wf = m;
      wbest = wp;
    }
  }
  // the trace queues at that shape: one head, or one per XCD
  for (uint32_t q : {1u, (uint32_t)XCD_COUNT}) {
    WfParams wp = wbest;
    wp.queues = q;
    double m = 0;
    if ((rc = probe_wf(wp, m))) return rc;
    tried++;
    if (m > m_wf) {
      m_wf = m;
      wbest = wp;
    }
  }
  // then the sort between vertices at that shape: the whole key, or bins of its top 10, 12 or 13 bits
  for (uint32_t sb : {0u, 10u, 12u, 13u}) {
    WfParams wp = wbest;
    wp.sort_bits = sb;
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
  // the form: the persistent winner once more at the wavefront probes' size, against the wavefront winner (the
  // real twin frame: the persistent form ahead on a 1/32 probe, the wavefront 5 % ahead on the whole frame)
  double mb_d = 0;
  if ((rc = probe_at(best, dstride, mb_d))) return rc;
  c->def_wf_batch = wbest.batch;
  c->def_wf_stack = wbest.slds;
  c->def_wf_refill = wbest.refill;
  c->def_wf_sort = wbest.sort_bits;
  c->def_wf_queues = wbest.queues;
  c->def_wf_shadow_sort = wbest.shadow_sort;
  c->def_mode = m_wf > mb_d ? (uint32_t)VT_RENDER_WAVEFRONT : (uint32_t)VT_RENDER_BATCHED;
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
  out->wf_sort_bits = wbest.sort_bits;
  out->probe_mrays_wavefront = m_wf;
  out->probe_fraction = (double)probe_waves / (double)std::max<uint64_t>(1, frame_waves);
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
  if (c->has_textured)
    return fail(c, VT_ERR_ARG, "backward: printed (textured) materials are not differentiated; upload the scene without "
                               "VT_MESH_TEXTURED meshes");
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

