// The wavefront form of the camera path tracer: a batch of paths advances one vertex at a time through
// separate kernels (extension trace, shade, shadow trace, NEE gather), and between vertices the live paths are
// re-ordered by a radix sort on (direction cell, Morton code of the ray origin), so a wave traces rays that
// head the same way and start close together. The estimator, its random numbers and the order of every sum
// are those of path_walk (vt_render.hip); the per-pixel sums use the same butterfly over 64-item waves.
#include <rocprim/device/device_radix_sort.hpp>

#include "vt_internal.h"

namespace vt {

namespace {

constexpr uint32_t DEAD_KEY = 0xffffffffu;
// sort key: KEY_DIR_BITS bits per axis of the direction's octahedral cell, then KEY_MORTON_BITS bits per axis of
// the origin's Morton code; the dead key's bit KEY_BITS puts ended paths last, so a radix sort over KEY_BITS + 1
// bits orders everything. 16 x 16 cells and 5 bits per axis make 24 sorted bits, three radix passes instead of
// four: on the full-scale frame 1.4 % more throughput than 7 bits per axis (finer origin order traced 1-4 %
// faster, sorted 29 ms per frame slower), and ahead of 32 x 32 cells with 4 or 5 bits (2.1 %, 3.9 %), 8 x 8 with 6
// (3.1 %) and 64 x 64 with 4 (7.8 %).
constexpr uint32_t KEY_DIR_BITS = 4, KEY_MORTON_BITS = 5;
constexpr uint32_t KEY_BITS = 2 * KEY_DIR_BITS + 3 * KEY_MORTON_BITS;

// the sample of frame item `item`, given its camera
struct ItemPix {
  uint32_t x, y, i;
  bool valid;
};
__device__ __forceinline__ ItemPix item_pix(const DevCam &C, uint64_t item) {
  ItemPix r;
  const uint64_t q = item - C.item_begin;
  const uint64_t pix = q / C.nspp;
  r.i = C.spp_begin + (uint32_t)(q - pix * C.nspp);
  const uint32_t t = (uint32_t)(pix >> 6), s = (uint32_t)(pix & 63u);
  r.x = (t % C.tiles_x) * 8u + (s & 7u);
  r.y = (t / C.tiles_x) * 8u + (s >> 3);
  r.valid = r.x < C.width && r.y < C.height;
  return r;
}

// frame item of path p: batch items are 64-item waves of the frame queue, every wave_stride-th one (1: all)
__device__ __forceinline__ uint64_t path_item(const WfLaunch &W, uint32_t p) {
  const uint64_t q = W.batch_begin + p;
  return (q >> 6) * W.wave_stride * 64ull + (q & 63ull);
}

__device__ __forceinline__ uint32_t spread3(uint32_t v) {   // 9 bits -> every third bit of 27
  v &= 0x1ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
// sort key of a ray: its direction's cell in a 16 x 16 octahedral map (8 bits), then the Morton code of its origin
// in the scene box at KEY_MORTON_BITS bits per axis: rays of a wave head the same way first, start nearby second
// (direction-major keys traced 8 % faster than origin-major ones, and 16 x 16 cells than 2 x 2 x 2 octants)
__device__ __forceinline__ uint32_t ray_key(v3 o, v3 d, const WfLaunch &W) {
  const float s = 1.0f / (fabsf(d.x) + fabsf(d.y) + fabsf(d.z));
  float px = d.x * s, py = d.y * s;
  if (d.z < 0.0f) {
    const float qx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
    const float qy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
    px = qx;
    py = qy;
  }
  const float cells = (float)(1u << KEY_DIR_BITS);
  const uint32_t cx = (uint32_t)fminf(fmaxf((px * 0.5f + 0.5f) * cells, 0.0f), cells - 1.0f);
  const uint32_t cy = (uint32_t)fminf(fmaxf((py * 0.5f + 0.5f) * cells, 0.0f), cells - 1.0f);
  const float fx = fminf(fmaxf((o.x - W.bmin.x) * W.bscale.x, 0.0f), 127.0f);
  const float fy = fminf(fmaxf((o.y - W.bmin.y) * W.bscale.y, 0.0f), 127.0f);
  const float fz = fminf(fmaxf((o.z - W.bmin.z) * W.bscale.z, 0.0f), 127.0f);
  const uint32_t m = spread3((uint32_t)fx) | (spread3((uint32_t)fy) << 1) | (spread3((uint32_t)fz) << 2);
  return ((cy << KEY_DIR_BITS | cx) << (3 * KEY_MORTON_BITS)) | (m >> (3 * (7 - KEY_MORTON_BITS)));
}

__device__ __forceinline__ uint32_t wave_count(bool pred) {
  return (uint32_t)__popcll(__ballot(pred));
}

// ---------------------------------------------------------------- kernels
__global__ __launch_bounds__(256) void k_wf_gen(WfLaunch W) {
  if (blockIdx.x == 0 && threadIdx.x < WF_COUNT_WORDS) W.count[threadIdx.x] = threadIdx.x == 0 ? W.n : 0u;
  if (W.sh_q)
    for (uint32_t s = blockIdx.x * 256 + threadIdx.x; s < W.n * W.nl_max; s += gridDim.x * 256) W.sh_idx[s] = s;
  for (uint32_t p = blockIdx.x * 256 + threadIdx.x; p < W.n; p += gridDim.x * 256) {
    const uint64_t item = path_item(W, p);
    uint32_t c = 0;
    while (c + 1 < W.ncam && item >= W.cams[c + 1].item_begin) c++;
    const DevCam &C = W.cams[c];
    const ItemPix ip = item_pix(C, item);
    v3 d = mk(0.f, 0.f, 1.f);
    float time = 0.0f;
    if (ip.valid) {
      PixSeeds P;
      pix_seeds(C, ip.x, ip.y, P);
      d = camera_dir(C, P, ip.x, ip.y, ip.i);
      time = sample_time(C, P, ip.i);
    }
    float4 *ps = W.ps + 4 * (size_t)p;
    ps[0] = make_float4(C.pos[0], C.pos[1], C.pos[2], 0.0f);     // w: pdf of the BSDF sample that led here
    ps[1] = make_float4(d.x, d.y, d.z, time);                     // w: the sample's time in the exposure
    ps[2] = make_float4(1.f, 1.f, 1.f, ip.valid ? 1.0f : 0.0f);  // w: 1 while the path is alive
    ps[3] = make_float4(__uint_as_float(c), 0.f, 0.f, 0.f);
    W.L[p] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// Persistent trace with dynamic ray fetch (Aila and Laine, HPG 2009): each lane holds one ray of the queue and
// walks it with the while-while loop of traverse(); when at least `refill` lanes of the wave have finished, the
// wave takes that many new rays from the queue head with one atomic and starts them, so lanes do not idle until
// the slowest ray of a fixed group ends. ANY = false: extension rays of the live slots (count[0], head count[2]),
// result into hit[slot]; ANY = true: shadow slots (count[0] * nl_max, head count[3]), visibility into sh_c.w.
// SLDS: stack entries per lane kept in LDS (the rest of the STACK_TOTAL spill to global memory); a shallow
// LDS stack lets more waves share a CU, which hides the node fetches better than the deeper stack saves.
// Built for 6 waves per SIMD (at most 80 VGPRs): 7 and 8 spill and ran slower. The body-walking build (posed
// bodies: the instance transforms on top) spills some 50 registers at that bound and still ran the real twin frame
// 22 % faster than at its natural 135 VGPRs and 3 waves (1,637 -> 2,003 Mrays/s; 4 and 5 waves 1,859 and 1,954,
// 7 and 8 waves 1,504 and 1,317); a 16-entry LDS stack keeps it at 3 waves (six blocks do not fit the LDS).
template <bool ANY, int SLDS, bool INST>
__global__ __launch_bounds__(256, 6) void k_wf_trace_dyn(WfLaunch W, uint32_t refill) {
  __shared__ uint32_t s_node[SLDS * 256];
  __shared__ float s_t[SLDS * 256];
  Stack st{(lds_u32 *)s_node, (lds_f32 *)s_t, W.spill, threadIdx.x, 256, SLDS, STACK_TOTAL - SLDS, W.spill_stride,
           (uint64_t)blockIdx.x * 256 + threadIdx.x, W.overflow};
  const uint32_t lane = threadIdx.x & 63u;
  // shadow rays: every slot of the vertex in slot order, or (sh_q) the count[4] live ones in key order
  const uint32_t n = ANY ? (W.sh_q ? W.count[4] : W.count[0] * W.nl_max) : W.count[0];
  uint32_t *head = &W.count[ANY ? 3 : 2];
  // per-XCD queues (xcd_queues = 8): queue x is the x-th contiguous eighth of the rays, drained first by the
  // blocks of XCD x (consecutive blocks go to consecutive XCDs), so an XCD's L2 holds the nodes of fewer
  // direction cells; a wave whose queue is used up moves on to the next one
  const uint32_t nq = W.xcd_queues;
  uint32_t *qhead = &W.count[ANY ? 16 : 8];
  const uint32_t seg = (n + nq - 1) / nq;
  uint32_t qi = blockIdx.x % nq, tried = 0;   // lane 0's queue and how many it found used up

  bool active = false, exhausted = false;
  uint32_t k = 0;
  // the lane's ray and traversal state
  const __amdgpu_buffer_rsrc_t NR = node_rsrc(W.S.nodes);
  Slab B;
  B.inv = mk(0.f, 0.f, 0.f); B.oi = mk(0.f, 0.f, 0.f); B.near = 0x402000u;
  RayW R;
  R.o = mk(0.f, 0.f, 0.f); R.kx = 0; R.ky = 1; R.kz = 2; R.Sx = 0.f; R.Sy = 0.f; R.Sz = 0.f;
  float best = 0.0f, tmax0 = 0.0f, time = 0.0f;
  uint32_t skip = 0;   // bodies this ray ignores (posed-body kernels only)
  v3 dir = mk(0.f, 0.f, 1.f);
  bool found = false;
  Hit h{0.f, 0.f, 0.f, 0u, NO_INST};
  uint32_t sp = 0, ref = TRAV_DONE, leaf = TRAV_DONE;
  uint32_t nrays = 0;
  // start lane's ray `kk` (a shadow ray through the sorted indices when sh_q): its slot in k, its traversal state
  auto start_ray = [&](uint32_t kk) {
    k = kk;
    v3 o, d;
    float tmax;
    bool ok;
    if (ANY) {
      if (W.sh_q) {
        k = W.sh_q[k];    // the slot the result belongs to
        ok = true;
      } else {
        ok = k % W.nl_max < W.nee[k / W.nl_max];
      }
      const float4 o4 = W.sh_o[k], d4 = W.sh_d[k];
      o = mk(o4.x, o4.y, o4.z);
      d = mk(d4.x, d4.y, d4.z);
      tmax = o4.w;
      time = d4.w;
      if (INST) skip = W.S.shadow_skip;
    } else {
      const uint32_t p = W.q ? W.q[k] : k;
      const float4 *ps = W.ps + 4 * (size_t)p;
      ok = ps[2].w > 0.0f;
      const float4 o4 = ps[0], d4 = ps[1];
      o = mk(o4.x, o4.y, o4.z);
      d = mk(d4.x, d4.y, d4.z);
      tmax = INFINITY;
      time = d4.w;
      if (INST) skip = W.bounce == 0 ? W.cams[__float_as_uint(ps[3].x)].skip : 0u;
      if (!ok) W.hit[k] = make_float4(0.f, 0.f, 0.f, __uint_as_float(MISS));
    }
    if (ok) {
      B = make_slab(o, d);
      R = ray_w(o, d);
      dir = d;
      best = tmax;
      tmax0 = tmax;
      found = false;
      sp = 0;
      ref = W.S.root;
      leaf = TRAV_DONE;
      active = true;
      nrays++;
    }
  };
  for (;;) {
    // refill: the wave's idle lanes take consecutive queue entries
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (!exhausted && nidle >= refill) {
      uint32_t base = 0, cnt = 0;
      if (lane == 0) {
        if (nq == 1) {
          base = atomicAdd(head, nidle);
          cnt = base < n ? min(nidle, n - base) : 0u;
          if (base + nidle >= n) tried = 1;
        } else {
          while (tried < nq) {
            const uint32_t s0 = qi * seg, s1 = min(n, s0 + seg);
            const uint32_t b = s0 < s1 ? atomicAdd(&qhead[qi], nidle) : 0u;
            if (s0 + b < s1) {
              base = s0 + b;
              cnt = min(nidle, s1 - base);
              break;
            }
            qi = qi + 1 == nq ? 0u : qi + 1;
            tried++;
          }
        }
      }
      base = __shfl(base, 0, 64);
      cnt = __shfl(cnt, 0, 64);
      if (__shfl(tried, 0, 64) >= nq || cnt == 0) exhausted = true;
      if (!active) {
        const uint32_t r = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
        k = base + r;
        if (r < cnt) start_ray(k);
      }
    }
    if (!__any(active)) {
      if (exhausted) break;
      continue;
    }
    // traverse until this lane's ray ends, or until `refill` lanes of the wave are idle
    if (active) {
      for (;;) {
        while (ref != TRAV_DONE && !(ref & LEAF_BIT)) {
          float tn[4];
          uint32_t rr[4];
          node4_test(NR, ref, B, 0.0f, best * 1.00000024f, tn, rr);
          if (tn[0] < INFINITY) {
            if (tn[3] < INFINITY) push(st, sp, rr[3], tn[3]);
            if (tn[2] < INFINITY) push(st, sp, rr[2], tn[2]);
            if (tn[1] < INFINITY) push(st, sp, rr[1], tn[1]);
            ref = rr[0];
          } else {
            ref = pop_next(st, sp, best);
          }
          if (is_leaf_ref(ref) && leaf == TRAV_DONE) {
            leaf = ref;
            ref = pop_next(st, sp, best);
          }
          if (!__any(leaf == TRAV_DONE)) break;
        }
        while (leaf != TRAV_DONE) {
          leaf_test<ANY>(W.S.tri, R, leaf, 0.0f, best, h, found);
          if (ANY && found) {
            leaf = TRAV_DONE;
            ref = TRAV_DONE;
            break;
          }
          leaf = TRAV_DONE;
          if (is_leaf_ref(ref)) {
            leaf = ref;
            ref = pop_next(st, sp, best);
          }
        }
        const bool done = ref == TRAV_DONE && leaf == TRAV_DONE;
        if (done) {
          h.inst = NO_INST;
          // then the posed bodies (few, small: walked one after the other with the lane's own stack)
          if (INST && W.S.inst_count && !(ANY && found)) trace_instances<ANY>(W.S, R.o, dir, 0.0f, tmax0, time, skip, h, found,
                                                                                  st);
          if (ANY) {
            if (!found) W.sh_c[k].w = 1.0f;
          } else {
            // the hit's body rides in the triangle word's top bits (slots are < 2^26)
            const uint32_t bits = h.tri | (h.inst == NO_INST ? 0u : (h.inst + 1u) << 26);
            W.hit[k] = found ? make_float4(h.t, h.u, h.v, __uint_as_float(bits))
                             : make_float4(0.f, 0.f, 0.f, __uint_as_float(MISS));
          }
          active = false;
        }
        if (!active || (uint32_t)__popcll(__ballot(!active)) >= refill) break;
      }
    }
  }
  for (int off = 32; off >= 1; off >>= 1) nrays += __shfl_xor(nrays, off, 64);
  if (lane == 0 && nrays) {
    atomicAdd(&W.counters[1], (unsigned long long)nrays);
    atomicAdd(&W.counters[ANY ? 4 : 3], (unsigned long long)nrays);
  }
}

// one vertex of path_walk<false>: emitter hit, NEE samples (their shadow rays left to k_wf_shadow), continuation
// (112-125 VGPRs, 4 waves per SIMD: bounding it to 5, 6 or 8 waves spills and ran 5 %, 8 % and 12 % slower)
template <bool INST>
__global__ __launch_bounds__(256) void k_wf_shade(WfLaunch W) {
  // every slot of the batch gets a key: live paths of this vertex theirs, the rest (and the slots past the
  // live count) the dead key, so the sort of all n slots puts exactly the next vertex's paths first
  const uint32_t n = W.count[0];
  const uint32_t b = W.bounce;
  const float A = W.S.emit_area;
  uint32_t wave_live = 0;   // the wave's live paths of the next vertex, over the grid-stride loop
  uint32_t lane_nee = 0;    // the lane's shadow slots written (the shadow sort's live count)
  for (uint32_t k0 = blockIdx.x * 256; k0 < W.n; k0 += gridDim.x * 256) {
    const uint32_t k = k0 + threadIdx.x;
    bool live = false;
    uint32_t nee_k = 0;   // shadow slots this slot wrote
    if (k >= n && k < W.n) {
      W.key_next[k] = DEAD_KEY;
      W.q_next[k] = k;
    }
    if (k < n) {
      const uint32_t p = W.q ? W.q[k] : k;
      float4 *ps = W.ps + 4 * (size_t)p;
      const float4 ps3 = ps[3];
      const DevCam &C = W.cams[__float_as_uint(ps3.x)];
      const uint32_t nl = C.nl < 1 ? 1 : (C.nl > 16 ? 16 : C.nl);
      const float fnl = (float)nl;
      const uint32_t Dp = C.depth < 1 ? 1 : C.depth;
      uint32_t nee = 0;   // shadow slots written
      const float4 h4 = W.hit[k];
      const uint32_t bits = __float_as_uint(h4.w);
      const uint32_t tri = bits == MISS ? MISS : (bits & 0x03ffffffu);
      const uint32_t inst = bits == MISS ? NO_INST : (bits >> 26) - 1u;   // NO_INST when the top bits are 0
      if (tri != MISS && b < Dp) {
        const ItemPix ip = item_pix(C, path_item(W, p));
        PixSeeds P;
        pix_seeds(C, ip.x, ip.y, P);
        const uint32_t i = ip.i, hs = P.hs;
        const float4 o4 = ps[0], d4 = ps[1];
        const v3 o = mk(o4.x, o4.y, o4.z), d = mk(d4.x, d4.y, d4.z);
        const float pB = o4.w, time = d4.w;
        const float4 t4 = ps[2];
        const v3 thr = mk(t4.x, t4.y, t4.z);
        // the gather's weight for this vertex's NEE, w: 1 / light samples (the gather needs no camera record)
        W.thr_slot[k] = make_float4(t4.x, t4.y, t4.z, 1.0f / fnl);
        const float4 *sh = W.S.tshade + 4 * (size_t)tri;
        const float4 n0 = sh[0];
        const uint4 ms = W.S.meshes[__float_as_uint(n0.w)];
        v3 ng = ld3(sh[3]), ns, nsi;
        {
          const float w0 = 1.0f - h4.y - h4.z;
          nsi = normalize(ld3(n0) * w0 + ld3(sh[1]) * h4.y + ld3(sh[2]) * h4.z);
        }
        if (INST && inst != NO_INST) {
          const Pose Pb = inst_pose(W.S, inst, time);
          ng = rot(Pb, ng);
          nsi = rot(Pb, nsi);
        }
        ns = nsi;
        if (dot(ng, d) > 0.0f) ng = -ng;
        if (dot(ns, ng) < 0.0f) ns = -ns;
        if (dot(ns, -d) < 1e-4f) ns = ng;
        const v3 pt = o + d * h4.x;
        if (ms.y & F_EMITTER) {
          const float cosE = -dot(d, nsi);
          if (cosE > 0.0f) {
            const v3 Le = ld3(W.S.lights[ms.z]) * light_profile(W.S, ms.z, emit_local(emitter_pose(W.S, time), -d));
            float w = 1.0f;
            if (b > 0) {
              const float pL = h4.x * h4.x / (cosE * A);
              w = pB * pB / (pB * pB + fnl * fnl * pL * pL);
            }
            float4 Lp = W.L[p];
            const v3 add = thr * Le * w;
            Lp.x += add.x; Lp.y += add.y; Lp.z += add.z;
            W.L[p] = Lp;
          }
        } else {
          const v3 wo = -d;
          const bool last = b + 1 == Dp;
          if (W.S.emit_count > 0) {
            for (uint32_t j = 0; j < nl; j++) {
              float u0, u1, u2;
              if (b == 0) {
                const uint32_t li = i * nl + j;
                u0 = kr(li, KL0, P.s[3]); u1 = kr(li, KL1, P.s[4]); u2 = kr(li, KL2, P.s[5]);
              } else {
                u0 = u01(hash3(hs, i, 256u * b + 3u * j));
                u1 = u01(hash3(hs, i, 256u * b + 3u * j + 1u));
                u2 = u01(hash3(hs, i, 256u * b + 3u * j + 2u));
              }
              v3 pl, nle;
              const EmitPose E = emitter_pose(W.S, time);
              const uint32_t e = sample_emitter(W.S, E, u0, u1, u2, pl, nle);
              v3 wi = pl - pt;
              const float d2 = dot(wi, wi), dist = sqrtf(d2);
              wi = wi * (1.0f / dist);
              const float cosE = -dot(wi, nle);
              if (cosE <= 0.0f) continue;
              const float cosS = dot(ns, wi);
              if (cosS <= 0.0f || dot(ng, wi) <= 0.0f) continue;
              const uint32_t lid = __float_as_uint(W.S.emit[4 * e + 2].w);
              const v3 Le = ld3(W.S.lights[lid]) * light_profile(W.S, lid, emit_local(E, -wi));
              if (maxc(Le) <= 0.0f) continue;
              const Mat M = hit_mat(W.S, ms, tri, h4.y, h4.z);
              const v3 f = eval_bsdf(M, ns, wi, wo);
              if (maxc(f) <= 0.0f) continue;
              const float pW = d2 / (cosE * A);
              float w = 1.0f;
              if (!last) {
                const float pb = pdf_bsdf(M, ns, wo, wi);
                w = fnl * fnl * pW * pW / (fnl * fnl * pW * pW + pb * pb);
              }
              const float kk = cosS * w / pW;
              const v3 contrib = f * Le * kk;
              const v3 so = pt + ng * (dot(ng, wi) > 0.0f ? EPS_OFF : -EPS_OFF);
              // slots in j order; the gather adds the visible ones in that order
              const size_t s = (size_t)k * W.nl_max + nee;
              W.sh_o[s] = make_float4(so.x, so.y, so.z, dist - 2e-5f);
              W.sh_d[s] = make_float4(wi.x, wi.y, wi.z, time);
              W.sh_c[s] = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
              if (W.sh_q) W.sh_key[s] = ray_key(so, wi, W);
              nee++;
            }
          }
          if (!last) {
            float u0, u1, u2;
            if (b == 0) {
              u0 = kr(i, KB0, P.s[6]); u1 = kr(i, KB1, P.s[7]); u2 = kr(i, KB2, P.s[8]);
            } else {
              u0 = u01(hash3(hs, i, 256u * b + 200u));
              u1 = u01(hash3(hs, i, 256u * b + 201u));
              u2 = u01(hash3(hs, i, 256u * b + 202u));
            }
            const Mat M = hit_mat(W.S, ms, tri, h4.y, h4.z);
            v3 l;
            const float pdf = sample_bsdf(M, ns, wo, u0, u1, u2, l);
            if (pdf > 0.0f && dot(ng, l) > 0.0f) {
              const v3 f = eval_bsdf(M, ns, l, wo);
              const float cp = dot(ns, l) / pdf;
              const v3 a = f * cp;
              const v3 tn = thr * a;
              if (maxc(tn) > 0.0f) {
                const v3 no = pt + ng * (dot(ng, l) > 0.0f ? EPS_OFF : -EPS_OFF);
                // the whole 64-B record, so the line is written in full
                ps[0] = make_float4(no.x, no.y, no.z, pdf);
                ps[1] = make_float4(l.x, l.y, l.z, time);
                ps[2] = make_float4(tn.x, tn.y, tn.z, 1.0f);
                ps[3] = ps3;
                W.key_next[k] = ray_key(no, l, W);
                W.q_next[k] = p;
                live = true;
              }
            }
          }
        }
      }
      W.nee[k] = nee;
      nee_k = nee;
      if (!live) {
        W.key_next[k] = DEAD_KEY;
        W.q_next[k] = p;
      }
    }
    wave_live += wave_count(live);
    if (W.sh_q) {
      // the unused shadow slots sort last; the live ones are counted
      if (k < W.n)
        for (uint32_t j = nee_k; j < W.nl_max; j++) W.sh_key[(size_t)k * W.nl_max + j] = DEAD_KEY;
      lane_nee += nee_k;
    }
  }
  // one atomic per wave for the whole grid-stride loop: same-address atomics serialise across the XCDs (one per
  // wave and iteration made this kernel take twice as long)
  if ((threadIdx.x & 63u) == 0 && wave_live) atomicAdd(&W.count[1], wave_live);
  if (W.sh_q) {
    for (int off = 32; off >= 1; off >>= 1) lane_nee += __shfl_xor(lane_nee, off, 64);
    if ((threadIdx.x & 63u) == 0 && lane_nee) atomicAdd(&W.count[4], lane_nee);
  }
}

// L += thr * (sum of the visible NEE contributions, in sample order) / nl; then the continuation's throughput
__global__ __launch_bounds__(256) void k_wf_gather(WfLaunch W) {
  const uint32_t n = W.count[0];
  for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
    const uint32_t p = W.q ? W.q[k] : k;
    const uint32_t ne = W.nee[k];
    if (ne) {
      v3 Ln = mk(0.f, 0.f, 0.f);
      for (uint32_t j = 0; j < ne; j++) {
        const float4 c = W.sh_c[(size_t)k * W.nl_max + j];
        if (c.w > 0.0f) Ln = Ln + mk(c.x, c.y, c.z);
      }
      const float4 t4 = W.thr_slot[k];
      const v3 add = mk(t4.x, t4.y, t4.z) * Ln * t4.w;
      float4 Lp = W.L[p];
      Lp.x += add.x; Lp.y += add.y; Lp.z += add.z;
      W.L[p] = Lp;
    }
  }
}

// after the sort: the next vertex's live count
__global__ void k_wf_swap_counts(uint32_t *count) {
  const uint32_t next = count[1];
  __syncthreads();
  if (threadIdx.x < WF_COUNT_WORDS) count[threadIdx.x] = threadIdx.x == 0 ? next : 0u;
}

// per-pixel sums of the batch: the butterfly of run_wave over each 64-item wave
__global__ __launch_bounds__(256) void k_wf_accum(WfLaunch W) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t waves = W.n / 64;
  uint32_t nsamp = 0;
  for (uint32_t w = (blockIdx.x * 256 + threadIdx.x) / 64; w < waves; w += gridDim.x * 4) {
    const uint32_t p = w * 64 + lane;
    const uint64_t item = path_item(W, p);
    uint32_t c = 0;   // the camera from the item's place in the frame queue, as k_wf_gen finds it
    while (c + 1 < W.ncam && item >= W.cams[c + 1].item_begin) c++;
    const DevCam &C = W.cams[c];
    const ItemPix ip = item_pix(C, item);
    const float4 L4 = W.L[p];
    const v3 L = ip.valid ? mk(L4.x, L4.y, L4.z) : mk(0.f, 0.f, 0.f);
    const uint32_t nspp = C.nspp;
    const bool pow2 = (nspp & (nspp - 1)) == 0 && nspp <= 64;
    if (pow2) {
      float sx = L.x, sy = L.y, sz = L.z;
      for (uint32_t off = 1; off < nspp; off <<= 1) {
        sx += __shfl_xor(sx, (int)off, 64);
        sy += __shfl_xor(sy, (int)off, 64);
        sz += __shfl_xor(sz, (int)off, 64);
      }
      if (ip.valid && (lane & (nspp - 1)) == 0) {
        float4 *a = C.acc + (size_t)ip.y * C.width + ip.x;
        float4 v = *a;
        v.x += sx; v.y += sy; v.z += sz; v.w += (float)nspp;
        *a = v;
      }
    } else if (ip.valid) {
      float4 *a = C.acc + (size_t)ip.y * C.width + ip.x;
      atomicAdd(&a->x, L.x);
      atomicAdd(&a->y, L.y);
      atomicAdd(&a->z, L.z);
      atomicAdd(&a->w, 1.0f);
    }
    nsamp += ip.valid ? 1u : 0u;
  }
  for (int off = 32; off >= 1; off >>= 1) nsamp += __shfl_xor(nsamp, off, 64);
  if (lane == 0 && nsamp) atomicAdd(&W.counters[2], (unsigned long long)nsamp);
}

uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap) {
  const uint64_t g = (n + block - 1) / block;
  return (uint32_t)(g < 1 ? 1 : (g > cap ? cap : g));
}

}  // namespace

namespace {
// the trace kernel for an LDS stack depth, with or without the posed-body walk
template <bool ANY, bool INST>
void (*trace_kernel(uint32_t slds))(WfLaunch, uint32_t) {
  switch (slds) {
    case 4: return k_wf_trace_dyn<ANY, 4, INST>;
    case 8: return k_wf_trace_dyn<ANY, 8, INST>;
    default: return k_wf_trace_dyn<ANY, 16, INST>;
  }
}
template <bool ANY>
void (*trace_kernel(uint32_t slds, bool inst))(WfLaunch, uint32_t) {
  return inst ? trace_kernel<ANY, true>(slds) : trace_kernel<ANY, false>(slds);
}
}  // namespace

hipError_t wf_trace_occupancy(uint32_t slds, bool inst, int *blocks_per_cu) {
  if (slds != 4 && slds != 8 && slds != 16) return hipErrorInvalidValue;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, trace_kernel<false>(slds, inst), 256, 0);
}

// ---------------------------------------------------------------- the bin sort between vertices
// A counting sort of the live slots by the top `bits` bits of their key (direction cell first, then the coarse
// origin), the dead slots in one last bin: a histogram pass, one scan, and a scatter in which each block reserves
// its share of every bin with one atomic and ranks its own slots in LDS. It is not stable -- the order within a bin
// follows the atomics -- which leaves every path's result unchanged (a path's work depends on its own state only)
// and moves no more than 16 B per slot; it issues kernels only, so the frame can be captured as a graph.
constexpr uint32_t BIN_SORT_CHUNK = 4096;     // slots per scatter block
__device__ __forceinline__ uint32_t key_bin(uint32_t key, uint32_t bits) {
  return key == DEAD_KEY ? (1u << bits) : key >> (KEY_BITS - bits);
}

__global__ __launch_bounds__(256) void k_bin_zero(uint32_t *hist, uint32_t nbins) {
  for (uint32_t b = blockIdx.x * 256 + threadIdx.x; b < nbins; b += gridDim.x * 256) hist[b] = 0u;
}

__global__ __launch_bounds__(256) void k_bin_hist(const uint32_t *keys, uint32_t n, uint32_t bits, uint32_t *hist) {
  extern __shared__ uint32_t s_cnt[];
  const uint32_t nbins = (1u << bits) + 1u;
  for (uint32_t b = threadIdx.x; b < nbins; b += 256) s_cnt[b] = 0u;
  __syncthreads();
  const uint32_t k0 = blockIdx.x * BIN_SORT_CHUNK, k1 = min(n, k0 + BIN_SORT_CHUNK);
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += 256) atomicAdd(&s_cnt[key_bin(keys[k], bits)], 1u);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nbins; b += 256)
    if (s_cnt[b]) atomicAdd(&hist[b], s_cnt[b]);
}

// exclusive prefix sum of the bin counts, in place (one block of 1024 lanes, each summing a contiguous run)
__global__ __launch_bounds__(1024) void k_bin_scan(uint32_t *hist, uint32_t nbins) {
  __shared__ uint32_t s_part[1024];
  const uint32_t per = (nbins + 1023) / 1024;
  const uint32_t b0 = threadIdx.x * per, b1 = min(nbins, b0 + per);
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; b++) sum += hist[b];
  s_part[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    const uint32_t v = threadIdx.x >= off ? s_part[threadIdx.x - off] : 0u;
    __syncthreads();
    s_part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? s_part[threadIdx.x - 1] : 0u;
  for (uint32_t b = b0; b < b1; b++) {
    const uint32_t c = hist[b];
    hist[b] = run;
    run += c;
  }
}

__global__ __launch_bounds__(256) void k_bin_scatter(const uint32_t *keys, const uint32_t *vals, uint32_t n,
                                                     uint32_t bits, uint32_t *cursor, uint32_t *out) {
  extern __shared__ uint32_t s_cnt[];
  const uint32_t nbins = (1u << bits) + 1u;
  for (uint32_t b = threadIdx.x; b < nbins; b += 256) s_cnt[b] = 0u;
  __syncthreads();
  const uint32_t k0 = blockIdx.x * BIN_SORT_CHUNK, k1 = min(n, k0 + BIN_SORT_CHUNK);
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += 256) atomicAdd(&s_cnt[key_bin(keys[k], bits)], 1u);
  __syncthreads();
  // the block's share of each bin: s_cnt becomes its first position there
  for (uint32_t b = threadIdx.x; b < nbins; b += 256) {
    const uint32_t c = s_cnt[b];
    s_cnt[b] = c ? atomicAdd(&cursor[b], c) : 0u;
  }
  __syncthreads();
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += 256) {
    const uint32_t pos = atomicAdd(&s_cnt[key_bin(keys[k], bits)], 1u);
    out[pos] = vals[k];
  }
}

hipError_t launch_bin_sort(const uint32_t *keys, const uint32_t *vals, uint32_t n, uint32_t bits, uint32_t *hist,
                           uint32_t *out, hipStream_t s) {
  if (bits < 4 || bits > WF_SORT_BITS_MAX) return hipErrorInvalidValue;
  const uint32_t nbins = (1u << bits) + 1u;
  const size_t lds = (size_t)nbins * 4;
  const uint32_t blocks = std::max<uint32_t>(1, (n + BIN_SORT_CHUNK - 1) / BIN_SORT_CHUNK);
  hipLaunchKernelGGL(k_bin_zero, dim3((nbins + 255) / 256), dim3(256), 0, s, hist, nbins);
  hipLaunchKernelGGL(k_bin_hist, dim3(blocks), dim3(256), lds, s, keys, n, bits, hist);
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, hist, nbins);
  hipLaunchKernelGGL(k_bin_scatter, dim3(blocks), dim3(256), lds, s, keys, vals, n, bits, hist, out);
  return hipGetLastError();
}

size_t wf_sort_temp_bytes(uint32_t n) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                  (const uint32_t *)nullptr, (uint32_t *)nullptr, n, 0, KEY_BITS + 1);
  return std::max<size_t>(bytes, ((size_t)(1u << WF_SORT_BITS_MAX) + 1u) * 4);
}

// One batch: generate, then per vertex trace / shade / shadow / gather, and a sort between vertices.
hipError_t launch_wavefront_batch(WfLaunch W, uint32_t max_depth, uint32_t grid_cap, uint32_t trace_grid_cap,
                                  uint32_t refill, uint32_t slds, uint32_t sort_bits, uint32_t *keys_sorted,
                                  uint32_t *q_sorted, void *sort_temp, size_t sort_temp_bytes, bool sort_shadows,
                                  hipEvent_t *ev_ext, hipStream_t s) {
  if (!sort_shadows) W.sh_q = nullptr;
  const uint32_t nsh = W.n * W.nl_max;
  if (refill < 1 || refill > 64 || (slds != 4 && slds != 8 && slds != 16)) return hipErrorInvalidValue;
  const uint32_t g = grid_for(W.n, 256, grid_cap);
  const uint32_t gt = grid_for(W.n, 256, trace_grid_cap);
  const uint32_t gs = grid_for((uint64_t)W.n * W.nl_max, 256, trace_grid_cap);
  const bool inst = W.S.inst_count != 0;
  void (*trace_ext)(WfLaunch, uint32_t) = trace_kernel<false>(slds, inst);
  void (*trace_shadow)(WfLaunch, uint32_t) = trace_kernel<true>(slds, inst);
  hipLaunchKernelGGL(k_wf_gen, dim3(g), dim3(256), 0, s, W);
  W.q = nullptr;   // vertex 1: the camera rays in item order (coherent already)
  for (uint32_t b = 0; b < max_depth; b++) {
    W.bounce = b;
    if (ev_ext) {
      hipError_t e = hipEventRecord(ev_ext[2 * b], s);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(trace_ext, dim3(gt), dim3(256), 0, s, W, refill);
    if (ev_ext) {
      hipError_t e = hipEventRecord(ev_ext[2 * b + 1], s);
      if (e != hipSuccess) return e;
    }
    if (inst) hipLaunchKernelGGL(k_wf_shade<true>, dim3(g), dim3(256), 0, s, W);
    else hipLaunchKernelGGL(k_wf_shade<false>, dim3(g), dim3(256), 0, s, W);
    if (W.sh_q) {
      // the shadow slots by their rays' keys, so a wave traces shadow rays that start close together and head the
      // same way (the visibility of a slot does not depend on the order)
      hipError_t e;
      if (sort_bits) {
        e = launch_bin_sort(W.sh_key, W.sh_idx, nsh, sort_bits, (uint32_t *)sort_temp, W.sh_q, s);
      } else {
        size_t tb = sort_temp_bytes;
        e = rocprim::radix_sort_pairs(sort_temp, tb, W.sh_key, W.sh_keys_sorted, W.sh_idx, W.sh_q, nsh, 0, KEY_BITS + 1, s);
      }
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(trace_shadow, dim3(gs), dim3(256), 0, s, W, refill);
    hipLaunchKernelGGL(k_wf_gather, dim3(g), dim3(256), 0, s, W);
    if (b + 1 == max_depth) break;
    hipError_t e;
    if (sort_bits) {
      e = launch_bin_sort(W.key_next, W.q_next, W.n, sort_bits, (uint32_t *)sort_temp, q_sorted, s);
    } else {
      size_t tb = sort_temp_bytes;
      e = rocprim::radix_sort_pairs(sort_temp, tb, W.key_next, keys_sorted, W.q_next, q_sorted, W.n, 0, KEY_BITS + 1, s);
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_wf_swap_counts, dim3(1), dim3(WF_COUNT_WORDS), 0, s, W.count);
    W.q = q_sorted;
  }
  hipLaunchKernelGGL(k_wf_accum, dim3(grid_for(W.n, 256, grid_cap)), dim3(256), 0, s, W);
  return hipGetLastError();
}

}  // namespace vt
