// Camera path tracing on gfx950: one lane per (pixel, sub-sample) item, 64 items per wave, items of every
// camera of a frame in one queue. The persistent form takes waves from the queue with an atomic head (a
// frame of uneven cameras is one launch); the plain form maps one wave to each 64 items of one camera.
// A wave's lanes hold consecutive sub-samples of the same pixel(s), so primary rays are coherent and the
// pixel sum is a butterfly reduction within the wave (deterministic; one writer per pixel).
#include "vt_dual.h"
#include "vt_internal.h"
#include "vt_nodes.h"

namespace vt {

// Gradient sink of the adjoint: per-block LDS accumulators (material colour rgb, light radiance rgb), or the
// global buffer directly when the tables are too large for LDS.
struct GradSink {
  float *mat;      // n_mat * 3
  float *light;    // n_light * 3
  v3 delta;        // dLoss/dI of this sample's pixel times scale / spp
};
__device__ __forceinline__ void grad_add(float *base, uint32_t idx, v3 g) {
  if (g.x != 0.0f) atomicAdd(base + 3 * idx, g.x);
  if (g.y != 0.0f) atomicAdd(base + 3 * idx + 1, g.y);
  if (g.z != 0.0f) atomicAdd(base + 3 * idx + 2, g.z);
}
__device__ __forceinline__ float ratio(float d, float a) { return a > 0.0f ? d / a : 0.0f; }

// The camera estimator (the reference's pieces composed into paths of `depth` vertices, as the oracle states):
// NEE with the power heuristic at every vertex but the last, NEE alone at the last, emitter hits by BSDF rays
// MIS-weighted (gpu/wgsl/shade.wgsl.js:135-188, 224-236), Kronecker dimensions at the first vertex
// (gpu/wgsl/kernels.wgsl.js:42-53) and counter hashes after it.
// GRAD: the path-replay walk of the adjoint. The same path is re-traced from the same random numbers; every
// term's derivative with respect to the lights and the material colours is added to the sink, sampling and
// MIS weights detached; a continuation factor a = f cos / pdf at vertex k scales everything after it, whose sum
// is Ltot - (what was accumulated up to k), so dL/dc gets (da / a) * (Ltot - L_acc) there.
template <bool GRAD, bool INST>
__device__ __forceinline__ v3 path_walk(const SceneDev &S, const DevCam &C, const PixSeeds &P, uint32_t x, uint32_t y,
                                        uint32_t i, Stack &st, uint32_t &nrays, GradSink *gs = nullptr,
                                        v3 Ltot = v3{0.f, 0.f, 0.f}) {
  // Register pressure: every value a shadow ray or an extension ray waits for is computed before its traversal
  // (the NEE term and its MIS weight, the derivative pieces), and the material record is re-read from the
  // cached table after it, so only the ray and the path's small state stay live across the traversals.
  const uint32_t hs = P.hs;
  const float time = sample_time(C, P, i);
  v3 o = mk(C.pos[0], C.pos[1], C.pos[2]);
  v3 d = camera_dir(C, P, x, y, i);
  v3 thr = mk(1.f, 1.f, 1.f), L = mk(0.f, 0.f, 0.f);
  float pB = 0.0f;
  const uint32_t nl = C.nl < 1 ? 1 : (C.nl > 16 ? 16 : C.nl);
  const float fnl = (float)nl;
  const uint32_t D = C.depth < 1 ? 1 : C.depth;
  const float A = S.emit_area;
  for (uint32_t b = 0; b < D; b++) {
    Hit h;
    nrays++;
    if (!trace_scene<false, INST>(S, o, d, 0.0f, INFINITY, time, b == 0 ? C.skip : 0u, h, st)) break;
    const float4 *sh = S.tshade + 4 * (size_t)h.tri;
    const float4 n0 = sh[0];
    const uint4 ms = S.meshes[__float_as_uint(n0.w)];
    v3 ng = ld3(sh[3]), ns, nsi;
    {
      const float w0 = 1.0f - h.u - h.v;
      nsi = normalize(ld3(n0) * w0 + ld3(sh[1]) * h.u + ld3(sh[2]) * h.v);
    }
    if (INST && h.inst != NO_INST) {
      const Pose Pb = inst_pose(S, h.inst, time);
      ng = rot(Pb, ng);
      nsi = rot(Pb, nsi);
    }
    ns = nsi;
    if (dot(ng, d) > 0.0f) ng = -ng;
    if (dot(ns, ng) < 0.0f) ns = -ns;
    if (dot(ns, -d) < 1e-4f) ns = ng;
    const v3 p = o + d * h.t;
    if (ms.y & F_EMITTER) {
      const float cosE = -dot(d, nsi);
      if (cosE > 0.0f) {
        const float pe = light_profile(S, ms.z, emit_local(emitter_pose(S, time), -d));
        const v3 Le = ld3(S.lights[ms.z]) * pe;
        float w = 1.0f;
        if (b > 0) {
          const float pL = h.t * h.t / (cosE * A);
          w = pB * pB / (pB * pB + fnl * fnl * pL * pL);
        }
        L = L + thr * Le * w;
        if (GRAD) grad_add(gs->light, ms.z, gs->delta * thr * (w * pe));
      }
      break;
    }
    const v3 wo = -d;
    const bool last = b + 1 == D;
    v3 Ln = mk(0.f, 0.f, 0.f);
    if (S.emit_count > 0) {
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
        v3 so, wi, contrib, gl_part, gm_part;
        float tmax;
        uint32_t lid;
        {
          v3 pl, nl;
          const EmitPose E = emitter_pose(S, time);
          const uint32_t e = sample_emitter(S, E, u0, u1, u2, pl, nl);
          wi = pl - p;
          const float d2 = dot(wi, wi), dist = sqrtf(d2);
          wi = wi * (1.0f / dist);
          const float cosE = -dot(wi, nl);
          if (cosE <= 0.0f) continue;
          const float cosS = dot(ns, wi);
          if (cosS <= 0.0f || dot(ng, wi) <= 0.0f) continue;
          lid = __float_as_uint(S.emit[4 * e + 2].w);
          const float pe = light_profile(S, lid, emit_local(E, -wi));
          const v3 Le = ld3(S.lights[lid]) * pe;
          if (maxc(Le) <= 0.0f) continue;
          const Mat M = hit_mat(S, ms, h.tri, h.u, h.v);
          const v3 f = eval_bsdf(M, ns, wi, wo);
          if (maxc(f) <= 0.0f) continue;
          const float pW = d2 / (cosE * A);
          float w = 1.0f;
          if (!last) {
            const float pb = pdf_bsdf(M, ns, wo, wi);
            w = fnl * fnl * pW * pW / (fnl * fnl * pW * pW + pb * pb);
          }
          const float k = cosS * w / pW;
          contrib = f * Le * k;
          if (GRAD) {
            const v3 g = gs->delta * thr * (1.0f / fnl);
            gl_part = g * f * (k * pe);
            gm_part = g * eval_bsdf_dc(M, ns, wi, wo) * Le * k;
          }
          so = p + ng * (dot(ng, wi) > 0.0f ? EPS_OFF : -EPS_OFF);
          tmax = dist - 2e-5f;
        }
        Hit sh;
        nrays++;
        if (trace_scene<true, INST>(S, so, wi, 0.0f, tmax, time, S.shadow_skip, sh, st)) continue;
        Ln = Ln + contrib;
        if (GRAD) {
          grad_add(gs->light, lid, gl_part);
          grad_add(gs->mat, ms.x, gm_part);
        }
      }
    }
    L = L + thr * Ln * (1.0f / fnl);
    if (last) break;
    float u0, u1, u2;
    if (b == 0) {
      u0 = kr(i, KB0, P.s[6]); u1 = kr(i, KB1, P.s[7]); u2 = kr(i, KB2, P.s[8]);
    } else {
      u0 = u01(hash3(hs, i, 256u * b + 200u));
      u1 = u01(hash3(hs, i, 256u * b + 201u));
      u2 = u01(hash3(hs, i, 256u * b + 202u));
    }
    const Mat M = hit_mat(S, ms, h.tri, h.u, h.v);
    v3 l;
    const float pdf = sample_bsdf(M, ns, wo, u0, u1, u2, l);
    if (!(pdf > 0.0f) || dot(ng, l) <= 0.0f) break;
    const v3 f = eval_bsdf(M, ns, l, wo);
    const float cp = dot(ns, l) / pdf;
    const v3 a = f * cp;
    if (GRAD) {
      const v3 da = eval_bsdf_dc(M, ns, l, wo) * cp;
      const v3 rest = Ltot - L;
      grad_add(gs->mat, ms.x, gs->delta * mk(ratio(da.x, a.x) * rest.x, ratio(da.y, a.y) * rest.y, ratio(da.z, a.z) * rest.z));
    }
    thr = thr * a;
    if (maxc(thr) <= 0.0f) break;
    o = p + ng * (dot(ng, l) > 0.0f ? EPS_OFF : -EPS_OFF);
    d = l;
    pB = pdf;
  }
  return L;
}

template <bool INST>
__device__ __forceinline__ v3 path_radiance(const SceneDev &S, const DevCam &C, const PixSeeds &P, uint32_t x,
                                            uint32_t y, uint32_t i, Stack &st, uint32_t &nrays) {
  return path_walk<false, INST>(S, C, P, x, y, i, st, nrays);
}

// ------------------------------------------------------------------ the reference's camera estimator
// The hit's surface (surface, trace.wgsl.js:172-195): vertex ids, barycentrics, the outward interpolated normal
// (nsi, unflipped), the two-sided shading and geometric normals, the point; posed bodies in the world at `time`
struct Surf {
  v3 p, ng, ns, nsi, b;
  uint4 vi;
  uint4 ms;     // mesh record: material, flags, light, -
  uint32_t mesh;
};
template <bool INST>
__device__ __forceinline__ Surf surface_at(const SceneDev &S, const Hit &h, v3 o, v3 d, float time) {
  Surf s;
  const float4 *sh = S.tshade + 4 * (size_t)h.tri;
  const float4 n0 = sh[0];
  s.mesh = __float_as_uint(n0.w);
  s.ms = S.meshes[s.mesh];
  s.vi = S.tvid[h.tri];
  s.b = mk(1.0f - h.u - h.v, h.u, h.v);
  v3 ng = ld3(sh[3]);
  v3 nsi = normalize(ld3(n0) * s.b.x + ld3(sh[1]) * s.b.y + ld3(sh[2]) * s.b.z);
  if (INST && h.inst != NO_INST) {
    const Pose Pb = inst_pose(S, h.inst, time);
    ng = rot(Pb, ng);
    nsi = rot(Pb, nsi);
  }
  v3 ns = nsi;
  if (dot(ng, d) > 0.0f) ng = -ng;
  if (dot(ns, ng) < 0.0f) ns = -ns;
  if (dot(ns, -d) < 1e-4f) ns = ng;
  s.ng = ng;
  s.ns = ns;
  s.nsi = nsi;
  s.p = o + d * h.t;
  return s;
}
__device__ __forceinline__ uint32_t rad_slot(const SceneDev &S) { return __float_as_uint(S.radu[0].w); }
__device__ __forceinline__ bool is_static_hit(const Hit &h, const Surf &s) {
  return (s.ms.y & F_STATIC) && h.inst == NO_INST;
}

// cachedVisibility (shade.wgsl.js:108-130): 1 or 0 when the hit triangle's three nodes agree that every lit emitter
// sample is visible or hidden and no visible body's bounding capsule comes near the segment to the emitters'
// centroid (its radius + 0.5 mm added); -1: trace the shadow rays
template <bool INST>
__device__ __forceinline__ float cached_visibility(const SceneDev &S, const Surf &s, float time) {
  const float a = S.nstate[4 * (size_t)s.vi.x + 3].w, b = S.nstate[4 * (size_t)s.vi.y + 3].w,
              c = S.nstate[4 * (size_t)s.vi.z + 3].w;
  if (!(a == b && b == c && (a == 1.0f || a == 0.0f))) return -1.0f;
  if (INST) {
    const float4 ec = S.radu[1];
    const v3 d = emit_centre(S, emitter_pose(S, time)) - s.p;
    for (uint32_t i = 0; i < S.inst_count; i++) {
      const float4 *I = S.inst + INST_WORDS * (size_t)i;
      if ((__float_as_uint(I[3].w) & INST_HIDDEN) || ((S.shadow_skip >> i) & 1u)) continue;
      if (seg_near_capsule(s.p, d, 0.0f, 1.0f, ld3(I[6]), ld3(I[7]), I[6].w + ec.w + 5e-4f)) return -1.0f;
    }
  }
  return a;
}

// one emitter sample's unshadowed contribution and its shadow segment (lightEval, shade.wgsl.js:135-157): ok false
// when the sample cannot light the point (back of the emitter or the surface, outside the emission profile, a zero
// BSDF); the segment starts off the surface on the lit side and ends 20 um short of the emitter
struct LightEval {
  v3 c, so, wi;
  float tmax;
  bool ok;
};

__device__ __forceinline__ LightEval light_eval(const SceneDev &S, const Surf &s, const Mat &M, v3 wo, float u0,
                                                float u1, float u2, float time, bool mis, float fnl) {
  LightEval e;
  e.ok = false;
  v3 pl, nl;
  const EmitPose E = emitter_pose(S, time);
  const uint32_t em = sample_emitter(S, E, u0, u1, u2, pl, nl);
  v3 wi = pl - s.p;
  const float d2 = dot(wi, wi), dist = sqrtf(d2);
  wi = wi * (1.0f / dist);
  const float cosE = -dot(wi, nl);
  if (cosE <= 0.0f) return e;
  const float cosS = dot(s.ns, wi);
  if (cosS <= 0.0f || dot(s.ng, wi) <= 0.0f) return e;
  const uint32_t lid = __float_as_uint(S.emit[4 * em + 2].w);
  const v3 Le = ld3(S.lights[lid]) * light_profile(S, lid, emit_local(E, -wi));
  if (maxc(Le) <= 0.0f) return e;
  const v3 f = eval_bsdf(M, s.ns, wi, wo);
  if (maxc(f) <= 0.0f) return e;
  const float pW = d2 / (cosE * S.emit_area);
  float w = 1.0f;
  if (mis) {
    const float pb = pdf_bsdf(M, s.ns, wo, wi);
    w = fnl * fnl * pW * pW / (fnl * fnl * pW * pW + pb * pb);
  }
  e.c = f * Le * (cosS * w / pW);
  e.so = s.p + s.ng * EPS_OFF;      // dot(ng, wi) > 0 here: the lit side
  e.wi = wi;
  e.tmax = dist - 2e-5f;
  e.ok = true;
  return e;
}

// one emitter sample's contribution at the surface (directLight, shade.wgsl.js:160-188): `cached` 1 / 0 answers the
// shadow test (cachedVisibility), -1 traces it; a traced shadow ray marks the pixel's key (0x20000)
template <bool INST>
__device__ __forceinline__ v3 direct_light(const SceneDev &S, const Surf &s, const Mat &M, v3 wo, float time, float u0,
                                           float u1, float u2, bool mis, float fnl, float cached, Stack &st,
                                           uint32_t &nrays, uint32_t &key) {
  const LightEval e = light_eval(S, s, M, wo, u0, u1, u2, time, mis, fnl);
  if (!e.ok || cached == 0.0f) return mk(0.f, 0.f, 0.f);
  if (cached < 0.0f) {
    Hit sh;
    nrays++;
    key |= 0x20000u;
    if (trace_scene<true, INST>(S, e.so, e.wi, 0.0f, e.tmax, time, S.shadow_skip, sh, st)) return mk(0.f, 0.f, 0.f);
  }
  return e.c;
}

// the nL light samples of a static surface under a still light (shade.wgsl.js:272-301): each lit sample is tested
// against the static scene -- the answer kept in the camera's occlusion cache word for this sub-sample and pixel
// when there is one (bit j, bit 31 = written) -- and, if clear, against the visible bodies. The key's 0x20000 marks
// a body's shadow, a penumbra (some lit samples hidden, some not) or visible samples whose luminances spread by
// more than 20 % of their midpoint. Returns the sum of the visible samples.
template <bool INST>
__device__ __forceinline__ v3 still_light(const SceneDev &S, const DevCam &C, const PixSeeds &P, const Surf &s,
                                          const Mat &M, v3 wo, float time, uint32_t x, uint32_t y, uint32_t i,
                                          uint32_t nl, Stack &st, uint32_t &nrays, uint32_t &key) {
  uint32_t *const sc = C.scache ? C.scache + ((size_t)i * C.height + y) * C.width + x : nullptr;
  const uint32_t word = sc ? *sc : 0u;
  const bool valid = (word & 0x80000000u) != 0u;
  const bool bodies = INST && S.inst_count;
  const float fnl = (float)nl;
  uint32_t nw = 0x80000000u, nvis = 0, nlit = 0;
  float lmin = INFINITY, lmax = 0.0f;
  v3 L = mk(0.f, 0.f, 0.f);
  for (uint32_t j = 0; j < nl; j++) {
    const uint32_t li = i * nl + j;
    const LightEval e = light_eval(S, s, M, wo, kr(li, KL0, P.s[3]), kr(li, KL1, P.s[4]), kr(li, KL2, P.s[5]),
                                   time, false, fnl);
    if (!e.ok) continue;
    nlit++;
    Hit sh;
    bool occ;
    if (valid) {
      occ = ((word >> j) & 1u) != 0u;
    } else {
      occ = traverse<true>(S, S.root, e.so, e.wi, 0.0f, e.tmax, sh, st);
      if (occ) nw |= 1u << j;
    }
    if (!valid || (bodies && !occ)) nrays++;      // one shadow segment, whichever parts of the scene it met
    if (occ) continue;
    if (bodies) {
      bool hit = false;
      trace_instances<true>(S, e.so, e.wi, 0.0f, e.tmax, time, S.shadow_skip, sh, hit, st);
      if (hit) {
        key |= 0x20000u;
        continue;
      }
    }
    nvis++;
    L = L + e.c;
    const float yl = luminance(e.c);
    lmin = fminf(lmin, yl);
    lmax = fmaxf(lmax, yl);
  }
  if (sc && !valid) *sc = nw;
  if (nvis > 0 && nvis < nlit) key |= 0x20000u;
  if (nvis > 1 && lmax - lmin > 0.2f * (lmax + lmin) * 0.5f) key |= 0x20000u;
  return L;
}

// radiance along a body's BSDF-sampled secondary ray (secondary, shade.wgsl.js:222-243): the emitters with MIS,
// the cached source of a static surface toward the ray, or the mean irradiance for another body
template <bool INST>
__device__ __forceinline__ v3 secondary(const SceneDev &S, v3 o, v3 l, float time, float pdfB, uint32_t opts, float fnl,
                                        Stack &st, uint32_t &nrays) {
  Hit h;
  nrays++;
  if (!trace_scene<false, INST>(S, o, l, 0.0f, INFINITY, time, 0u, h, st)) return mk(0.f, 0.f, 0.f);
  const Surf s = surface_at<INST>(S, h, o, l, time);
  if (s.ms.y & F_EMITTER) {
    if (opts & OPT_NO_MIS) return mk(0.f, 0.f, 0.f);
    const float cosE = -dot(l, s.nsi);
    if (cosE <= 0.0f) return mk(0.f, 0.f, 0.f);
    const v3 Le = ld3(S.lights[s.ms.z]) * light_profile(S, s.ms.z, emit_local(emitter_pose(S, time), -l));
    const float pL = h.t * h.t / (cosE * S.emit_area);
    return Le * (pdfB * pdfB / (pdfB * pdfB + fnl * fnl * pL * pL));
  }
  if (opts & (OPT_DIRECT_ONLY | OPT_NO_INDIRECT)) return mk(0.f, 0.f, 0.f);
  if (is_static_hit(h, s)) {
    const float mu = fminf(fmaxf(dot(s.ns, -l), 0.0f), 1.0f);
    const uint32_t slot = rad_slot(S);
    return albedo_at(S.alb, ALB_BINS, s.ms.x, mu) * node_interp(S, s.vi, s.b, slot, slot);
  }
  return ld3(S.hh[s.ms.x]) * ld3(S.radu[0]) * INV_PI_F;
}

// shadeHit (shade.wgsl.js:250-334) of a camera ray: an emitter's radiance; else nL light samples (MIS on posed
// bodies only), then on static surfaces the node cache E_free x capsule visibility x A(mu) / pi, on bodies one
// BSDF bounce into `secondary`. Light sample j uses kr(i nL + j); the bounce the first BSDF dimensions. key: the
// first hit's mesh (0xfffe: a miss), 0x10000 on a posed body, 0x20000 once a shadow ray is traced.
template <bool INST>
__device__ __forceinline__ v3 ref_radiance(const SceneDev &S, const DevCam &C, const PixSeeds &P, uint32_t x,
                                           uint32_t y, uint32_t i, uint32_t nl_in, Stack &st, uint32_t &nrays,
                                           uint32_t &key) {
  const float time = sample_time(C, P, i);
  const v3 o = mk(C.pos[0], C.pos[1], C.pos[2]);
  const v3 d = camera_dir(C, P, x, y, i);
  Hit h;
  nrays++;
  key = 0xfffeu;
  bool found;
  if (C.hcache) {
    // the static first hit from the camera's cache (written on the first frame after a change), then the bodies
    float4 *hc = C.hcache + ((size_t)i * C.height + y) * C.width + x;
    const float4 c = *hc;
    if (c.x > 0.0f) {
      found = __float_as_uint(c.y) != MISS;
      h.t = c.x;
      h.tri = __float_as_uint(c.y);
      h.u = c.z;
      h.v = c.w;
    } else {
      found = traverse<false>(S, S.root, o, d, 0.0f, INFINITY, h, st);
      *hc = found ? make_float4(h.t, __uint_as_float(h.tri), h.u, h.v) : make_float4(1e30f, __uint_as_float(MISS), 0.f, 0.f);
    }
    h.inst = NO_INST;
    if (INST && S.inst_count) trace_instances<false>(S, o, d, 0.0f, INFINITY, time, C.skip, h, found, st);
  } else {
    found = trace_scene<false, INST>(S, o, d, 0.0f, INFINITY, time, C.skip, h, st);
  }
  if (!found) return mk(0.f, 0.f, 0.f);
  const Surf s = surface_at<INST>(S, h, o, d, time);
  key = (s.mesh & 0xffffu) | ((h.inst != NO_INST && !(s.ms.y & F_STATIC)) ? 0x10000u : 0u);
  const v3 wo = -d;
  const uint32_t opts = C.opts;
  if (opts & OPT_NORMALS) return s.ns * 0.5f + mk(0.5f, 0.5f, 0.5f);
  if (s.ms.y & F_EMITTER) {
    if (dot(wo, s.nsi) <= 0.0f) return mk(0.f, 0.f, 0.f);
    return ld3(S.lights[s.ms.z]) * light_profile(S, s.ms.z, emit_local(emitter_pose(S, time), wo));
  }
  const Mat M = hit_mat(S, s.ms, h.tri, h.u, h.v);
  const bool stat = is_static_hit(h, s);
  const bool mis = !stat && !(opts & OPT_NO_MIS);
  const uint32_t nl = nl_in < 1 ? 1 : (nl_in > 16 ? 16 : nl_in);
  const float fnl = (float)nl;
  const float cached = (stat && (opts & OPT_CULL_SHADOWS) && !(opts & OPT_SHADOW_CACHE))
                           ? cached_visibility<INST>(S, s, time) : -1.0f;
  v3 L = mk(0.f, 0.f, 0.f);
  if (S.emit_count > 0) {
    if (stat && (opts & OPT_SHADOW_CACHE)) {
      L = still_light<INST>(S, C, P, s, M, wo, time, x, y, i, nl, st, nrays, key);
    } else {
      for (uint32_t j = 0; j < nl; j++) {
        const uint32_t li = i * nl + j;
        L = L + direct_light<INST>(S, s, M, wo, time, kr(li, KL0, P.s[3]), kr(li, KL1, P.s[4]), kr(li, KL2, P.s[5]),
                                   mis, fnl, cached, st, nrays, key);
      }
    }
  }
  L = L * (1.0f / fnl);
  if (stat) {
    if (!(opts & (OPT_NO_INDIRECT | OPT_DIRECT_ONLY))) {
      v3 E = node_interp(S, s.vi, s.b, rad_slot(S), 0);
      if (INST && !(opts & OPT_NO_CAPSULE)) E = E * capsule_vis(S, s.p, s.ns);
      const float mu = fmaxf(dot(s.ns, wo), 0.0f);
      L = L + E * INV_PI_F * albedo_at(S.alb, ALB_BINS, s.ms.x, mu);
    }
  } else {
    v3 l;
    const float pdf = sample_bsdf(M, s.ns, wo, kr(i, KB0, P.s[6]), kr(i, KB1, P.s[7]), kr(i, KB2, P.s[8]), l);
    if (pdf > 0.0f && dot(s.ng, l) > 0.0f) {
      const v3 thr = eval_bsdf(M, s.ns, l, wo) * (dot(s.ns, l) / pdf);
      const v3 so = s.p + s.ng * (dot(s.ng, l) > 0.0f ? EPS_OFF : -EPS_OFF);
      L = L + thr * secondary<INST>(S, so, l, time, pdf, opts, fnl, st, nrays);
    }
  }
  return L;
}

// ------------------------------------------------------------------ camera adjoint (interior term)
// The sample's image point (the u, v camera_dir traces)
__device__ __forceinline__ void camera_uv(const PixSeeds &P, uint32_t x, uint32_t y, uint32_t i, float &u, float &v) {
  u = ((float)x + kr(i, R2A, P.s[0])) - 0.5f;
  v = ((float)y + kr(i, R2B, P.s[1])) - 0.5f;
}

// d(camera-frame ray direction) / d(fx, fy, cx, cy, k1, k2, k3, k4) at image point (u, v): lens.wgsl.js:21-29
// differentiated, the Newton inverse of theta_d(theta) by its implicit derivative
// d theta = (d theta_d - sum_j theta^(2j+1) d k_j) / theta_d'(theta).
__device__ __forceinline__ void camdir_jac(const DevCam &C, float u, float v, v3 J[8]) {
  const float *L = C.lens;
  const float fx = L[0], fy = L[1], cx = L[2], cy = L[3], al = L[8];
#pragma unroll
  for (int j = 0; j < 8; j++) J[j] = mk(0.f, 0.f, 0.f);
  if (C.mode == 0) {
    const float k1 = L[4], k2 = L[5], k3 = L[6], k4 = L[7];
    const float yd = (v - cy) / fy, xd = (u - cx) / fx - al * yd;
    const float thd = sqrtf(xd * xd + yd * yd);
    if (thd < 1e-20f) return;     // the optical axis itself (measure zero)
    float th = thd;
#pragma unroll
    for (int it = 0; it < 8; it++) {
      const float t2 = th * th;
      const float f = th * (1.0f + t2 * (k1 + t2 * (k2 + t2 * (k3 + t2 * k4)))) - thd;
      const float fp = 1.0f + t2 * (3.0f * k1 + t2 * (5.0f * k2 + t2 * (7.0f * k3 + t2 * 9.0f * k4)));
      th = th - f / fp;
    }
    const float t2 = th * th;
    const float gp = 1.0f + t2 * (3.0f * k1 + t2 * (5.0f * k2 + t2 * (7.0f * k3 + t2 * 9.0f * k4)));
    float sn, cs;
    sincosf(th, &sn, &cs);
    const float s = sn / thd;
    const float dxd[4] = {-(u - cx) / (fx * fx), al * yd / fy, -1.0f / fx, al / fy};
    const float dyd[4] = {0.0f, -yd / fy, 0.0f, -1.0f / fy};
#pragma unroll
    for (int p = 0; p < 4; p++) {
      const float dthd = (xd * dxd[p] + yd * dyd[p]) / thd;
      const float dth = dthd / gp;
      const float ds = (cs * dth) / thd - sn * dthd / (thd * thd);
      J[p] = mk(dxd[p] * s + xd * ds, dyd[p] * s + yd * ds, -sn * dth);
    }
    float pw = th * t2;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float dth = -pw / gp;
      const float ds = cs * dth / thd;
      J[4 + j] = mk(xd * ds, yd * ds, -sn * dth);
      pw *= t2;
    }
  } else {
    const float qx = (u - cx) / fx, qy = (v - cy) / fy;
    const float inv = 1.0f / sqrtf(qx * qx + qy * qy + 1.0f);
    const v3 dc = mk(qx * inv, qy * inv, inv);
    const v3 dq[4] = {mk(-qx / fx, 0.f, 0.f), mk(0.f, -qy / fy, 0.f), mk(-1.0f / fx, 0.f, 0.f), mk(0.f, -1.0f / fy, 0.f)};
#pragma unroll
    for (int p = 0; p < 4; p++) J[p] = (dq[p] - dc * dot(dc, dq[p])) * inv;
  }
}

// a world direction with tangents in the emitter's frame (the pose does not depend on the camera)
__device__ __forceinline__ DV d_emit_local(const EmitPose &E, const DV &w) {
  if (!E.posed) return w;
  return {ddot(w, E.P.c0), ddot(w, E.P.c1), ddot(w, E.P.c2)};
}

// light_profile with the tangents of the direction: theta = acos(w . axis), d theta = -d(w . axis) / sin(theta);
// the same value as light_profile, zero tangents where p is flat or clamped
__device__ __forceinline__ D d_light_profile(const SceneDev &S, uint32_t light, const DV &w) {
  const float4 a = S.lprof[light];
  const uint32_t cnt = __float_as_uint(a.w);
  if (cnt == 0) return dk(1.0f);
  const float2 *K = S.knots + (size_t)light * MAX_KNOTS;
  const D c = ddot(w, mk(a.x, a.y, a.z));
  const float cv = fminf(fmaxf(c.v, -1.0f), 1.0f);
  const float th = acosf(cv);
  const float2 last = K[cnt - 1];
  if (th >= last.x) return dk(th == last.x ? last.y : 0.0f);
  uint32_t lo = 0, hi = cnt - 1;
  for (int it = 0; it < 8; it++) {
    if (hi - lo <= 1) break;
    const uint32_t m = (lo + hi) >> 1;
    if (K[m].x <= th) lo = m; else hi = m;
  }
  const float2 ka = K[lo], kb = K[hi];
  const float den = fmaxf(kb.x - ka.x, 1e-12f);
  const float f = (th - ka.x) / den;
  if (!(f > 0.0f)) return dk(ka.y);
  if (!(f < 1.0f)) return dk(kb.y);
  D r = c * ((kb.y - ka.y) / den * (-1.0f / sqrtf(fmaxf(1.0f - cv * cv, 1e-12f))));
  r.v = ka.y + (kb.y - ka.y) * f;
  return r;
}

// The estimator of path_walk with forward-mode tangents of the camera ray (origin o0, direction d0): the same
// random numbers, the same triangles (each hit point is re-derived on its triangle's plane), every branch taken on
// the value, the continuations' directions and pdfs at their sampled values (sampling detached: each vertex moves
// along its fixed incoming direction, and its shading, emission and NEE follow the moved point). Tangents: origin
// along e1, e2, direction along e1, e2 (e1, e2 perpendicular to d0). Returns the radiance value; ws gets the four
// tangents of sum_ch delta_ch L_ch.
template <bool INST>
__device__ __forceinline__ v3 path_walk_tan(const SceneDev &S, const DevCam &C, const PixSeeds &P, uint32_t i,
                                            v3 o0, v3 d0, v3 e1, v3 e2, v3 delta, Stack &st, uint32_t &nrays,
                                            float ws[ND]) {
  const uint32_t hs = P.hs;
  const float time = sample_time(C, P, i);
  DV o = dv(o0), d = dv(d0);
  o.x.g[0] = e1.x; o.y.g[0] = e1.y; o.z.g[0] = e1.z;
  o.x.g[1] = e2.x; o.y.g[1] = e2.y; o.z.g[1] = e2.z;
  d.x.g[2] = e1.x; d.y.g[2] = e1.y; d.z.g[2] = e1.z;
  d.x.g[3] = e2.x; d.y.g[3] = e2.y; d.z.g[3] = e2.z;
  DV thr = dv(mk(1.f, 1.f, 1.f));
  v3 L = mk(0.f, 0.f, 0.f);
  D Ls = dk(0.0f);                // sum_ch delta_ch L_ch
  D pB = dk(0.0f);
  const uint32_t nl = C.nl < 1 ? 1 : (C.nl > 16 ? 16 : C.nl);
  const float fnl = (float)nl;
  const uint32_t Dp = C.depth < 1 ? 1 : C.depth;
  const float A = S.emit_area;
  for (uint32_t b = 0; b < Dp; b++) {
    Hit h;
    nrays++;
    if (!trace_scene<false, INST>(S, val(o), val(d), 0.0f, INFINITY, time, b == 0 ? C.skip : 0u, h, st)) break;
    const float4 *sh = S.tshade + 4 * (size_t)h.tri;
    const float4 n0 = sh[0];
    const uint4 ms = S.meshes[__float_as_uint(n0.w)];
    const float4 *tp = S.tri + 3 * (size_t)h.tri;
    v3 p0 = ld3(tp[0]), e1 = ld3(tp[1]) - p0, e2 = ld3(tp[2]) - p0;
    v3 nl0 = ld3(n0), nl1 = ld3(sh[1]), nl2 = ld3(sh[2]), ngl = ld3(sh[3]);
    if (INST && h.inst != NO_INST) {
      // a posed body: its triangle and normals in the world at the sample's time (the pose does not depend on the
      // camera)
      const Pose Pb = inst_pose(S, h.inst, time);
      p0 = rot(Pb, p0) + Pb.p;
      e1 = rot(Pb, e1);
      e2 = rot(Pb, e2);
      nl0 = rot(Pb, nl0); nl1 = rot(Pb, nl1); nl2 = rot(Pb, nl2); ngl = rot(Pb, ngl);
    }
    const v3 nt = cross(e1, e2);
    // the hit on the triangle's plane and its barycentrics, as functions of the ray
    const D t = ddot(p0 - o, nt) / ddot(d, nt);
    const DV p = o + d * t;
    const DV q = p - p0;
    const float inn = 1.0f / dot(nt, nt);
    const D bu = ddot(dcross(q, dv(e2)), nt) * inn;
    const D bv = ddot(dcross(dv(e1), q), nt) * inn;
    v3 ng = ngl;
    const DV nsi = dnormalize(nl0 * (1.0f - bu - bv) + nl1 * bu + nl2 * bv);
    DV ns = nsi;
    const v3 dval = val(d);
    if (dot(ng, dval) > 0.0f) ng = -ng;
    if (dot(val(ns), ng) < 0.0f) ns = -ns;
    if (-dot(val(ns), dval) < 1e-4f) ns = dv(ng);
    if (ms.y & F_EMITTER) {
      const D cosE = -ddot(d, nsi);
      if (cosE.v > 0.0f) {
        const v3 Le = ld3(S.lights[ms.z]);
        const D pe = d_light_profile(S, ms.z, d_emit_local(emitter_pose(S, time), -d));
        D w = dk(1.0f);
        if (b > 0) {
          const D pL = t * t / (cosE * A);
          const D pb2 = pB * pB;
          w = pb2 / (pb2 + (fnl * fnl) * (pL * pL));
        }
        const DV c = (thr * Le) * pe;
        L = L + val(c) * w.v;
        Ls = Ls + (c.x * delta.x + c.y * delta.y + c.z * delta.z) * w;
      }
      break;
    }
    const DV wo = -d;
    const bool last = b + 1 == Dp;
    DV Ln = dv(mk(0.f, 0.f, 0.f));
    if (S.emit_count > 0) {
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
        const EmitPose E = emitter_pose(S, time);
        const uint32_t e = sample_emitter(S, E, u0, u1, u2, pl, nle);
        DV wi = pl - p;
        const D d2 = ddot(wi, wi);
        const D dist = dsqrt(d2);
        wi = wi * rcp(dist);
        const D cosE = -ddot(wi, nle);
        if (cosE.v <= 0.0f) continue;
        const D cosS = ddot(ns, wi);
        const v3 wiv = val(wi);
        if (cosS.v <= 0.0f || dot(ng, wiv) <= 0.0f) continue;
        const uint32_t lid = __float_as_uint(S.emit[4 * e + 2].w);
        const D pe = d_light_profile(S, lid, d_emit_local(E, -wi));
        const v3 Le = ld3(S.lights[lid]) * pe.v;
        if (maxc(Le) <= 0.0f) continue;
        const Mat M = load_mat(S.mats, ms.x);
        const DV f = d_eval_bsdf(M, ns, wi, wo);
        if (maxc(val(f)) <= 0.0f) continue;
        const D pW = d2 / (cosE * A);
        D w = dk(1.0f);
        if (!last) {
          const D pb = d_pdf_bsdf(M, ns, wo, wi);
          const D a2 = (fnl * fnl) * (pW * pW);
          w = a2 / (a2 + pb * pb);
        }
        const D k = cosS * w / pW;
        const DV contrib = (f * ld3(S.lights[lid])) * (k * pe);
        const v3 so = val(p) + ng * (dot(ng, wiv) > 0.0f ? EPS_OFF : -EPS_OFF);
        Hit sh;
        nrays++;
        if (trace_scene<true, INST>(S, so, wiv, 0.0f, dist.v - 2e-5f, time, S.shadow_skip, sh, st)) continue;
        Ln = Ln + contrib;
      }
    }
    {
      const DV c = thr * Ln * (1.0f / fnl);
      L = L + val(c);
      Ls = Ls + c.x * delta.x + c.y * delta.y + c.z * delta.z;
    }
    if (last) break;
    float u0, u1, u2;
    if (b == 0) {
      u0 = kr(i, KB0, P.s[6]); u1 = kr(i, KB1, P.s[7]); u2 = kr(i, KB2, P.s[8]);
    } else {
      u0 = u01(hash3(hs, i, 256u * b + 200u));
      u1 = u01(hash3(hs, i, 256u * b + 201u));
      u2 = u01(hash3(hs, i, 256u * b + 202u));
    }
    const Mat M = load_mat(S.mats, ms.x);
    // the continuation's direction and pdf are the sample's values (sampling detached): the vertex moves with the
    // camera along the fixed world direction, whose weight f cos / pdf follows the moved point's shading normal and
    // incoming direction; turning the direction with the camera would feed back through every later bounce
    v3 lv;
    const float pdf = sample_bsdf(M, val(ns), val(wo), u0, u1, u2, lv);
    if (!(pdf > 0.0f) || dot(ng, lv) <= 0.0f) break;
    const DV l = dv(lv);
    const DV f = d_eval_bsdf(M, ns, l, wo);
    const D cp = ddot(ns, l) * (1.0f / pdf);
    thr = thr * (f * cp);
    if (maxc(val(thr)) <= 0.0f) break;
    o = p + ng * (dot(ng, lv) > 0.0f ? EPS_OFF : -EPS_OFF);
    d = l;
    pB = dk(pdf);
  }
#pragma unroll
  for (int k = 0; k < ND; k++) ws[k] = Ls.g[k];
  return L;
}

__device__ __forceinline__ float wave_sum_xor(float v, uint32_t width) {
  for (uint32_t off = 1; off < width; off <<= 1) v += __shfl_xor(v, (int)off, 64);
  return v;
}

// one 64-item wave of the queue
// nrays_acc / nsamp_acc: the lane's running counts, added to the launch counters once when the wave leaves
// REF: the launch's cameras use the reference estimator (a launch holds cameras of one estimator)
template <bool INST, bool REF>
__device__ __forceinline__ void run_wave(const RenderLaunch &Lc, uint64_t wave, Stack &st, uint32_t &nrays_acc,
                                         uint32_t &nsamp_acc) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t g0 = Lc.item_base + wave * 64ull;
  // the wave's camera: item ranges are multiples of 64, so every lane shares it
  uint32_t c = 0;
  while (c + 1 < Lc.ncam && g0 >= Lc.cams[c + 1].item_begin) c++;
  const DevCam &C = Lc.cams[c];
  const uint64_t q = g0 - C.item_begin + lane;
  const uint32_t nspp = C.nspp;
  const uint64_t pix = q / nspp;
  const uint32_t i = C.spp_begin + (uint32_t)(q - pix * nspp);
  const uint32_t t = (uint32_t)(pix >> 6), r = (uint32_t)(pix & 63u);
  const uint32_t x = (t % C.tiles_x) * 8u + (r & 7u), y = (t / C.tiles_x) * 8u + (r >> 3);
  const bool valid = x < C.width && y < C.height;
  v3 L = mk(0.f, 0.f, 0.f);
  uint32_t nrays = 0;
  if (valid) {
    PixSeeds P;
    pix_seeds(C, x, y, P);
    if constexpr (REF) {
      uint32_t key;
      L = ref_radiance<INST>(Lc.S, C, P, x, y, i, C.nl, st, nrays, key);
      if (C.keys) {
        // the pixel's first-pass key (camMain, kernels.wgsl.js:202-206): sub-sample 0's mesh, the mesh range over
        // the sub-samples (a change marks an edge), the body / traced-shadow bits of any
        uint4 *k = C.keys + (size_t)y * C.width + x;
        const uint32_t m = key & 0xffffu;
        if (i == 0) k->x = m;
        atomicMin(&k->y, m);
        atomicMax(&k->z, m);
        if (key & 0x30000u) atomicOr(&k->w, key & 0x30000u);
      }
    } else {
      L = path_radiance<INST>(Lc.S, C, P, x, y, i, st, nrays);
    }
  }
  // per-pixel sums: aligned power-of-two runs reduce by butterfly; other run shapes add atomically
  const bool pow2 = (nspp & (nspp - 1)) == 0 && nspp <= 64;
  if (pow2) {
    float sx = wave_sum_xor(L.x, nspp), sy = wave_sum_xor(L.y, nspp), sz = wave_sum_xor(L.z, nspp);
    if (valid && (lane & (nspp - 1)) == 0) {
      float4 *a = C.acc + (size_t)y * C.width + x;
      float4 v = *a;
      v.x += sx; v.y += sy; v.z += sz; v.w += (float)nspp;
      *a = v;
    }
  } else if (valid) {
    float4 *a = C.acc + (size_t)y * C.width + x;
    atomicAdd(&a->x, L.x);
    atomicAdd(&a->y, L.y);
    atomicAdd(&a->z, L.z);
    atomicAdd(&a->w, 1.0f);
  }
  nrays_acc += nrays;
  nsamp_acc += valid ? 1u : 0u;
}

// a wave's ray and sample counts into the launch counters (one atomic pair per wave and launch)
__device__ __forceinline__ void flush_counts(const RenderLaunch &Lc, uint32_t nrays, uint32_t nsamp) {
  unsigned long long rs = nrays, ns = nsamp;
  for (int off = 32; off >= 1; off >>= 1) {
    rs += __shfl_xor(rs, off, 64);
    ns += __shfl_xor(ns, off, 64);
  }
  if ((threadIdx.x & 63u) == 0) {
    if (rs) atomicAdd(&Lc.counters[1], rs);
    if (ns) atomicAdd(&Lc.counters[2], ns);
  }
}

// the sample of lane `lane` in queue wave `wave`: its camera (index c in the launch), pixel and sub-sample
struct WaveItem {
  uint32_t c, x, y, i;
  bool valid;
};
__device__ __forceinline__ WaveItem wave_item(const RenderLaunch &Lc, uint64_t wave, uint32_t lane) {
  const uint64_t g0 = Lc.item_base + wave * 64ull;
  WaveItem it;
  it.c = 0;
  while (it.c + 1 < Lc.ncam && g0 >= Lc.cams[it.c + 1].item_begin) it.c++;
  const DevCam &C = Lc.cams[it.c];
  const uint64_t q = g0 - C.item_begin + lane;
  const uint64_t pix = q / C.nspp;
  it.i = C.spp_begin + (uint32_t)(q - pix * C.nspp);
  const uint32_t t = (uint32_t)(pix >> 6), r = (uint32_t)(pix & 63u);
  it.x = (t % C.tiles_x) * 8u + (r & 7u);
  it.y = (t / C.tiles_x) * 8u + (r >> 3);
  it.valid = it.x < C.width && it.y < C.height;
  return it;
}

// backward wave: replay each sample's path twice (its value, then the gradient walk)
template <bool INST>
__device__ __forceinline__ void run_wave_bwd(const RenderLaunch &Lc, uint64_t wave, Stack &st, GradSink &gs) {
  const uint32_t lane = threadIdx.x & 63u;
  const WaveItem it = wave_item(Lc, wave, lane);
  const DevCam &C = Lc.cams[it.c];
  uint32_t nrays = 0;
  if (it.valid) {
    const float4 a = C.adj[(size_t)it.y * C.width + it.x];
    const float s = C.scale / (float)C.nspp;
    gs.delta = mk(a.x * s, a.y * s, a.z * s);
    if (gs.delta.x != 0.0f || gs.delta.y != 0.0f || gs.delta.z != 0.0f) {
      PixSeeds P;
      pix_seeds(C, it.x, it.y, P);
      v3 Ltot;
      if (Lc.ltot) {
        const float4 l4 = Lc.ltot[wave * 64 + lane];
        Ltot = mk(l4.x, l4.y, l4.z);
      } else {
        Ltot = path_walk<false, INST>(Lc.S, C, P, it.x, it.y, it.i, st, nrays);
      }
      path_walk<true, INST>(Lc.S, C, P, it.x, it.y, it.i, st, nrays, &gs, Ltot);
    }
  }
  uint32_t rs = nrays;
  for (int off = 32; off >= 1; off >>= 1) rs += __shfl_xor(rs, off, 64);
  if (lane == 0) atomicAdd(&Lc.counters[1], (unsigned long long)rs);
}

// camera wave: the tangent walk of each sample, contracted with the ray's derivative with respect to its
// camera's intrinsics and pose; the wave's sums go to scam[camera * CAM_PARAMS + k]
template <bool INST>
__device__ __forceinline__ void run_wave_cam(const RenderLaunch &Lc, uint64_t wave, Stack &st, float *scam) {
  const uint32_t lane = threadIdx.x & 63u;
  const WaveItem it = wave_item(Lc, wave, lane);
  const DevCam &C = Lc.cams[it.c];
  uint32_t nrays = 0;
  float gc[CAM_PARAMS];
#pragma unroll
  for (int k = 0; k < CAM_PARAMS; k++) gc[k] = 0.0f;
  if (it.valid) {
    const float4 a = C.adj[(size_t)it.y * C.width + it.x];
    const float s = C.scale / (float)C.nspp;
    const v3 delta = mk(a.x * s, a.y * s, a.z * s);
    if (delta.x != 0.0f || delta.y != 0.0f || delta.z != 0.0f) {
      PixSeeds P;
      pix_seeds(C, it.x, it.y, P);
      const v3 o0 = mk(C.pos[0], C.pos[1], C.pos[2]);
      const v3 d0 = camera_dir(C, P, it.x, it.y, it.i);
      v3 e1, e2;
      basis(d0, e1, e2);
      float ws[ND];
      const v3 Lv = path_walk_tan<INST>(Lc.S, C, P, it.i, o0, d0, e1, e2, delta, st, nrays, ws);
      if (Lc.ltot) Lc.ltot[wave * 64 + lane] = make_float4(Lv.x, Lv.y, Lv.z, 0.0f);
      const v3 wo = e1 * ws[0] + e2 * ws[1];
      const v3 wd = e1 * ws[2] + e2 * ws[3];
      // direction: rotation r gives d' = d + r x d; intrinsics through d = R dc
      const v3 gr = cross(d0, wd);
      float u, v;
      camera_uv(P, it.x, it.y, it.i, u, v);
      v3 J[8];
      camdir_jac(C, u, v, J);
      const v3 rw = mk(C.R[0] * wd.x + C.R[1] * wd.y + C.R[2] * wd.z, C.R[3] * wd.x + C.R[4] * wd.y + C.R[5] * wd.z,
                       C.R[6] * wd.x + C.R[7] * wd.y + C.R[8] * wd.z);
#pragma unroll
      for (int k = 0; k < 8; k++) gc[k] = dot(rw, J[k]);
      gc[8] = gr.x; gc[9] = gr.y; gc[10] = gr.z;
      gc[11] = wo.x; gc[12] = wo.y; gc[13] = wo.z;
    }
  }
#pragma unroll
  for (int k = 0; k < CAM_PARAMS; k++) {
    float v = gc[k];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0 && v != 0.0f) atomicAdd(scam + it.c * CAM_PARAMS + k, v);
  }
  uint32_t rs = nrays;
  for (int off = 32; off >= 1; off >>= 1) rs += __shfl_xor(rs, off, 64);
  if (lane == 0) atomicAdd(&Lc.counters[1], (unsigned long long)rs);
}

template <int BLOCK, bool INST>
__global__ __launch_bounds__(BLOCK, 4) void k_backward(RenderLaunch Lc, float *gmat, float *glight, uint32_t n_mat,
                                                    uint32_t n_light, uint32_t use_lds) {
  __shared__ uint32_t s_node[STACK_LDS * BLOCK];
  __shared__ float s_t[STACK_LDS * BLOCK];
  extern __shared__ float sgrad[];
  const uint32_t ng = 3 * (n_mat + n_light);
  if (use_lds) {
    for (uint32_t k = threadIdx.x; k < ng; k += BLOCK) sgrad[k] = 0.0f;
  }
  __syncthreads();
  Stack st{(lds_u32 *)s_node, (lds_f32 *)s_t, Lc.spill, threadIdx.x, BLOCK, STACK_LDS, STACK_SPILL, Lc.spill_stride,
           (uint64_t)blockIdx.x * BLOCK + threadIdx.x, Lc.overflow};
  GradSink gs{use_lds ? sgrad : gmat, use_lds ? sgrad + 3 * n_mat : glight, v3{0.f, 0.f, 0.f}};
  for (;;) {
    unsigned long long base = 0;
    if ((threadIdx.x & 63u) == 0) base = atomicAdd(&Lc.counters[0], (unsigned long long)Lc.chunk);
    base = __shfl(base, 0, 64);
    if (base >= Lc.total_waves) break;
    const uint64_t end = base + Lc.chunk < Lc.total_waves ? base + Lc.chunk : Lc.total_waves;
    for (uint64_t w = base; w < end; w++) run_wave_bwd<INST>(Lc, w, st, gs);
  }
  __syncthreads();
  if (use_lds) {
    for (uint32_t k = threadIdx.x; k < ng; k += BLOCK) {
      const float v = sgrad[k];
      if (v != 0.0f) atomicAdd(k < 3 * n_mat ? gmat + k : glight + (k - 3 * n_mat), v);
    }
  }
}

// Built for 4 waves per SIMD: the tangent-carrying walk spills about half its registers there and still ran the
// differentiation frame's camera walk faster than at 2 waves (backward 957 -> 807 ms; 1 wave 1,428 ms, 3 waves
// 815 ms, 5 and 6 waves 1,430 ms)
template <int BLOCK, bool INST>
__global__ __launch_bounds__(BLOCK, 4) void k_camgrad(RenderLaunch Lc, float *gcam) {
  __shared__ uint32_t s_node[STACK_LDS * BLOCK];
  __shared__ float s_t[STACK_LDS * BLOCK];
  __shared__ float scam[MAX_CAMS * CAM_PARAMS];
  for (uint32_t k = threadIdx.x; k < Lc.ncam * CAM_PARAMS; k += BLOCK) scam[k] = 0.0f;
  __syncthreads();
  Stack st{(lds_u32 *)s_node, (lds_f32 *)s_t, Lc.spill, threadIdx.x, BLOCK, STACK_LDS, STACK_SPILL, Lc.spill_stride,
           (uint64_t)blockIdx.x * BLOCK + threadIdx.x, Lc.overflow};
  for (;;) {
    unsigned long long base = 0;
    if ((threadIdx.x & 63u) == 0) base = atomicAdd(&Lc.counters[0], (unsigned long long)Lc.chunk);
    base = __shfl(base, 0, 64);
    if (base >= Lc.total_waves) break;
    const uint64_t end = base + Lc.chunk < Lc.total_waves ? base + Lc.chunk : Lc.total_waves;
    for (uint64_t w = base; w < end; w++) run_wave_cam<INST>(Lc, w, st, scam);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < Lc.ncam * CAM_PARAMS; k += BLOCK)
    if (scam[k] != 0.0f) atomicAdd(gcam + k, scam[k]);
}

hipError_t launch_backward(const RenderLaunch &L, float *gmat, float *glight, uint32_t n_mat, uint32_t n_light,
                           hipStream_t s) {
  const size_t lds = (size_t)3 * (n_mat + n_light) * sizeof(float);
  const uint32_t use_lds = lds <= 16384 ? 1u : 0u;
  if (L.S.inst_count)
    hipLaunchKernelGGL((k_backward<256, true>), dim3(L.grid), dim3(256), use_lds ? lds : 0, s, L, gmat, glight, n_mat,
                       n_light, use_lds);
  else
    hipLaunchKernelGGL((k_backward<256, false>), dim3(L.grid), dim3(256), use_lds ? lds : 0, s, L, gmat, glight,
                       n_mat, n_light, use_lds);
  return hipGetLastError();
}

hipError_t launch_camgrad(const RenderLaunch &L, float *gcam, hipStream_t s) {
  if (L.ncam > MAX_CAMS) return hipErrorInvalidValue;
  if (L.S.inst_count)
    hipLaunchKernelGGL((k_camgrad<256, true>), dim3(L.grid), dim3(256), 0, s, L, gcam);
  else
    hipLaunchKernelGGL((k_camgrad<256, false>), dim3(L.grid), dim3(256), 0, s, L, gcam);
  return hipGetLastError();
}

template <int BLOCK, bool INST, bool REF>
__global__ __launch_bounds__(BLOCK, 4) void k_render(RenderLaunch Lc) {
  // the per-lane stacks: [stack_lds * BLOCK] node ids, then [stack_lds * BLOCK] entry distances
  extern __shared__ uint32_t s_stack[];
  Stack st;
  st.lds_node = (lds_u32 *)s_stack;
  st.lds_t = (lds_f32 *)(s_stack + Lc.stack_lds * BLOCK);
  st.spill = Lc.spill;
  st.tid = threadIdx.x;
  st.block = BLOCK;
  st.slds = Lc.stack_lds;
  st.spill_depth = STACK_TOTAL - Lc.stack_lds;
  st.spill_stride = Lc.spill_stride;
  st.gtid = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  st.overflow = Lc.overflow;
  const uint32_t wib = threadIdx.x >> 6;
  uint32_t nrays = 0, nsamp = 0;
  if (Lc.persistent) {
    // Queue q owns waves [total * q / nq, total * (q + 1) / nq). A block starts on its XCD's queue (blocks
    // are dispatched round-robin over the XCDs), so neighbouring tiles share that XCD's L2, and moves on to
    // the other queues when its own is empty. Every wave leaves once all nq queues are exhausted.
    const uint32_t nq = Lc.queues;
    const uint32_t home = nq > 1 ? blockIdx.x % nq : 0u;
    for (uint32_t k = 0; k < nq; k++) {
      const uint32_t qi = (home + k) % nq;
      const uint64_t q0 = Lc.total_waves * qi / nq, q1 = Lc.total_waves * (qi + 1) / nq;
      unsigned long long *head = nq > 1 ? &Lc.counters[8 + qi] : &Lc.counters[0];
      for (;;) {
        unsigned long long base = 0;
        if ((threadIdx.x & 63u) == 0) base = atomicAdd(head, (unsigned long long)Lc.chunk);
        base = __shfl(base, 0, 64);
        if (base >= q1 - q0) break;
        const uint64_t b = q0 + base;
        const uint64_t end = b + Lc.chunk < q1 ? b + Lc.chunk : q1;
        for (uint64_t w = b; w < end; w++) run_wave<INST, REF>(Lc, w * Lc.wave_stride, st, nrays, nsamp);
      }
    }
  } else {
    // static schedule: the waves of the grid stride over the queue
    const uint64_t stride = (uint64_t)gridDim.x * (BLOCK / 64);
    for (uint64_t w = (uint64_t)blockIdx.x * (BLOCK / 64) + wib; w < Lc.total_waves; w += stride)
      run_wave<INST, REF>(Lc, w * Lc.wave_stride, st, nrays, nsamp);
  }
  flush_counts(Lc, nrays, nsamp);
}

// The reference estimator's refinement pass (camRefine, kernels.wgsl.js:218-252). First each pixel decides from the
// first pass's keys: an edge (another first-hit mesh among its sub-samples or in its 3x3 block), a posed body, or
// (VT_OPT_REFINE_SHADOWS) traced shadow rays; the pixels that refine are appended to one list for every camera.
__global__ __launch_bounds__(256) void k_refine_mark(const DevCam *cam, uint32_t c, uint2 *list, uint32_t *count) {
  // the block's 16 x 16 pixels and a one-pixel rim of sub-sample-0 meshes in LDS (edge-clamped, as the 3x3 test)
  __shared__ uint32_t sk[18][18];
  __shared__ uint32_t s_wave[4], s_base;
  const DevCam &C = *cam;
  const uint32_t W = C.width, H = C.height;
  const int x0 = (int)blockIdx.x * 16 - 1, y0 = (int)blockIdx.y * 16 - 1;
  for (uint32_t q = threadIdx.x; q < 18 * 18; q += 256) {
    const int xx = min(max(x0 + (int)(q % 18), 0), (int)W - 1), yy = min(max(y0 + (int)(q / 18), 0), (int)H - 1);
    sk[q / 18][q % 18] = C.keys[(size_t)yy * W + xx].x;
  }
  __syncthreads();
  const uint32_t tx = threadIdx.x & 15u, ty = threadIdx.x >> 4;
  const uint32_t x = blockIdx.x * 16 + tx, y = blockIdx.y * 16 + ty;
  bool refine = false;
  if (x < W && y < H) {
    const uint4 k = C.keys[(size_t)y * W + x];
    refine = (k.w & 0x10000u) || k.y != k.z || ((k.w & 0x20000u) && (C.opts & OPT_REFINE_SHADOWS));
    for (uint32_t dy = 0; dy < 3; dy++)
      for (uint32_t dx = 0; dx < 3; dx++) refine |= sk[ty + dy][tx + dx] != k.x;
  }
  // one append per block (same-address atomics serialise across the chip): each wave counts its refining lanes,
  // the first thread reserves the block's entries, every lane writes at its rank in block order
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t m = __ballot(refine);
  if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t tot = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    s_base = tot ? atomicAdd(count, tot) : 0u;
  }
  __syncthreads();
  if (!refine) return;
  uint32_t at = s_base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  for (uint32_t w = 0; w < wave; w++) at += s_wave[w];
  list[at] = make_uint2(c, y * W + x);
}

// Then every listed pixel takes `refine` more sub-samples (indices spp .. spp + refine - 1, nl_refine light samples
// each): one lane per sub-sample, a pixel's sub-samples on consecutive lanes (groups of R = the frame's largest
// refine), and the group's first lane adds them to the first pass's sum in sub-sample order.
template <bool INST>
__global__ __launch_bounds__(64, 4) void k_refine(RenderLaunch Lc, const uint2 *list, const uint32_t *count, uint32_t R) {
  extern __shared__ uint32_t s_stack[];
  Stack st;
  st.lds_node = (lds_u32 *)s_stack;
  st.lds_t = (lds_f32 *)(s_stack + Lc.stack_lds * 64);
  st.spill = Lc.spill;
  st.tid = threadIdx.x;
  st.block = 64;
  st.slds = Lc.stack_lds;
  st.spill_depth = STACK_TOTAL - Lc.stack_lds;
  st.spill_stride = Lc.spill_stride;
  st.gtid = (uint64_t)blockIdx.x * 64 + threadIdx.x;
  st.overflow = Lc.overflow;
  const uint32_t lane = threadIdx.x, G = 64u / R, grp = lane / R, k = lane - grp * R;
  const uint32_t n = *count;
  uint32_t nrays = 0, nsamp = 0;
  for (uint32_t g0 = blockIdx.x * G; g0 < n; g0 += gridDim.x * G) {
    const uint32_t g = g0 + grp;
    const bool in = grp < G && g < n;
    uint2 item = make_uint2(0u, 0u);
    if (in) item = list[g];
    const DevCam &C = Lc.cams[item.x];
    const uint32_t x = item.y % C.width, y = item.y / C.width;
    const bool run = in && k < C.refine;
    v3 L = mk(0.f, 0.f, 0.f);
    if (run) {
      PixSeeds P;
      pix_seeds(C, x, y, P);
      uint32_t key;
      L = ref_radiance<INST>(Lc.S, C, P, x, y, C.nspp + k, C.nl_refine, st, nrays, key);
    }
    // the group's sum in sub-sample order, onto the first pass's (every lane takes part in the shuffles)
    const uint32_t base = grp * R;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    const bool writer = in && k == 0;
    if (writer) v = C.acc[item.y];
    for (uint32_t j = 0; j < R; j++) {
      const int src = (int)min(base + j, 63u);
      const float lx = __shfl(L.x, src, 64), ly = __shfl(L.y, src, 64), lz = __shfl(L.z, src, 64);
      if (j < C.refine) {
        v.x += lx;
        v.y += ly;
        v.z += lz;
      }
    }
    if (writer) {
      v.w += (float)C.refine;
      C.acc[item.y] = v;
      nsamp += C.refine;
    }
  }
  flush_counts(Lc, nrays, nsamp);
}

__global__ void k_fill_u4(uint4 *p, uint64_t n, uint4 v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// per-slot shading record: vertex normals (gathered through the slot's vertex ids), mesh id, geometric normal
__global__ void k_tri_shading(const float4 *tri, const uint4 *attr, const float4 *vnorm, uint32_t T, float4 *out) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < T; k += gridDim.x * blockDim.x) {
    const uint4 at = attr[k];
    const float4 *tp = tri + 3 * (size_t)k;
    const v3 p0 = ld3(tp[0]);
    const v3 ng = normalize(cross(ld3(tp[1]) - p0, ld3(tp[2]) - p0));
    const float4 a = vnorm[at.x], b = vnorm[at.y], c = vnorm[at.z];
    out[4 * (size_t)k] = make_float4(a.x, a.y, a.z, __uint_as_float(at.w));
    out[4 * (size_t)k + 1] = make_float4(b.x, b.y, b.z, 0.0f);
    out[4 * (size_t)k + 2] = make_float4(c.x, c.y, c.z, 0.0f);
    out[4 * (size_t)k + 3] = make_float4(ng.x, ng.y, ng.z, 0.0f);
  }
}
hipError_t launch_tri_shading(const float4 *tri, const uint4 *attr, const float4 *vnorm, uint32_t T, float4 *out,
                              hipStream_t s) {
  const uint32_t g = (T + 255) / 256 < 65535u ? (T + 255) / 256 : 65535u;
  hipLaunchKernelGGL(k_tri_shading, dim3(g ? g : 1), dim3(256), 0, s, tri, attr, vnorm, T, out);
  return hipGetLastError();
}

// y = x over n float4 (a plain copy: the bandwidth probe)
__global__ void k_copy_f4(const float4 *__restrict__ x, float4 *__restrict__ y, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    y[i] = x[i];
}
hipError_t launch_copy_f4(const float4 *x, float4 *y, uint64_t n, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_f4, dim3(grid), dim3(256), 0, s, x, y, n);
  return hipGetLastError();
}

// reads n uint4 once (grid-stride, four 16-B loads in flight per lane) and folds them into one word, written only
// when it equals a value the probe's fill never produces: the read-bandwidth probe
__global__ void k_read_u4(const uint4 *__restrict__ x, uint64_t n, uint32_t *sink) {
  uint32_t f = 0;
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x * 4;
  for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x * 4 + threadIdx.x; b < n; b += step) {
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint64_t i = b + (uint64_t)u * blockDim.x;
      v[u] = i < n ? x[i] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) f ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  }
  if (f == 0x9e3779b9u) sink[0] = f;
}
hipError_t launch_read_u4(const uint4 *x, uint64_t n, uint32_t grid, uint32_t *sink, hipStream_t s) {
  hipLaunchKernelGGL(k_read_u4, dim3(grid), dim3(256), 0, s, x, n, sink);
  return hipGetLastError();
}

__global__ void k_zero_u64(unsigned long long *p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}
__global__ void k_zero_f4(float4 *p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

hipError_t launch_zero_u64(unsigned long long *p, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}
hipError_t launch_zero_f4(float4 *p, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_f4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

namespace {
template <int BLOCK>
void (*render_kernel(bool inst, bool ref))(RenderLaunch) {
  if (ref) return inst ? k_render<BLOCK, true, true> : k_render<BLOCK, false, true>;
  return inst ? k_render<BLOCK, true, false> : k_render<BLOCK, false, false>;
}
}  // namespace

hipError_t launch_render(const RenderLaunch &L, hipStream_t s) {
  if (L.stack_lds < STACK_LDS_MIN || L.stack_lds > STACK_LDS_MAX || L.wave_stride == 0 ||
      (L.queues != 1 && L.queues != XCD_COUNT) || L.chunk == 0)
    return hipErrorInvalidValue;
  const size_t lds = (size_t)L.stack_lds * L.block * 8;
  const bool inst = L.S.inst_count != 0, ref = L.ref != 0;
  switch (L.block) {
    case 64: hipLaunchKernelGGL(render_kernel<64>(inst, ref), dim3(L.grid), dim3(64), lds, s, L); break;
    case 128: hipLaunchKernelGGL(render_kernel<128>(inst, ref), dim3(L.grid), dim3(128), lds, s, L); break;
    case 256: hipLaunchKernelGGL(render_kernel<256>(inst, ref), dim3(L.grid), dim3(256), lds, s, L); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_refine_mark(const DevCam *cam, uint32_t c, uint32_t width, uint32_t height, uint2 *list,
                              uint32_t *count, hipStream_t s) {
  hipLaunchKernelGGL(k_refine_mark, dim3((width + 15) / 16, (height + 15) / 16), dim3(256), 0, s, cam, c, list, count);
  return hipGetLastError();
}
hipError_t launch_refine(const RenderLaunch &L, const uint2 *list, const uint32_t *count, uint64_t max_items,
                         uint32_t R, hipStream_t s) {
  if (L.stack_lds < STACK_LDS_MIN || L.stack_lds > STACK_LDS_MAX || L.spill_stride < 64 || R == 0 || R > 64)
    return hipErrorInvalidValue;
  if (max_items == 0) return hipSuccess;
  const uint64_t G = 64 / R, want = (max_items + G - 1) / G;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(want, L.spill_stride / 64);
  const size_t lds = (size_t)L.stack_lds * 64 * 8;
  if (L.S.inst_count) hipLaunchKernelGGL(k_refine<true>, dim3(grid), dim3(64), lds, s, L, list, count, R);
  else hipLaunchKernelGGL(k_refine<false>, dim3(grid), dim3(64), lds, s, L, list, count, R);
  return hipGetLastError();
}
hipError_t launch_fill_u4(uint4 *p, uint64_t n, uint4 v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_keys(uint4 *keys, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, keys, n,
                     make_uint4(0xfffeu, 0xffffffffu, 0u, 0u));
  return hipGetLastError();
}

// resident blocks per CU of the render kernel at this block size and LDS stack depth
hipError_t render_occupancy(uint32_t block, uint32_t stack_lds, bool inst, bool ref, int *blocks_per_cu) {
  const size_t lds = (size_t)stack_lds * block * 8;
  switch (block) {
    case 64:
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, render_kernel<64>(inst, ref), 64, lds);
    case 128:
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, render_kernel<128>(inst, ref), 128, lds);
    case 256:
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, render_kernel<256>(inst, ref), 256, lds);
    default: return hipErrorInvalidValue;
  }
}

// ------------------------------------------------------------------ test kernels
__global__ __launch_bounds__(DEFAULT_BLOCK) void k_trace_test(TestLaunch Lt, const float *rays, uint32_t n,
                                                              int any_hit, float *hits) {
  __shared__ uint32_t s_node[STACK_LDS * DEFAULT_BLOCK];
  __shared__ float s_t[STACK_LDS * DEFAULT_BLOCK];
  Stack st{(lds_u32 *)s_node, (lds_f32 *)s_t, Lt.spill, threadIdx.x, DEFAULT_BLOCK, STACK_LDS, STACK_SPILL, Lt.spill_stride,
           (uint64_t)blockIdx.x * DEFAULT_BLOCK + threadIdx.x, Lt.overflow};
  for (uint32_t k = blockIdx.x * DEFAULT_BLOCK + threadIdx.x; k < n; k += gridDim.x * DEFAULT_BLOCK) {
    const float *r = rays + 8 * (size_t)k;
    Hit h;
    bool hit = any_hit ? traverse<true>(Lt.S, Lt.S.root, mk(r[0], r[1], r[2]), mk(r[4], r[5], r[6]), r[3], r[7], h, st)
                       : traverse<false>(Lt.S, Lt.S.root, mk(r[0], r[1], r[2]), mk(r[4], r[5], r[6]), r[3], r[7], h, st);
    float *o = hits + 4 * (size_t)k;
    o[0] = hit ? h.t : INFINITY;
    o[1] = __uint_as_float(hit ? h.tri : MISS);
    o[2] = hit ? h.u : 0.0f;
    o[3] = hit ? h.v : 0.0f;
  }
}
hipError_t launch_trace_test(const TestLaunch &L, const float *rays, uint32_t n, int any_hit, float *hits,
                             hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint32_t blocks = (n + DEFAULT_BLOCK - 1) / DEFAULT_BLOCK;
  if (blocks > L.max_blocks) blocks = L.max_blocks;
  hipLaunchKernelGGL(k_trace_test, dim3(blocks), dim3(DEFAULT_BLOCK), 0, s, L, rays, n, any_hit, hits);
  return hipGetLastError();
}

// visible (gpu/renderer.js:1238-1258): an any-hit query from the camera centre to tol short of each point, over the
// static scene and the posed bodies at exposure time `time`, the camera's own bodies skipped
template <bool INST>
__global__ __launch_bounds__(DEFAULT_BLOCK) void k_visible(TestLaunch Lt, float3 c, uint32_t skip, float tol,
                                                           float time, const float *pts, uint32_t n, uint8_t *out) {
  __shared__ uint32_t s_node[STACK_LDS * DEFAULT_BLOCK];
  __shared__ float s_t[STACK_LDS * DEFAULT_BLOCK];
  Stack st{(lds_u32 *)s_node, (lds_f32 *)s_t, Lt.spill, threadIdx.x, DEFAULT_BLOCK, STACK_LDS, STACK_SPILL, Lt.spill_stride,
           (uint64_t)blockIdx.x * DEFAULT_BLOCK + threadIdx.x, Lt.overflow};
  const v3 o = mk(c.x, c.y, c.z);
  for (uint32_t k = blockIdx.x * DEFAULT_BLOCK + threadIdx.x; k < n; k += gridDim.x * DEFAULT_BLOCK) {
    const v3 v = mk(pts[3 * (size_t)k], pts[3 * (size_t)k + 1], pts[3 * (size_t)k + 2]) - o;
    const float dist = sqrtf(dot(v, v));
    uint8_t vis = 1;
    if (dist > tol) {
      Hit h;
      vis = trace_scene<true, INST>(Lt.S, o, v * (1.0f / dist), 0.0f, dist - tol, time, skip, h, st) ? 0 : 1;
    }
    out[k] = vis;
  }
}
hipError_t launch_visible(const TestLaunch &L, const float c[3], uint32_t skip, float tol, float time, const float *pts,
                          uint32_t n, uint8_t *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint32_t blocks = (n + DEFAULT_BLOCK - 1) / DEFAULT_BLOCK;
  if (blocks > L.max_blocks) blocks = L.max_blocks;
  const float3 cc = make_float3(c[0], c[1], c[2]);
  if (L.S.inst_count)
    hipLaunchKernelGGL(k_visible<true>, dim3(blocks), dim3(DEFAULT_BLOCK), 0, s, L, cc, skip, tol, time, pts, n, out);
  else
    hipLaunchKernelGGL(k_visible<false>, dim3(blocks), dim3(DEFAULT_BLOCK), 0, s, L, cc, skip, tol, time, pts, n, out);
  return hipGetLastError();
}

__global__ void k_lens_test(const float *lens9, const float *uv, uint32_t n, float *rays) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float L[9];
  for (int j = 0; j < 9; j++) L[j] = lens9[j];
  v3 r = kb4_unproject(L, uv[2 * k], uv[2 * k + 1]);
  rays[3 * k] = r.x; rays[3 * k + 1] = r.y; rays[3 * k + 2] = r.z;
}
hipError_t launch_lens_test(const float *lens9, const float *uv, uint32_t n, float *rays, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_lens_test, dim3((n + 255) / 256), dim3(256), 0, s, lens9, uv, n, rays);
  return hipGetLastError();
}

__global__ void k_camray_test(const DevCam *cam, const uint32_t *xyi, uint32_t n, float *out) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const DevCam &C = *cam;
  PixSeeds P;
  pix_seeds(C, xyi[3 * k], xyi[3 * k + 1], P);
  v3 d = camera_dir(C, P, xyi[3 * k], xyi[3 * k + 1], xyi[3 * k + 2]);
  float *o = out + 6 * (size_t)k;
  o[0] = C.pos[0]; o[1] = C.pos[1]; o[2] = C.pos[2];
  o[3] = d.x; o[4] = d.y; o[5] = d.z;
}
hipError_t launch_camray_test(const DevCam *cam, const uint32_t *xyi, uint32_t n, float *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_camray_test, dim3((n + 255) / 256), dim3(256), 0, s, cam, xyi, n, out);
  return hipGetLastError();
}

// capsuleOcc on caller cases: 16 floats each (p.xyz, -, n.xyz, -, a.xyz, r, b.xyz, -) -> occlusion
__global__ void k_capsule_test(const float *cases, uint32_t n, float *out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float *c = cases + 16 * (size_t)k;
  out[k] = capsule_occ(mk(c[0], c[1], c[2]), mk(c[4], c[5], c[6]), mk(c[8], c[9], c[10]), mk(c[12], c[13], c[14]), c[11]);
}
hipError_t launch_capsule_test(const float *cases, uint32_t n, float *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_capsule_test, dim3((n + 255) / 256), dim3(256), 0, s, cases, n, out);
  return hipGetLastError();
}

__global__ void k_bsdf_test(const float4 *mats, const float *cases, uint32_t n, float *out) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float *c = cases + 16 * (size_t)k;
  v3 nn = mk(c[0], c[1], c[2]), l = mk(c[4], c[5], c[6]), v = mk(c[8], c[9], c[10]);
  const Mat M = load_mat(mats, (uint32_t)c[15]);
  v3 f = eval_bsdf(M, nn, l, v);
  float pdf = pdf_bsdf(M, nn, v, l);
  v3 ls = mk(0.f, 0.f, 0.f);
  float ps = sample_bsdf(M, nn, v, c[12], c[13], c[14], ls);
  float *o = out + 8 * (size_t)k;
  o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = pdf;
  o[4] = ps > 0.0f ? ls.x : 0.0f; o[5] = ps > 0.0f ? ls.y : 0.0f; o[6] = ps > 0.0f ? ls.z : 0.0f; o[7] = ps;
}
hipError_t launch_bsdf_test(const float4 *mats, const float *cases, uint32_t n, float *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bsdf_test, dim3((n + 255) / 256), dim3(256), 0, s, mats, cases, n, out);
  return hipGetLastError();
}

}  // namespace vt
