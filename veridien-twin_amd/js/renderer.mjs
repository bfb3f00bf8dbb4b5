// TwinRenderer over the MI355X library: the method names, argument meanings and return shapes of the
// reference's gpu/renderer.js (create / setPoses / renderCamera / renderFrame / sensor / readback / project /
// timings / idle / destroy), backed by the N-API addon napi/vt_napi.node and libvt.so's HIP kernels.
// There is no WebGPU device: `device` is a HIP device ordinal, and renderCamera returns an opaque handle where
// the reference returned a GPUTexture (gpu/renderer.js:855-866). readback returns fresh typed arrays as the
// reference does (gpu/renderer.js:1191-1231): 'linear' Float32Array w*h*4, 'raw' Float32Array DN, 'rgb8',
// 'rgba8', 'yuv420' {y, cb, cr}.
import { createRequire } from 'module';
import { sensorParams, AutoExposure, updateAWB, nativeSensor, nativeIsp } from './sensor.mjs';

const require = createRequire(import.meta.url);
const native = require('../napi/vt_napi.node');
/** triBody entry of a static-scene triangle */
export const STATIC_BODY = 0xffffffff;

// camera option bits (gpu/wgsl/common.wgsl.js:45-58), plus the estimator choice and the shadow culling
export const OPTION_BITS = Object.freeze({ NO_INDIRECT: 4, DIRECT_ONLY: 8, NO_CAPSULE: 32, NORMALS: 64, NO_MIS: 128,
  REFINE_SHADOW: 1024, HIT_CACHE: 2048, SHADOW_CACHE: 8192, REFERENCE: 1 << 16, CULL_SHADOWS: 1 << 17 });
export const DEFAULTS = Object.freeze({
  spp: 16,            // sub-samples per pixel
  maxDepth: 3,        // path vertices (camera hit = 1), path tracer only
  lightSamples: 1,    // emitter samples per path vertex
  seed: 0,
  batched: true,      // every camera of a frame in one persistent launch
  graph: true,        // batched frames run as a hipGraph, captured once and replayed
  estimator: 'path',  // 'path': depth-D path tracing; 'reference': the reference's shadeHit over the node cache
  gatherK: 64,        // node cache (gpu/renderer.js:31-58 DEFAULTS): gather rays per node
  nodeSamples: 4,     // emitter samples per node per update
  itersPerFrame: 4,   // warm-started Jacobi steps per pose change
  itersInitial: 96,   // on the first solve
  smoothSigma: 0.003, // Gaussian sigma (m) of the node irradiance smoothing; 0 = off
  smoothRadius: 0.006,
  nbrK: 16,
  cullShadows: true,  // skip shadow rays whose answer the three nodes and the bodies' capsules already give
  refine: 6,          // reference estimator: extra sub-samples on edges, posed bodies and traced-shadow pixels
  refineLightSamples: 1,
  refineShadows: true,
  hitCache: true,     // fixed cameras: each sub-sample's static first hit is traced once and kept (exact)
  shadowCache: true,  // still light: each sub-sample's static occlusion per light sample is kept (exact)
  interiorBVH: true,  // after the node cache build, walk a static tree without the shell's outer skin
  radiosityEvery: 1,  // update the node cache every n-th pose change
});
// the reference's operating points for its camera estimator (gpu/renderer.js:63-67), spelled out in full
export const PRESETS = Object.freeze({
  judge: { spp: 1, lightSamples: 4, refine: 6, refineLightSamples: 1, refineShadows: true, radiosityEvery: 1, itersPerFrame: 4 },
  balanced: { spp: 1, lightSamples: 2, refine: 2, refineLightSamples: 1, refineShadows: false, radiosityEvery: 2, itersPerFrame: 1 },
  live: { spp: 1, lightSamples: 1, refine: 0, radiosityEvery: 2, itersPerFrame: 1 },
});

const KIND = { linear: 0, raw: 1, ri: 2, rgb8: 3, rgba8: 4, yuv420: 5 };
const KIND_STATS = 6;
// Node 12 has no `??`: the same rule as a function
const dflt = (a, b) => (a === undefined || a === null ? b : a);
const MESH_EMITTER = 2;
const MESH_TEXTURED = 128;   // MESH_FLAG.TEXTURED (gpu/scene_gpu.js:14)

/** gpu/lens.js:8-16: scene lens {model, K, D, size} -> [fx fy cx cy k1..k4 alpha] */
export function lensParams(lens) {
  if (Array.isArray(lens) || ArrayBuffer.isView(lens)) return Array.from(lens);
  if (!lens || (lens.model && lens.model !== 'kb4')) throw new Error(`lens model ${lens && lens.model} is not kb4`);
  const K = lens.K, fx = K[0][0], fy = K[1][1];
  const D = (lens.D.flat ? lens.D.flat() : lens.D).map(Number);
  if (D.length !== 4) throw new Error('kb4 needs 4 coefficients');
  return [fx, fy, K[0][2], K[1][2], D[0], D[1], D[2], D[3], K[0][1] / fx];
}

const rowMajor4 = (T) => (Array.isArray(T[0]) ? T : [[T[0], T[4], T[8], T[12]], [T[1], T[5], T[9], T[13]], [T[2], T[6], T[10], T[14]], [T[3], T[7], T[11], T[15]]]);

/** KB4 projection (gpu/lens.js:43-51), FP64 */
export function projectKB4(L, X) {
  const [fx, fy, cx, cy, k1, k2, k3, k4, alpha] = L;
  const rho = Math.hypot(X[0], X[1]);
  if (rho < 1e-300) return [cx, cy];
  const th = Math.atan2(rho, X[2]), t2 = th * th;
  const s = th * (1 + t2 * (k1 + t2 * (k2 + t2 * (k3 + t2 * k4)))) / rho;
  const xd = X[0] * s, yd = X[1] * s;
  return [fx * (xd + alpha * yd) + cx, fy * yd + cy];
}

/**
 * The reference's prepared scene (gpu/scene_gpu.js:461-469 prepareScene output: nodes, tris, verts, meshTable,
 * materials, emitter, blas) -> this library's upload arrays: one light (L0) with the emitter's angular profile
 * p(theta) about its axis, the bodies' triangle ranges as triBody. The emitter records and axis are in the scope's
 * frame, as the reference's, and the light rides on the scope body (emitBody; shade.wgsl.js:31-47).
 */
/** cavityCapsule's cut (gpu/scene_gpu.js:486-503) for a body of `kind` through `port` (a scene.json ports[]
 * entry), as vt_instance's cavity fields: the port's pivot and inward axis, s >= 0 with the sheath's radius for a
 * cannula ('fixture'), s >= the cannula's toe for an instrument part or the scope; null without a port (lens fronts
 * and boards have no capsule at all). */
export function cavityOf(kind, port) {
  if (!port || kind === 'fixed' || kind === 'board') return null;
  const fixture = kind === 'fixture';
  const toe = Number(port.cannula_toe_insertion), rr = Number(port.cannula_outer_radius);
  return { origin: Array.from(port.pivot), axis: Array.from(port.axis_out, (x) => -x),
    s0: fixture || !Number.isFinite(toe) ? 0 : toe, r: fixture && Number.isFinite(rr) ? rr : 0 };
}

export function fromPrepared(S, ports = []) {
  const meshes = new Uint32Array(S.meshTable.length * 4);
  S.meshTable.forEach((m, i) => { meshes[4 * i] = m.mat; meshes[4 * i + 1] = m.flags >>> 0; meshes[4 * i + 2] = 0; });
  const materials = new Float32Array(S.materials.length * 16);
  S.materials.forEach((m, i) => {
    const kind = m.coated ? 2 : m.metallic ? 1 : 0;
    const c = kind ? (m.f0 || [0.56, 0.57, 0.58]) : (m.rho || [0.5, 0.5, 0.5]);
    // words 8-11: the coat (tau, alpha_c), or a printed material's ink (rho_d, roughInk^2; gpu/renderer.js:291-293)
    const x = m.coated || !m.texture ? [...(m.tau || [1, 1, 1]), m.alphaC || 0]
      : [...m.texture.ink, m.texture.roughInk * m.texture.roughInk];
    materials.set([c[0], c[1], c[2], m.alpha, m.ior || 1.5, kind, m.fbar || 0, m.roughness,
      x[0], x[1], x[2], x[3], m.fbarComputed || 0, 0, 0, 0], 16 * i);
  });
  // texture coordinates of the textured meshes' vertices: the reference keeps them in a dynamic vertex's spare
  // words (u in pos.w, v in normal.w; gpu/scene_gpu.js:390-392)
  let uvs;
  if (S.meshTable.some((m) => m.flags & MESH_TEXTURED)) {
    uvs = new Float32Array((S.verts.length / 8) * 2);
    const tu = new Uint32Array(S.tris.buffer, S.tris.byteOffset, S.tris.length);
    for (let k = 0; k < S.tris.length / 16; k++) {
      if (!(S.meshTable[tu[16 * k + 12]].flags & MESH_TEXTURED)) continue;
      for (const w of [3, 7, 11]) {
        const v = tu[16 * k + w];
        uvs[2 * v] = S.verts[8 * v + 3];
        uvs[2 * v + 1] = S.verts[8 * v + 7];
      }
    }
  }
  const em = S.emitter || { tris: [], area: 0, L0: [0, 0, 0] };
  const emit = new Float32Array(Math.max(1, em.tris.length) * 16), eu = new Uint32Array(emit.buffer);
  em.tris.forEach((t, k) => {
    emit.set([...t.v0, t.cdf, ...t.e1, t.area, ...t.e2, 0, ...t.n, 0], 16 * k);
    eu[16 * k + 11] = 0;
  });
  const L0 = em.L0 || [0, 0, 0];
  // the outlet's angular profile L_e = L0 p(theta from the axis) (gpu/scene_gpu.js:205-214, shade.wgsl.js:48-52)
  const profiles = em.theta && em.p && em.axis ? [{ light: 0, axis: em.axis, theta: em.theta, p: em.p }] : [];
  const out = {
    profiles,
    nodes: S.nodes, tris: S.tris, verts: S.verts, meshes, materials,
    lights: new Float32Array([L0[0], L0[1], L0[2], 0]), emit: em.tris.length ? emit : new Float32Array(0),
    emitArea: em.area || 0, ...(uvs ? { uvs } : {}),
  };
  // bodies (S.blas: contiguous triangle ranges in their own frames): the library builds every BVH itself
  if (S.blas && S.blas.length) {
    const triBody = new Uint32Array(S.tris.length / 16).fill(STATIC_BODY);
    S.blas.forEach((b, i) => triBody.fill(i, b.triOff, b.triOff + b.triCount));
    Object.assign(out, { nodes: undefined, triBody,
      bodies: S.blas.map((b) => {
        const oc = b.occCapsule || b.capsule;   // the occlusion capsule of capsuleVis, in the body frame
        const pid = ((S.bodies || []).find((x) => x.key === b.body) || {}).port;
        const cavity = cavityOf(b.kind, pid ? (ports || []).find((p) => p.id === pid) : null);
        return { key: b.body, kind: b.kind, owner: b.owner || null, rhoEff: b.rhoEff || 0,
          occ: oc ? { a: Array.from(oc.a), b: Array.from(oc.b), r: oc.r } : null, ...(cavity ? { cavity } : {}) };
      }) });
    // shadow rays leave through the scope's housing when the emitter sits at its front (gpu/renderer.js:648-649)
    const scopeI = S.blas.findIndex((b) => b.body === 'scope');
    out.shadowSkip = em.shadowSkipsScope && scopeI >= 0 ? (1 << scopeI) >>> 0 : 0;
    // the outlet's triangles are part of the scope's BLAS: the light rides on the scope (SceneU.info.y)
    if (scopeI >= 0 && em.tris.length) out.emitBody = scopeI;
  }
  return out;
}

/** Unit quaternion (x, y, z, w) of the rotation in the top-left 3x3 of a column-major 4x4 (Shepperd's method:
 * the largest of 1 + trace and the three 1 + 2 r_ii - trace picks the component to divide by). */
export function quatFromMat(m) {
  const r = (i, j) => m[4 * j + i];
  const tr = r(0, 0) + r(1, 1) + r(2, 2);
  const cand = [1 + tr, 1 + 2 * r(0, 0) - tr, 1 + 2 * r(1, 1) - tr, 1 + 2 * r(2, 2) - tr];
  let k = 0;
  for (let i = 1; i < 4; i++) if (cand[i] > cand[k]) k = i;
  const h = 0.5 * Math.sqrt(cand[k]), f = 0.25 / h;
  let q;
  if (k === 0) q = [(r(2, 1) - r(1, 2)) * f, (r(0, 2) - r(2, 0)) * f, (r(1, 0) - r(0, 1)) * f, h];
  else if (k === 1) q = [h, (r(0, 1) + r(1, 0)) * f, (r(0, 2) + r(2, 0)) * f, (r(2, 1) - r(1, 2)) * f];
  else if (k === 2) q = [(r(0, 1) + r(1, 0)) * f, h, (r(1, 2) + r(2, 1)) * f, (r(0, 2) - r(2, 0)) * f];
  else q = [(r(0, 2) + r(2, 0)) * f, (r(1, 2) + r(2, 1)) * f, h, (r(1, 0) - r(0, 1)) * f];
  const l = Math.hypot(...q);
  return q.map((x) => x / l);
}

export class TwinRenderer {
  /**
   * opts: {scene: upload arrays ({nodes?, tris, verts, meshes, materials, lights, emit, emitArea}) or
   * {prepared: the reference's prepareScene output, ports?: scene.json's ports}, cameras: [{name, lens (scene lens block or 9 numbers),
   * T_D_C (4x4) | {R (9, columns), pos}, size: [w, h]}], physics? (physics.yaml's sensor / isp blocks), sensor?
   * (overrides of sensorParams), device?, preset?, log?, ...DEFAULTS}
   */
  static async create(opts) {
    const r = new TwinRenderer();
    r._init(opts);
    return r;
  }

  _init(opts) {
    this.opts = { ...DEFAULTS, ...(PRESETS[opts.preset] || {}), ...opts };
    this.log = opts.log || (() => {});
    const src = opts.scene && opts.scene.prepared ? fromPrepared(opts.scene.prepared, opts.scene.ports) : opts.scene;
    if (!src || !src.tris) throw new Error('scene: pass the upload arrays or {prepared}');
    this.ctx = native.create(opts.device || 0);
    this.device = { hip: opts.device || 0 };
    this.bodies = src.triBody ? (src.bodies || []).map((b) => (typeof b === 'string' ? { key: b, kind: 'body' } : b)) : [];
    native.uploadScene(this.ctx, { ...src, bodyCount: this.bodies.length, shadowSkip: src.shadowSkip || 0 });
    this.stats = { triangles: src.tris.length / 16, bodies: this.bodies.length };
    for (const pr of src.profiles || []) this.setLightProfile(pr.light, pr);
    // the first board's print mask (_loadBoardTexture, gpu/renderer.js:395-412): {mask: Uint8Array, width, height}
    // decoded by the caller (the PNG's red channel, row 0 the image's top); one paper texel without it
    const bt = opts.boardTexture || (opts.scene && opts.scene.boardTexture);
    if (bt) this.setBoardTexture(bt.mask, bt.width, bt.height);
    this.materials = Float32Array.from(src.materials);
    this.lights = Float32Array.from(src.lights);
    this.sensorP = sensorParams(opts.physics || {}, opts.sensor || {});
    this.cameras = {};
    (opts.cameras || []).forEach((c, i) => this._addCamera(c, i));
    this.frame = 0;
    this.poses0 = null;
    this.poses1 = null;
    this.posesVersion = 0;
    this.solvedVersion = -1;
    this.radiosity = null;
    this.nodeCount = src.verts.length / 8;
    this.eAvg = [0, 0, 0];
    if (this.bodies.length) this.setPoses(null);
  }

  /** camera option bits of a render (gpu/renderer.js:922-933 _options) */
  _options(o = {}) {
    let bits = 0;
    if (o.noIndirect) bits |= OPTION_BITS.NO_INDIRECT;
    if (o.directOnly) bits |= OPTION_BITS.DIRECT_ONLY;
    if (o.noCapsule) bits |= OPTION_BITS.NO_CAPSULE;
    if (o.normals) bits |= OPTION_BITS.NORMALS;
    if (o.noMIS) bits |= OPTION_BITS.NO_MIS;
    if (dflt(o.estimator, this.opts.estimator) === 'reference') {
      bits |= OPTION_BITS.REFERENCE;
      if (!o.exactShadows && dflt(o.cullShadows, this.opts.cullShadows)) bits |= OPTION_BITS.CULL_SHADOWS;
      if (dflt(o.refineShadows, this.opts.refineShadows)) bits |= OPTION_BITS.REFINE_SHADOW;
      // the caches of a fixed camera (gpu/renderer.js:944-957): the occlusion cache only with the hit cache
      if (this.opts.hitCache !== false && o.hitCache !== false) {
        bits |= OPTION_BITS.HIT_CACHE;
        if (this.opts.shadowCache !== false && o.shadowCache !== false && !o.exactShadows) bits |= OPTION_BITS.SHADOW_CACHE;
      }
    }
    return bits >>> 0;
  }

  /**
   * Bodies whose pose record differs from the one of the last node update, as a bit mask (_movedSinceSolve,
   * gpu/renderer.js:756-769); null when nothing was solved yet.
   */
  _movedSinceSolve() {
    const last = this._solvedInst;
    if (!last) return null;
    let mask = 0;
    (this.instances || []).forEach((I, i) => {
      const L = last[i];
      const eq = (a, b, tol) => a.every((x, k) => Math.abs(x - b[k]) < tol);
      const same = (!I && !L) || (I && L && eq(I.q0, L.q0, 1e-9) && eq(I.q1, L.q1, 1e-9) && eq(I.p0, L.p0, 1e-10) &&
        eq(I.p1, L.p1, 1e-10));
      if (!same) mask |= 1 << i;
    });
    return mask >>> 0;
  }

  /**
   * The node cache for the current poses (_ensureRadiosity, gpu/renderer.js:771-795): built on first use (gather
   * table, activity, neighbours, albedo tables), then the direct light and itersInitial Jacobi steps; after a
   * pose change the direct light of the nodes near the moved bodies (or of every node when the light's body
   * moved) and itersPerFrame steps; nothing moved, no work. stats.lastRadiosity: 'full', 'incremental' or
   * 'reused'.
   */
  _ensureRadiosity() {
    const o = this.opts;
    if (!this.radiosity) {
      this.radiosity = native.radiosityBuild(this.ctx, { gatherK: o.gatherK, nodeSamples: o.nodeSamples, nbrK: o.nbrK,
        smoothRadius: o.smoothRadius, smoothSigma: o.smoothSigma });
      this.stats.activeNodes = this.radiosity.active;
      this.stats.gatherBuildMs = this.radiosity.ms;
      // the interior tree for every later walk (_buildInteriorBVH, gpu/renderer.js:529, 533-571)
      this.stats.interiorTris = o.interiorBVH !== false ? native.interiorBvh(this.ctx, true) : undefined;
    }
    if (this.solvedVersion === this.posesVersion && !this.shadeChanged) return;
    // radiosityEvery n: a pose change updates the cache every n-th time (gpu/renderer.js:36); a material, light or
    // profile change always does, with the initial iteration count (the solution moves as a whole)
    if (!this.shadeChanged && this.solvedVersion >= 0 &&
        this.posesVersion - this.solvedVersion < Math.max(1, o.radiosityEvery || 1)) return;
    const fresh = this.solvedVersion < 0 || this.shadeChanged;
    const moved = fresh ? null : this._movedSinceSolve();
    if (moved === null) {
      native.radiosityUpdate(this.ctx, true, fresh ? o.itersInitial : o.itersPerFrame);
      this.stats.lastRadiosity = 'full';
    } else if (moved) {
      const ran = native.radiosityUpdateMoved(this.ctx, moved, o.itersPerFrame);
      this.stats.lastRadiosity = ran === 'none' ? 'reused' : ran;
    } else this.stats.lastRadiosity = 'reused';
    if (moved === null || moved) {
      this._solvedInst = (this.instances || []).map((I) => (I ? { q0: [...I.q0], q1: [...I.q1], p0: [...I.p0],
        p1: [...I.p1] } : null));
    }
    this.solvedVersion = this.posesVersion;
    this.shadeChanged = false;
  }

  /** Jacobi steps to convergence (converge, gpu/renderer.js:804-822), then nodeSummary. */
  async converge(iters = 200) {
    if (typeof iters !== 'number' || !Number.isFinite(iters)) iters = iters === false ? 0 : 200;
    this._ensureRadiosity();
    native.radiosityUpdate(this.ctx, false, iters);
    return this.nodeSummary();
  }

  /** The node state and eAvg, the area-weighted mean irradiance (nodeSummary, gpu/renderer.js:824-852). */
  async nodeSummary() {
    if (!this.radiosity) this._ensureRadiosity();
    const r = native.radiosityState(this.ctx, this.nodeCount);
    this.eAvg = r.eAvg;
    return { eAvg: r.eAvg, state: r.state, activeNodes: this.radiosity.active };
  }

  _addCamera(c, index) {
    const lens = lensParams(c.lens);
    let R, pos;
    if (c.T_D_C) {
      const T = rowMajor4(c.T_D_C);
      R = [T[0][0], T[1][0], T[2][0], T[0][1], T[1][1], T[2][1], T[0][2], T[1][2], T[2][2]];
      pos = [T[0][3], T[1][3], T[2][3]];
    } else { R = Array.from(c.R); pos = Array.from(c.pos); }
    const size = c.size || (c.lens && c.lens.size) || [c.width, c.height];
    this.cameras[c.name] = { index, name: c.name, lens, R, pos, size, mode: c.mode || 0, manual: null, awb: null,
      ae: new AutoExposure(this.sensorP, { exposureUnits: this.sensorP.exposureUnits0 }) };
  }

  _slot(name) {
    const cam = this.cameras[name];
    if (!cam) throw new Error(`unknown camera ${name}`);
    return cam;
  }

  _pack(cam, o) {
    const spp = dflt(o.spp, this.opts.spp);
    native.setCamera(this.ctx, cam.index, {
      R: cam.R, pos: cam.pos, lens: cam.lens, mode: cam.mode, width: cam.size[0], height: cam.size[1], spp,
      seed: dflt(o.seed, this.opts.seed), camIndex: cam.index, lightSamples: dflt(o.lightSamples, this.opts.lightSamples),
      maxDepth: dflt(o.maxDepth, this.opts.maxDepth), scale: dflt(o.scale, 1), shutter: o.shutter || [0, 1],
      skipBodies: this._skipMask(cam.name), options: this._options(o),
      refine: dflt(o.estimator, this.opts.estimator) === 'reference' ? dflt(o.refine, this.opts.refine) : 0,
      refineLightSamples: dflt(o.refineLightSamples, this.opts.refineLightSamples),
    });
  }

  /** Angular emission profile of a light: {axis, theta (radians, ascending), p}; null restores a Lambertian light. */
  setLightProfile(light, profile) {
    native.setLightProfile(this.ctx, light, profile ? { axis: Array.from(profile.axis), theta: Array.from(profile.theta),
      p: Array.from(profile.p) } : null);
    this.shadeChanged = true;
  }

  /** The printed board's mask (_loadBoardTexture, gpu/renderer.js:395-412): width x height bytes, 255 = paper,
   * row 0 the image's top. Textured meshes sample it bilinearly (clamp to edge). */
  setBoardTexture(mask, width, height) {
    native.setBoardTexture(this.ctx, mask instanceof Uint8Array ? mask : Uint8Array.from(mask), width >>> 0, height >>> 0);
    this.stats.boardTexture = [width, height];
  }

  /** Bodies a camera's own rays ignore: those it owns, e.g. its lens front (gpu/renderer.js:881-886). */
  _skipMask(name) {
    let m = 0;
    this.bodies.forEach((b, i) => { if (b.owner && b.owner === name) m |= 1 << i; });
    return m >>> 0;
  }

  /**
   * visible(name, pointsD, {tol, time}) (gpu/renderer.js:1238-1258): per world point, whether nothing lies between
   * the camera centre and tol short of it (posed bodies at exposure time `time`, the camera's own bodies ignored).
   * One any-hit launch on the GPU answers the whole list.
   */
  visible(name, pointsD, { tol = 2e-4, time = 0.5 } = {}) {
    const cam = this._slot(name);
    this._pack(cam, {});
    const pts = new Float32Array(pointsD.length * 3);
    pointsD.forEach((p, k) => pts.set([p[0], p[1], p[2]], 3 * k));
    return Array.from(native.visible(this.ctx, cam.index, pts, tol, time), (v) => v === 1);
  }

  /**
   * poses.bodies[key] = column-major 4x4 T_D_body at shutter open (poses0) and close (poses1, default poses0)
   * (gpu/renderer.js:619-633). Fixed bodies keep the identity; a body without a pose is hidden. A sub-sample at
   * exposure time t (opts.shutter of a render, default [0, 1]) sees slerp / lerp of the two poses.
   */
  setPoses(poses0, poses1 = null) {
    const b0 = (poses0 && poses0.bodies) || {}, b1 = ((poses1 || poses0) && (poses1 || poses0).bodies) || {};
    for (const k of Object.keys(b0)) if (!this.bodies.some((b) => b.key === k)) throw new Error(`unknown body ${k}`);
    this.instances = this.bodies.map((bl) => {
      // capsuleVis's capsule: none for lens fronts and boards, cut to the cavity for a body through a port
      const flush = bl.kind === 'fixed' || bl.kind === 'board';
      const occ = bl.occ ? { occA: bl.occ.a, occB: bl.occ.b, occR: flush ? 0 : bl.occ.r, rhoEff: dflt(bl.rhoEff, 0),
        ...(bl.cavity ? { cavity: bl.cavity } : {}) } : {};
      if (bl.kind === 'fixed') return { q0: [0, 0, 0, 1], q1: [0, 0, 0, 1], p0: [0, 0, 0], p1: [0, 0, 0], ...occ };
      const T0 = b0[bl.key], T1 = b1[bl.key] || T0;
      if (!T0) return null;
      return { q0: quatFromMat(T0), q1: quatFromMat(T1), p0: [T0[12], T0[13], T0[14]], p1: [T1[12], T1[13], T1[14]], ...occ };
    });
    this.posesVersion = (this.posesVersion || 0) + 1;
    if (this.bodies.length) native.setInstances(this.ctx, this.instances);
    this.poses0 = poses0; this.poses1 = poses1 || poses0;
  }

  /** Render one camera; resolves when enqueued. Returns an opaque handle (the reference returns a GPUTexture). */
  async renderCamera(name, opts = {}) {
    const cam = this._slot(name);
    this._pack(cam, opts);
    native.render(this.ctx, [cam.index], { mode: 1 });
    if (opts.exposure !== undefined) this._manualExposure(name, opts.exposure);
    return { camera: name, width: cam.size[0], height: cam.size[1] };
  }

  /** Every listed camera in one batched launch (gpu/renderer.js:986-996); optional sensor stage per camera. */
  renderFrame(opts = {}) {
    const names = opts.cameras || Object.keys(this.cameras);
    const cams = names.map((n) => this._slot(n));
    if (dflt(opts.estimator, this.opts.estimator) === 'reference') this._ensureRadiosity();
    cams.forEach((c) => this._pack(c, opts));
    // batched frames run in the form autotune() found fastest (the persistent launch, as a hipGraph, until then)
    const batched = dflt(opts.batched, this.opts.batched);
    native.render(this.ctx, cams.map((c) => c.index),
      { mode: batched ? 3 : 1, graph: batched && dflt(opts.graph, this.opts.graph) ? 1 : 0 });
    const out = {};
    for (const c of cams) out[c.name] = { linear: { camera: c.name, width: c.size[0], height: c.size[1] } };
    if (opts.sensor) for (const c of cams) out[c.name].raw = this.sensor(c.name, opts.sensorOpts || {});
    this.frame++;
    out.done = native.sync(this.ctx);
    return out;
  }

  /** Manual exposure for a camera ({time_s, gain} or seconds), or back to auto with null / 'auto'
   * (gpu/renderer.js:1043-1047). */
  _manualExposure(name, e) {
    const cam = this._slot(name);
    if (e === null || e === 'auto') { cam.manual = null; return; }
    cam.manual = typeof e === 'number' ? { time_s: e, gain: 1 }
      : { time_s: dflt(e.time_s, dflt(e.exposure_s, e.time)), gain: dflt(e.gain, 1) };
  }

  /** The exposure the next sensor pass of this camera uses: auto exposure's or the manual one
   * (gpu/renderer.js:1050-1054). */
  exposureOf(name) {
    const cam = this._slot(name);
    if (cam.manual) return { ...cam.manual, auto: false };
    return { time_s: cam.ae.exposure, gain: cam.ae.gain, gainUnits: cam.ae.gainUnits, units: cam.ae.units, auto: true };
  }

  /**
   * Radiance -> 8-bit camera RGB (gpu/renderer.js:1056-1099): the raw stage (optics, glare, exposure, noise,
   * ADC) and the ISP (demosaic, WB + CCM, NR, tone, sharpening, BT.601 YCbCr). opts: {seed, frame, noise,
   * exposure_s, gain, params} with params overriding sensorParams fields. The frame's AE / AWB statistics are
   * read back without blocking and steer the next frame, as the UVC loop does.
   */
  sensor(name, opts = {}) {
    const cam = this._slot(name);
    const p = { ...this.sensorP, ...(opts.params || {}) };
    if (p.awb && cam.awb && !(opts.params && opts.params.wb)) p.wb = p.wb.map((g, c) => g * cam.awb[c]);
    if (opts.noise === false) p.noise = false;
    const ex = this.exposureOf(name);
    const [width, height] = cam.size;
    native.sensor(this.ctx, cam.index, nativeSensor(p, { width, height, frame: dflt(opts.frame, this.frame),
      seed: dflt(opts.seed, 0), exposure: dflt(opts.exposure_s, ex.time_s), gain: dflt(opts.gain, ex.gain) }));
    native.isp(this.ctx, cam.index, nativeIsp(p));
    this._afterSensor(name);
    return { camera: name, kind: 'rgba8', width, height };
  }

  /** AE / AWB from the frame's statistics once they arrive (gpu/renderer.js:1075-1089). */
  _afterSensor(name) {
    const cam = this.cameras[name], [w, h] = cam.size;
    cam.pendingStats = native.readback(this.ctx, cam.index, KIND_STATS, w, h).then((st) => {
      cam.meanLuma = st[0] / (w * h * 255);
      cam.meanRGB = [st[1], st[2], st[3]].map((x) => x / (w * h * 1023));
      if (!cam.manual) cam.ae.update(cam.meanLuma);
      if (this.sensorP.awb) cam.awb = updateAWB(cam.awb, cam.meanRGB, this.sensorP.awbDamping);
      return st;
    }).catch(() => null);
  }

  /** Fresh host copy (gpu/renderer.js:1191-1231): 'linear' Float32Array w*h*4, 'raw' Float32Array of integer DN,
   * 'rgb8' Uint8Array w*h*3, 'rgba8' Uint8Array w*h*4, 'yuv420' {width, height, y, cb, cr}. */
  async readback(name, kind = 'rgb8') {
    const cam = this._slot(name);
    if (!(kind in KIND)) throw new Error(`readback kind ${kind} is not supported by this build`);
    const [w, h] = cam.size;
    const data = await native.readback(this.ctx, cam.index, KIND[kind], w, h);
    if (kind === 'raw') return Float32Array.from(data);
    if (kind === 'yuv420') {
      const cw = Math.ceil(w / 2), ch = Math.ceil(h / 2);
      return { width: w, height: h, y: data.slice(0, w * h), cb: data.slice(w * h, w * h + cw * ch),
        cr: data.slice(w * h + cw * ch, w * h + 2 * cw * ch) };
    }
    return data;
  }

  /** KB4 projection of D-frame points into a camera (gpu/renderer.js:1261-1265), FP64. */
  project(name, pointsD) {
    const cam = this._slot(name), R = cam.R, c = cam.pos;
    return pointsD.map((p) => {
      const d = [p[0] - c[0], p[1] - c[1], p[2] - c[2]];
      const X = [R[0] * d[0] + R[1] * d[1] + R[2] * d[2], R[3] * d[0] + R[4] * d[1] + R[5] * d[2], R[6] * d[0] + R[7] * d[1] + R[8] * d[2]];
      return projectKB4(cam.lens, X);
    });
  }

  /** Pick block size, LDS stack depth, queue chunk and queue layout for a frame of these cameras (no reference
   * counterpart); later frames use the result. Resolves to the chosen shape and the probe rates. */
  async autotune(opts = {}) {
    const names = opts.cameras || Object.keys(this.cameras);
    const cams = names.map((n) => this._slot(n));
    if (dflt(opts.estimator, this.opts.estimator) === 'reference') this._ensureRadiosity();
    cams.forEach((c) => this._pack(c, opts));
    this.tuned = await native.autotune(this.ctx, cams.map((c) => c.index), { probeStride: opts.probeStride || 0 });
    return this.tuned;
  }

  /** Replace the material table (16 floats per material) or the light radiances (4 per light) in place. */
  updateMaterials(materials) {
    native.updateMaterials(this.ctx, Float32Array.from(materials));
    this.materials = Float32Array.from(materials);
    this.shadeChanged = true;      // the node cache re-solves before its next reference-estimator frame
  }
  updateLights(lights) {
    native.updateLights(this.ctx, Float32Array.from(lights));
    this.lights = Float32Array.from(lights);
    this.shadeChanged = true;
  }

  /** Change a camera's intrinsics (9 lens numbers) and / or pose (R: 9, columns; pos: 3) for the next frame. */
  setCameraParams(name, { lens, R, pos } = {}) {
    const cam = this._slot(name);
    if (lens) cam.lens = lensParams(lens);
    if (R) cam.R = Array.from(R);
    if (pos) cam.pos = Array.from(pos);
  }

  /**
   * Adjoint of the last frame's linear images: adjoints = {name: Float32Array w*h*4 (dLoss/dI, rgb used)} for
   * cameras rendered with the same options. Path-replay gradients with respect to the material colours and
   * light radiances, and (opts.cameras, default true) each camera's [fx fy cx cy k1 k2 k3 k4 rx ry rz tx ty tz]
   * (rotation about the world axes, centre). Interior term only. Resolves to {materials, lights, cameras}.
   */
  async backward(adjoints, opts = {}) {
    const names = Object.keys(adjoints);
    const cams = names.map((n) => this._slot(n));
    const withCams = dflt(opts.cameras, true);
    const g = await native.backward(this.ctx, cams.map((c) => c.index), names.map((n) => adjoints[n]),
      this.materials.length / 16, this.lights.length / 4, { cameras: withCams ? 1 : 0 });
    const out = { materials: g.materials, lights: g.lights, cameras: {} };
    if (withCams) names.forEach((n, k) => { out.cameras[n] = g.cameras.slice(14 * k, 14 * k + 14); });
    return out;
  }

  async timings() {
    const st = await native.stats(this.ctx);
    return { gpu: true, last: st, stats: this.stats };
  }

  async idle() { await native.sync(this.ctx); }

  destroy() { if (this.ctx) { native.destroy(this.ctx); this.ctx = null; } }
}

/** Rotation exp([r]x) (Rodrigues) applied to R (9 numbers, columns): the pose update of the camera adjoint. */
export function rotateColumns(R, r) {
  const th = Math.hypot(r[0], r[1], r[2]);
  if (th === 0) return Array.from(R);
  const k = [r[0] / th, r[1] / th, r[2] / th], c = Math.cos(th), s = Math.sin(th), v = 1 - c;
  const M = [[c + k[0] * k[0] * v, k[0] * k[1] * v - k[2] * s, k[0] * k[2] * v + k[1] * s],
    [k[1] * k[0] * v + k[2] * s, c + k[1] * k[1] * v, k[1] * k[2] * v - k[0] * s],
    [k[2] * k[0] * v - k[1] * s, k[2] * k[1] * v + k[0] * s, c + k[2] * k[2] * v]];
  const out = new Array(9);
  for (let col = 0; col < 3; col++)
    for (let i = 0; i < 3; i++) out[3 * col + i] = M[i][0] * R[3 * col] + M[i][1] * R[3 * col + 1] + M[i][2] * R[3 * col + 2];
  return out;
}

/**
 * Autograd-style differentiable render over a TwinRenderer: forward(params) renders the listed cameras and
 * resolves to their linear images; backward(gradImages) resolves to the gradients of sum(grad . image) with
 * respect to the material colours, the light radiances and each camera's 14 parameters, by the path-replay
 * adjoint on the same samples. params: {materials?, lights?, cameras?: {name: {lens?, R?, pos?}}}.
 */
export class RenderFunction {
  constructor(renderer, names, opts = {}) {
    this.r = renderer;
    this.names = names || Object.keys(renderer.cameras);
    this.opts = opts;
  }

  async forward(params = {}) {
    if (params.materials) this.r.updateMaterials(params.materials);
    if (params.lights) this.r.updateLights(params.lights);
    for (const [n, p] of Object.entries(params.cameras || {})) this.r.setCameraParams(n, p);
    const out = this.r.renderFrame({ ...this.opts, cameras: this.names });
    await out.done;
    const images = {};
    for (const n of this.names) images[n] = await this.r.readback(n, 'linear');
    return images;
  }

  async backward(gradImages) {
    return this.r.backward(gradImages, { cameras: dflt(this.opts.cameraGradients, true) });
  }
}

export default TwinRenderer;
