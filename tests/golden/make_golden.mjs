// Writes tests/golden/ref_vectors.json: input/output vectors of the reference's own CPU functions, made by
// running them under Node in the build container (where the reference is mounted). The fixture holds data
// only: the inputs (generated here with the reference's own seeded mulberry32) and the reference's outputs.
//
//   node tests/golden/make_golden.mjs [/root/reference]
//
// Modules used: control/rng.js (mulberry32, Rng), gpu/bvh.js (buildBVH, packTriangles, traverse, bruteForce,
// intersectTri, sahCost), gpu/shade_ref.js (profile, capsuleOcc, discIrradiance, gaussLegendre). They have no
// imports, so each loads from its own text as a data: URL module (Node 12 has no `type: module` without a
// package.json, and the reference ships none).
import fs from 'fs';
import path from 'path';
import url from 'url';

const REF = process.argv[2] || '/root/reference';
const here = path.dirname(url.fileURLToPath(import.meta.url));
const load = (rel) => import('data:text/javascript;base64,' + Buffer.from(fs.readFileSync(path.join(REF, rel))).toString('base64'));

async function main() {
const rngM = await load('control/rng.js');
const bvhM = await load('gpu/bvh.js');
const shM = await load('gpu/shade_ref.js');

const out = { source: 'reference CPU functions under Node ' + process.version, rng: {}, bvh: [], shade: {} };

// ---------------------------------------------------------------- rng
for (const seed of [0, 1, 1001, 0xdeadbeef, 4294967295]) {
  const f = rngM.mulberry32(seed);
  out.rng['mulberry32_' + seed] = Array.from({ length: 64 }, () => Math.round(f() * 4294967296));
}
{
  const r = new rngM.Rng(1003);
  out.rng.rng1003 = {
    uniform: Array.from({ length: 8 }, () => r.uniform(-2, 5)),
    int: Array.from({ length: 8 }, () => r.int(0, 15)),
    normal: Array.from({ length: 9 }, () => r.normal(1, 2)),
    pick: Array.from({ length: 8 }, () => r.pick([1, 4, 16, 64])),
  };
}

// ---------------------------------------------------------------- bvh
const f32hex = (a) => Buffer.from(new Float32Array(a).buffer).toString('hex');
const u32hex = (a) => Buffer.from(new Uint32Array(a).buffer).toString('hex');

function soup(seed, n, spread, size) {
  const r = new rngM.Rng(seed);
  const pos = new Float32Array(n * 9), idx = new Uint32Array(n * 3);
  for (let t = 0; t < n; t++) {
    const c = [r.uniform(-spread, spread), r.uniform(-spread, spread), r.uniform(-spread, spread)];
    for (let v = 0; v < 3; v++) for (let k = 0; k < 3; k++) pos[9 * t + 3 * v + k] = c[k] + r.uniform(-size, size);
    idx[3 * t] = 3 * t; idx[3 * t + 1] = 3 * t + 1; idx[3 * t + 2] = 3 * t + 2;
  }
  return { pos, idx };
}
function sphere(nu, nv, rad) {
  const P = [], I = [];
  for (let i = 0; i <= nv; i++) for (let j = 0; j < nu; j++) {
    const th = Math.PI * i / nv, ph = 2 * Math.PI * j / nu;
    P.push(rad * Math.sin(th) * Math.cos(ph), rad * Math.sin(th) * Math.sin(ph), rad * Math.cos(th));
  }
  for (let i = 0; i < nv; i++) for (let j = 0; j < nu; j++) {
    const a = i * nu + j, b = i * nu + (j + 1) % nu, c = (i + 1) * nu + j, d = (i + 1) * nu + (j + 1) % nu;
    I.push(a, c, b, b, c, d);
  }
  return { pos: Float32Array.from(P), idx: Uint32Array.from(I) };
}
function same(n) {
  const pos = new Float32Array(n * 9), idx = new Uint32Array(n * 3);
  for (let t = 0; t < n; t++) {
    pos.set([0, 0, 0, 1, 0, 0, 0, 1, 0], 9 * t);
    idx.set([3 * t, 3 * t + 1, 3 * t + 2], 3 * t);
  }
  return { pos, idx };
}
const cases = [
  ['empty', { pos: new Float32Array(0), idx: new Uint32Array(0) }],
  ['one', soup(11, 1, 1, 0.5)], ['two', soup(12, 2, 1, 0.5)], ['three', soup(13, 3, 1, 0.5)],
  ['five', soup(14, 5, 1, 0.3)], ['seventeen', soup(15, 17, 1, 0.3)], ['soup100', soup(16, 100, 2, 0.2)],
  ['soup1000', soup(17, 1000, 3, 0.15)], ['coincident40', same(40)], ['coincident12', same(12)],
  ['sphere', sphere(40, 25, 3)],
];
for (const [name, { pos, idx }] of cases) {
  const n = idx.length / 3;
  const b = bvhM.buildBVH(pos, idx);
  const T = bvhM.packTriangles(pos, idx, b.order);
  const r = new rngM.Rng(777 + n);
  const rays = [], closest = [], any = [], brute = [];
  for (let k = 0; k < 64; k++) {
    let o, d;
    if (k % 2 === 0) { o = [r.uniform(-6, 6), r.uniform(-6, 6), r.uniform(-6, 6)]; const tgt = [r.uniform(-1, 1), r.uniform(-1, 1), r.uniform(-1, 1)]; d = [tgt[0] - o[0], tgt[1] - o[1], tgt[2] - o[2]]; }
    else { o = [r.uniform(-0.5, 0.5), r.uniform(-0.5, 0.5), r.uniform(-0.5, 0.5)]; d = [r.normal(), r.normal(), r.normal()]; }
    const l = Math.hypot(...d); d = d.map((x) => x / l);
    const of = Array.from(new Float32Array(o)), df = Array.from(new Float32Array(d));
    rays.push(...of, ...df);
    const c = bvhM.traverse(b.nodes, T, of, df, {});
    const a = bvhM.traverse(b.nodes, T, of, df, { anyHit: true, tmax: 4 });
    const bf = bvhM.bruteForce(T, n, of, df, {});
    closest.push([c.t === Infinity ? null : c.t, c.tri, c.u, c.v]);
    any.push([a.t === Infinity ? null : a.t, a.tri, a.u, a.v]);
    brute.push([bf.t === Infinity ? null : bf.t, bf.tri, bf.u, bf.v]);
  }
  out.bvh.push({
    name, ntri: n, pos: f32hex(pos), idx: u32hex(idx),
    nodes: u32hex(new Uint32Array(b.nodes.buffer, b.nodes.byteOffset, b.nodes.length)), nodeCount: b.nodeCount,
    order: u32hex(b.order), maxDepth: b.maxDepth, bmin: b.bmin, bmax: b.bmax,
    tris: f32hex(T), sah: bvhM.sahCost(b), rays: f32hex(rays), closest, any, brute,
  });
}

// ---------------------------------------------------------------- shade_ref
{
  const r = new rngM.Rng(4242);
  const knots = [0, 0.2, 0.5, 0.9, 1.2], p = [1, 0.9, 0.6, 0.1, 0.02];
  const th = [0, 0.2, 1.2, 1.3, ...Array.from({ length: 12 }, () => r.uniform(-0.1, 1.4))];
  out.shade.profile = { knots, p, theta: th, value: th.map((t) => shM.profile(t, knots, p)) };
  const caps = [];
  for (let k = 0; k < 24; k++) {
    const P = [r.uniform(-0.1, 0.1), r.uniform(-0.1, 0.1), 0];
    let n = [r.normal(), r.normal(), Math.abs(r.normal()) + 0.2]; const nl = Math.hypot(...n); n = n.map((x) => x / nl);
    const a = [r.uniform(-0.2, 0.2), r.uniform(-0.2, 0.2), r.uniform(0.01, 0.2)];
    const bb = [r.uniform(-0.2, 0.2), r.uniform(-0.2, 0.2), r.uniform(0.01, 0.3)];
    const rad = r.uniform(0.001, 0.02);
    caps.push({ p: P, n, a, b: bb, r: rad, value: shM.capsuleOcc(P, n, a, bb, rad) });
  }
  out.shade.capsule = caps;
  const gl = shM.gaussLegendre(12);
  out.shade.gaussLegendre12 = { x: Array.from(gl[0]), w: Array.from(gl[1]) };
  const disc = [];
  for (let k = 0; k < 6; k++) {
    const x = [r.uniform(-0.3, 0.3), r.uniform(-0.3, 0.3), 0], n = [0, 0, 1];
    const c = [r.uniform(-0.2, 0.2), r.uniform(-0.2, 0.2), r.uniform(0.3, 1.0)];
    let m = [r.uniform(-0.3, 0.3), r.uniform(-0.3, 0.3), -1]; const ml = Math.hypot(...m); m = m.map((v) => v / ml);
    const a = r.uniform(0.02, 0.2), L = r.uniform(1, 10);
    disc.push({ x, n, c, m, a, L, value: shM.discIrradiance(x, n, c, m, a, L, 24, 48) });
  }
  out.shade.disc = disc;
}

fs.writeFileSync(path.join(here, 'ref_vectors.json'), JSON.stringify(out));
console.log('wrote', path.join(here, 'ref_vectors.json'), fs.statSync(path.join(here, 'ref_vectors.json')).size, 'bytes');
}
main().catch((e) => { console.error(e); process.exit(1); });
