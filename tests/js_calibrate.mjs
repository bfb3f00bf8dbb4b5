// Calibration through the JavaScript autograd-style RenderFunction: a target frame from the true parameters,
// then gradient steps (Adam) on perturbed material colours, KB4 intrinsics and pose, using the adjoint.
// Usage: node tests/js_calibrate.mjs <dir> (scene written by tests/test_js_api.py); prints the loss per step.
import fs from 'fs';
import path from 'path';

const dir = process.argv[2];
const man = JSON.parse(fs.readFileSync(path.join(dir, 'manifest.json'), 'utf8'));
const f32 = (name) => { const b = fs.readFileSync(path.join(dir, name + '.bin')); return new Float32Array(b.buffer, b.byteOffset, b.byteLength / 4); };
const u32 = (name) => { const b = fs.readFileSync(path.join(dir, name + '.bin')); return new Uint32Array(b.buffer, b.byteOffset, b.byteLength / 4); };

async function main() {
  const m = await import(path.resolve(man.module));
  const scene = { nodes: f32('nodes'), tris: f32('tris'), verts: f32('verts'), meshes: u32('meshes'), materials: f32('materials'),
    lights: f32('lights'), emit: f32('emit'), emitArea: man.emitArea };
  const r = await m.TwinRenderer.create({ scene, cameras: man.cameras, spp: man.spp, maxDepth: man.maxDepth, seed: man.seed });
  const name = man.cameras[0].name;
  const fn = new m.RenderFunction(r, [name]);
  const target = (await fn.forward())[name];
  // perturb: material colours of the first nm materials, fx, k1, centre
  const mats = Float32Array.from(scene.materials);
  const nm = man.nMat;
  for (let i = 0; i < nm; i++) for (let c = 0; c < 3; c++) mats[16 * i + c] *= 1 + 0.2 * Math.sin(7 * i + c);
  const cam0 = r.cameras[name];
  const fxTrue = cam0.lens[0];
  const lens = Array.from(cam0.lens), pos = Array.from(cam0.pos);
  lens[0] *= 1.01; lens[1] *= 1.01; lens[4] += 0.01; pos[0] += 0.003;
  let R = Array.from(cam0.R);
  const p = { mats, lens, pos };
  // Adam over [colours (3 nm), fx fy cx cy k1..k4 (8), r (3), t (3)] with per-group step sizes
  const lr = [];
  for (let i = 0; i < 3 * nm; i++) lr.push(0.01);
  lr.push(0.5, 0.5, 0.25, 0.25, 0.002, 0.002, 0.002, 0.002, 0.0005, 0.0005, 0.0005, 0.0005, 0.0005, 0.0005);
  const mA = new Float64Array(lr.length), vA = new Float64Array(lr.length);
  const losses = [];
  for (let it = 1; it <= man.iters; it++) {
    const img = (await fn.forward({ materials: p.mats, cameras: { [name]: { lens: p.lens, R, pos: p.pos } } }))[name];
    const grad = new Float32Array(img.length);
    let loss = 0;
    for (let k = 0; k < img.length; k++) {
      if ((k & 3) === 3) continue;
      const d = img[k] - target[k];
      grad[k] = d; loss += 0.5 * d * d;
    }
    losses.push(loss);
    const g = await fn.backward({ [name]: grad });
    const gc = g.cameras[name];
    const gv = [];
    for (let i = 0; i < nm; i++) for (let c = 0; c < 3; c++) gv.push(g.materials[3 * i + c]);
    for (let k = 0; k < 14; k++) gv.push(gc[k]);
    const step = gv.map((x, k) => {
      mA[k] = 0.9 * mA[k] + 0.1 * x; vA[k] = 0.999 * vA[k] + 0.001 * x * x;
      const mh = mA[k] / (1 - 0.9 ** it), vh = vA[k] / (1 - 0.999 ** it);
      return lr[k] * mh / (Math.sqrt(vh) + 1e-12);
    });
    for (let i = 0; i < nm; i++) for (let c = 0; c < 3; c++) p.mats[16 * i + c] -= step[3 * i + c];
    for (let k = 0; k < 8; k++) p.lens[k] -= step[3 * nm + k];
    R = m.rotateColumns(R, [-step[3 * nm + 8], -step[3 * nm + 9], -step[3 * nm + 10]]);
    for (let k = 0; k < 3; k++) p.pos[k] -= step[3 * nm + 11 + k];
  }
  // parameter errors at the start and the end: material colours (mean |relative error|), fx, k1, centre (m)
  const matErr = (mm) => {
    let e = 0, n = 0;
    for (let i = 0; i < nm; i++) for (let c = 0; c < 3; c++) {
      const t = scene.materials[16 * i + c];
      if (t > 0) { e += Math.abs(mm[16 * i + c] / t - 1); n++; }
    }
    return e / Math.max(n, 1);
  };
  const mats0 = Float32Array.from(scene.materials);
  for (let i = 0; i < nm; i++) for (let c = 0; c < 3; c++) mats0[16 * i + c] *= 1 + 0.2 * Math.sin(7 * i + c);
  const posErr = (pp) => Math.hypot(pp[0] - cam0.pos[0], pp[1] - cam0.pos[1], pp[2] - cam0.pos[2]);
  console.log(JSON.stringify({ losses, fx: p.lens[0], fx_true: fxTrue, fx0: fxTrue * 1.01, k1: p.lens[4],
    k1_true: cam0.lens[4], mat_err0: matErr(mats0), mat_err: matErr(p.mats), pos_err0: 0.003, pos_err: posErr(p.pos) }));
  r.destroy();
}
main().catch((e) => { console.error(e); process.exit(1); });
