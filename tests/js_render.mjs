// Test driver for the JavaScript TwinRenderer: renders a scene written by tests/test_js_api.py and writes the
// linear readback (and the raw DN) next to it. Usage: node tests/js_render.mjs <dir>
import fs from 'fs';
import path from 'path';

const dir = process.argv[2];
const man = JSON.parse(fs.readFileSync(path.join(dir, 'manifest.json'), 'utf8'));
const f32 = (name) => { const b = fs.readFileSync(path.join(dir, name + '.bin')); return new Float32Array(b.buffer, b.byteOffset, b.byteLength / 4); };
const u32 = (name) => { const b = fs.readFileSync(path.join(dir, name + '.bin')); return new Uint32Array(b.buffer, b.byteOffset, b.byteLength / 4); };

async function main() {
  const { TwinRenderer } = await import(path.resolve(man.module));
  const scene = { tris: f32('tris'), verts: f32('verts'), meshes: u32('meshes'), materials: f32('materials'),
    lights: f32('lights'), emit: f32('emit'), emitArea: man.emitArea };
  if (man.bodies) Object.assign(scene, { triBody: u32('triBody'), bodies: man.bodies });
  else scene.nodes = f32('nodes');
  if (man.uvs) scene.uvs = f32('uvs');
  let boardTexture;
  if (man.boardTexture) {
    const b = fs.readFileSync(path.join(dir, 'mask.bin'));
    boardTexture = { mask: new Uint8Array(b.buffer, b.byteOffset, b.byteLength), width: man.boardTexture[0], height: man.boardTexture[1] };
  }
  const r = await TwinRenderer.create({ scene, cameras: man.cameras, spp: man.spp, maxDepth: man.maxDepth, seed: man.seed,
    estimator: man.estimator || 'path', boardTexture });
  if (man.poses0) r.setPoses(man.poses0, man.poses1 || null);
  const out = r.renderFrame({ sensor: true, sensorOpts: { params: man.sensor, seed: 1006, frame: 0, exposure_s: man.exposure, gain: 1 }, shutter: man.shutter });
  await out.done;
  for (const c of man.cameras) {
    const lin = await r.readback(c.name, 'linear');
    fs.writeFileSync(path.join(dir, c.name + '.linear.bin'), Buffer.from(lin.buffer, lin.byteOffset, lin.byteLength));
    const raw = await r.readback(c.name, 'raw');
    fs.writeFileSync(path.join(dir, c.name + '.raw.bin'), Buffer.from(raw.buffer, raw.byteOffset, raw.byteLength));
    const rgb = await r.readback(c.name, 'rgb8');
    fs.writeFileSync(path.join(dir, c.name + '.rgb8.bin'), Buffer.from(rgb.buffer, rgb.byteOffset, rgb.byteLength));
    const yuv = await r.readback(c.name, 'yuv420');
    fs.writeFileSync(path.join(dir, c.name + '.y.bin'), Buffer.from(yuv.y.buffer, yuv.y.byteOffset, yuv.y.byteLength));
  }
  if (man.visiblePoints) {
    const vis = r.visible(man.cameras[0].name, man.visiblePoints, { tol: 2e-4, time: 0.5 });
    fs.writeFileSync(path.join(dir, 'visible.json'), JSON.stringify(vis));
  }
  if (man.estimator === 'reference') {
    const s = await r.nodeSummary();
    fs.writeFileSync(path.join(dir, 'summary.json'), JSON.stringify({ eAvg: s.eAvg, active: s.activeNodes }));
  }
  const t = await r.timings();
  console.log(JSON.stringify({ rays: t.last.rays, overflows: t.last.stackOverflows }));
  r.destroy();
}
main().catch((e) => { console.error(e); process.exit(1); });
