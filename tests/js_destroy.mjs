// Test driver: destroy() while readback / timings / backward promises are pending. Each pending promise settles
// (resolves with its data), and a call on the destroyed context throws. Usage: node tests/js_destroy.mjs <dir>
import fs from 'fs';
import path from 'path';
import { createRequire } from 'module';

const dir = process.argv[2];
const man = JSON.parse(fs.readFileSync(path.join(dir, 'manifest.json'), 'utf8'));
const f32 = (name) => { const b = fs.readFileSync(path.join(dir, name + '.bin')); return new Float32Array(b.buffer, b.byteOffset, b.byteLength / 4); };
const u32 = (name) => { const b = fs.readFileSync(path.join(dir, name + '.bin')); return new Uint32Array(b.buffer, b.byteOffset, b.byteLength / 4); };

async function main() {
  const { TwinRenderer } = await import(path.resolve(man.module));
  const native = createRequire(import.meta.url)(path.resolve(path.dirname(man.module), '../napi/vt_napi.node'));
  const scene = { nodes: f32('nodes'), tris: f32('tris'), verts: f32('verts'), meshes: u32('meshes'),
    materials: f32('materials'), lights: f32('lights'), emit: f32('emit'), emitArea: man.emitArea };
  const r = await TwinRenderer.create({ scene, cameras: man.cameras, spp: man.spp, maxDepth: man.maxDepth, seed: 3 });
  const name = man.cameras[0].name;
  const [w, h] = man.cameras[0].size;
  r.renderCamera(name, {});
  const ctx = r.ctx;
  const pending = [r.readback(name, 'linear'), r.timings(), r.readback(name, 'linear')];
  r.destroy();                       // with three jobs queued
  const settled = await Promise.allSettled(pending);
  const out = { states: settled.map((s) => s.status), lengths: [] };
  for (const s of settled) if (s.status === 'fulfilled' && s.value && s.value.length) out.lengths.push(s.value.length);
  const lin = settled[0].status === 'fulfilled' ? settled[0].value : null;
  out.finite = lin ? lin.every(Number.isFinite) : false;
  out.mean = lin ? lin.reduce((a, b) => a + b, 0) / lin.length : 0;
  try { native.sync(ctx); out.afterDestroy = 'returned'; } catch (e) { out.afterDestroy = String(e.message); }
  try { native.destroy(ctx); out.secondDestroy = 'ok'; } catch (e) { out.secondDestroy = String(e.message); }
  out.expect = w * h * 4;
  console.log(JSON.stringify(out));
}
main().catch((e) => { console.error(e); process.exit(1); });
