"""The JavaScript host API (js/renderer.mjs over the N-API addon): module shape and FP64 projection on the CPU;
on the GPU, a frame rendered through TwinRenderer equals the same frame through the C ABI from Python."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from vtwin import lib, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODULE = os.path.join(ROOT, "veridien-twin_amd", "js", "renderer.mjs")
NODE = shutil.which("node")
needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(os.path.join(ROOT, "veridien-twin_amd", "napi", "vt_napi.node")),
                                reason="node or the N-API addon is not available")


def _js_sensor(sp):
    """scenes.default_sensor as sensorParams overrides (gpu/sensor.js field names) for the JavaScript renderer"""
    return dict(responseMatrix=[[float(v) for v in row] for row in np.asarray(sp["m"])], fNumber=2.0, transmission=0.9,
                esfSigmaPx=0.7, readNoise=sp["read_noise"], fullWell=sp["full_well"], conversionGain=sp["conv_gain"],
                adcBits=12, blackLevel=0.0, darkCurrent=0.0, cfa="RGGB", pupilExponent=1.0)


def _node(code):
    r = subprocess.run([NODE, "--input-type=module", "-e", code], capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    return r.stdout


@needs_node
def test_module_exports_the_renderer_api():
    out = _node("import('%s').then(m => { const p = m.TwinRenderer.prototype; console.log(JSON.stringify({exp: Object.keys(m).sort(), "
                "methods: Object.getOwnPropertyNames(p).sort(), create: typeof m.TwinRenderer.create})); })" % MODULE)
    d = json.loads(out)
    assert {"TwinRenderer", "DEFAULTS", "PRESETS", "OPTION_BITS", "lensParams", "fromPrepared", "RenderFunction",
            "rotateColumns", "quatFromMat", "STATIC_BODY"} <= set(d["exp"])
    for m in ("renderCamera", "renderFrame", "sensor", "readback", "setPoses", "project", "timings", "destroy", "idle",
              "backward", "autotune", "updateMaterials", "updateLights", "setCameraParams", "converge", "nodeSummary",
              "visible", "setLightProfile"):
        assert m in d["methods"], m
    assert d["create"] == "function"


@needs_node
def test_js_projection_matches_oracle():
    k = [0.031, -0.012, 0.004, -0.002]
    pts = [[0.1, 0.2, 1.0], [-0.4, 0.3, 0.5], [0.9, -0.7, 0.2], [0.0, 0.0, 1.0]]
    out = _node("import('%s').then(m => { const L = m.lensParams({model: 'kb4', K: [[310, 0.8, 320.5], [0, 305, 241.25], [0, 0, 1]], D: %s});"
                "console.log(JSON.stringify(%s.map(p => m.projectKB4(L, p)))); })" % (MODULE, json.dumps(k), json.dumps(pts)))
    got = json.loads(out)
    Ls = oracle.lens(310, 305, 320.5, 241.25, k, 0.8 / 310)
    for p, g in zip(pts, got):
        uv = np.zeros(2)
        oracle.lib().or_project(oracle.C.byref(Ls), oracle.ptr(np.array(p, float)), oracle.ptr(uv))
        assert np.allclose(g, uv, rtol=0, atol=1e-9)


@needs_node
def test_from_prepared_layout():
    S = ("{nodes: new Float32Array(16), tris: new Float32Array(16), verts: new Float32Array(24), "
         "meshTable: [{mat: 1, flags: 1}, {mat: 0, flags: 2}], materials: [{metallic: 0, rho: [0.5, 0.4, 0.3], alpha: 0.25, "
         "roughness: 0.5, ior: 1.5, fbar: 0.09, fbarComputed: 0.09}, {metallic: 1, f0: [0.9, 0.8, 0.7], alpha: 0.04, roughness: 0.2}], "
         "emitter: {tris: [{v0: [0, 0, 1], e1: [1, 0, 0], e2: [0, 1, 0], n: [0, 0, -1], area: 0.5, cdf: 1}], area: 0.5, L0: [3, 4, 5]}}")
    out = _node("import('%s').then(m => { const u = m.fromPrepared(%s); console.log(JSON.stringify({meshes: Array.from(u.meshes), "
                "mat: Array.from(u.materials), lights: Array.from(u.lights), emit: Array.from(u.emit), area: u.emitArea})); })" % (MODULE, S))
    d = json.loads(out)
    assert d["meshes"] == [1, 1, 0, 0, 0, 2, 0, 0]
    m = np.array(d["mat"], np.float32).reshape(2, 16)
    assert np.allclose(m[0, :8], [0.5, 0.4, 0.3, 0.25, 1.5, 0, 0.09, 0.5]) and m[1, 5] == 1 and np.allclose(m[1, :3], [0.9, 0.8, 0.7])
    assert d["lights"] == [3, 4, 5, 0] and d["area"] == 0.5
    assert np.allclose(d["emit"][:16], [0, 0, 1, 1, 1, 0, 0, 0.5, 0, 1, 0, 0, 0, 0, -1, 0])


@pytest.mark.gpu
@needs_node
def test_js_frame_equals_c_abi_frame(tmp_path):
    sc, cams = scenes.c1()
    cam = dict(cams[0])
    cam["spp"] = 4
    for k in ("nodes", "tris", "verts", "materials", "lights", "emit"):
        np.ascontiguousarray(sc[k], np.float32).tofile(tmp_path / (k + ".bin"))
    np.ascontiguousarray(sc["meshes"], np.uint32).tofile(tmp_path / "meshes.bin")
    sp = scenes.default_sensor()
    man = dict(module=MODULE, emitArea=sc["emit_area"], spp=4, maxDepth=cam["max_depth"], seed=cam["seed"],
               cameras=[dict(name="CAM1", lens=[float(x) for x in cam["lens"]], R=[float(x) for x in cam["R"]],
                             pos=[float(x) for x in cam["pos"]], size=[cam["width"], cam["height"]])],
               sensor=_js_sensor(sp), exposure=sp["exposure"])
    (tmp_path / "manifest.json").write_text(json.dumps(man))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js_render.mjs"), str(tmp_path)], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["overflows"] == 0 and info["rays"] > 0
    js_lin = np.fromfile(tmp_path / "CAM1.linear.bin", np.float32).reshape(cam["height"], cam["width"], 4)
    js_raw = np.fromfile(tmp_path / "CAM1.raw.bin", np.float32).reshape(cam["height"], cam["width"])
    ctx = lib.Context(0)
    ctx.upload_scene(sc)
    cam["cam_index"] = 0
    ctx.set_camera(0, cam)
    ctx.render([0])
    lin = ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"])
    ctx.sensor(0, dict(sp, seed=1006, frame=0))
    raw = ctx.readback(0, lib.VT_KIND_RAW, cam["width"], cam["height"])
    # sensorParams' ISP defaults: bilinear, sRGB, 4:2:0, and NR on ('none' matches the reference's /on/ test)
    ctx.isp(0, dict(nr=1))
    rgb = ctx.readback(0, lib.VT_KIND_RGB8, cam["width"], cam["height"])
    yuv = ctx.readback(0, lib.VT_KIND_YUV420, cam["width"], cam["height"])
    ctx.close()
    assert np.array_equal(js_lin, lin)
    assert np.array_equal(js_raw, raw.astype(np.float32))
    js_rgb = np.fromfile(tmp_path / "CAM1.rgb8.bin", np.uint8).reshape(cam["height"], cam["width"], 3)
    assert np.array_equal(js_rgb, rgb)
    assert np.array_equal(np.fromfile(tmp_path / "CAM1.y.bin", np.uint8).reshape(cam["height"], cam["width"]), yuv["y"])


@pytest.mark.gpu
@needs_node
def test_js_destroy_with_pending_jobs(tmp_path):
    """destroy() while readback and timings promises are queued: each settles with its data (the context goes
    when the last job has run), and a later call on the destroyed context throws instead of touching freed
    memory."""
    sc, cams = scenes.c1()
    cam = dict(cams[0])
    for k in ("nodes", "tris", "verts", "materials", "lights", "emit"):
        np.ascontiguousarray(sc[k], np.float32).tofile(tmp_path / (k + ".bin"))
    np.ascontiguousarray(sc["meshes"], np.uint32).tofile(tmp_path / "meshes.bin")
    man = dict(module=MODULE, emitArea=sc["emit_area"], spp=2, maxDepth=2,
               cameras=[dict(name="CAM1", lens=[float(x) for x in cam["lens"]], R=[float(x) for x in cam["R"]],
                             pos=[float(x) for x in cam["pos"]], size=[cam["width"], cam["height"]])])
    (tmp_path / "manifest.json").write_text(json.dumps(man))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js_destroy.mjs"), str(tmp_path)], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    print(out)
    assert out["states"] == ["fulfilled"] * 3
    assert out["lengths"] == [out["expect"], out["expect"]]
    assert out["finite"] and out["mean"] > 0
    assert "destroyed" in out["afterDestroy"]
    assert out["secondDestroy"] == "ok"


@pytest.mark.gpu
@needs_node
def test_js_calibration_lowers_the_loss(tmp_path):
    """The autograd-style RenderFunction drives a small calibration: material colours, KB4 intrinsics and the
    camera centre start perturbed; ten Adam steps on the adjoint's gradients lower the image loss and bring fx, k1
    and the camera centre back to their true values. (The camera gradient detaches the continuations' sampling, an
    unbiased estimate of the expected loss's gradient rather than the exact gradient of this 4-spp loss, so the loss
    is not driven to the common-random-number minimum; Adam's normalised steps move the colours of materials the
    camera barely sees by the full step size, so their error is printed, not bounded.)"""
    sc, cams = scenes.c1()
    cam = dict(cams[0])
    for k in ("nodes", "tris", "verts", "materials", "lights", "emit"):
        np.ascontiguousarray(sc[k], np.float32).tofile(tmp_path / (k + ".bin"))
    np.ascontiguousarray(sc["meshes"], np.uint32).tofile(tmp_path / "meshes.bin")
    man = dict(module=MODULE, emitArea=sc["emit_area"], spp=4, maxDepth=cam["max_depth"], seed=cam["seed"], iters=10,
               nMat=sc["materials"].size // 16,
               cameras=[dict(name="CAM1", lens=[float(x) for x in cam["lens"]], R=[float(x) for x in cam["R"]],
                             pos=[float(x) for x in cam["pos"]], size=[cam["width"], cam["height"]])])
    (tmp_path / "manifest.json").write_text(json.dumps(man))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js_calibrate.mjs"), str(tmp_path)], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    L = info["losses"]
    print("calibration losses:", ["%.4g" % x for x in L], info)
    assert L[-1] < 0.8 * L[0] and min(L) < 0.6 * L[0]
    assert abs(info["fx"] - info["fx_true"]) < 0.3 * abs(info["fx0"] - info["fx_true"])
    assert abs(info["k1"] - info["k1_true"]) < 0.1 * 0.01
    assert info["pos_err"] < 0.3 * info["pos_err0"]


@needs_node
def test_js_rotate_columns_is_rodrigues():
    R = scenes.look_at((0.3, -1.0, 0.5), (0.0, 0.2, 0.1))
    r = [0.01, -0.02, 0.005]
    out = _node("import('%s').then(m => console.log(JSON.stringify(m.rotateColumns(%s, %s))))" % (MODULE, json.dumps([float(x) for x in R]), json.dumps(r)))
    got = np.array(json.loads(out)).reshape(3, 3).T
    th = np.linalg.norm(r)
    k = np.array(r) / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    M = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
    assert np.allclose(got, M @ np.asarray(R, float).reshape(3, 3).T, atol=1e-7)


@needs_node
def test_js_quat_from_mat_matches_scipy():
    """quatFromMat (every branch of the largest-component choice) against scipy's Rotation.as_quat (x, y, z, w),
    up to the sign of the whole quaternion."""
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(11)
    mats = [Rotation.random(random_state=3 + k).as_matrix() for k in range(12)]
    for ax in np.eye(3):                                  # half turns: trace -1, the x, y, z branches
        mats.append(scenes.rot_axis(ax, np.pi))
    mats.append(np.eye(3))
    cols = []
    for M in mats:
        T = np.eye(4)
        T[:3, :3] = M
        T[:3, 3] = rng.uniform(-1, 1, 3)
        cols.append([float(x) for x in T.T.reshape(-1)])   # column-major
    out = _node("import('%s').then(m => console.log(JSON.stringify(%s.map((c) => m.quatFromMat(c)))))" % (MODULE, json.dumps(cols)))
    got = np.array(json.loads(out))
    for q, M in zip(got, mats):
        ref = Rotation.from_matrix(M).as_quat()
        assert min(np.abs(q - ref).max(), np.abs(q + ref).max()) < 1e-12
        assert abs(np.linalg.norm(q) - 1) < 1e-14


@needs_node
def test_cavity_of_rules():
    """cavityOf (cavityCapsule's rules, gpu/scene_gpu.js:486-503): a cannula keeps s >= 0 with the sheath radius,
    a body through the port s >= the toe, flush bodies and bodies without a port get no cut; fromPrepared attaches
    the cut of each body's port."""
    port = "{id: 'L', pivot: [0.1, 0.2, 0.3], axis_out: [0, 0, 1], cannula_toe_insertion: 0.099, cannula_outer_radius: 0.00425}"
    S = ("{tris: new Float32Array(16 * 3), verts: new Float32Array(24), meshTable: [{mat: 0, flags: 1}], "
         "materials: [{metallic: 0, rho: [0.5, 0.5, 0.5], alpha: 0.3, roughness: 0.5}], bodies: [{key: 'L/jaw', port: 'L'}, "
         "{key: 'L/cannula', port: 'L'}], blas: [{body: 'L/jaw', kind: 'instrument', triOff: 1, triCount: 1, "
         "capsule: {a: [0, 0, 0], b: [0, 0, 0.2], r: 0.003}}, {body: 'L/cannula', kind: 'fixture', triOff: 2, triCount: 1, "
         "capsule: {a: [0, 0, 0], b: [0, 0, 0.1], r: 0.005}}], emitter: {tris: [], area: 0, L0: [1, 1, 1]}}")
    out = _node("import('%s').then(m => { const P = %s; const u = m.fromPrepared(%s, [P]); console.log(JSON.stringify({"
                "fix: m.cavityOf('fixture', P), ins: m.cavityOf('instrument', P), sc: m.cavityOf('scope', P), "
                "lens: m.cavityOf('fixed', P), board: m.cavityOf('board', P), none: m.cavityOf('instrument', null), "
                "bodies: u.bodies.map((b) => b.cavity || null)})); })" % (MODULE, port, S))
    d = json.loads(out)
    assert d["fix"] == {"origin": [0.1, 0.2, 0.3], "axis": [0, 0, -1], "s0": 0, "r": 0.00425}
    assert d["ins"] == {"origin": [0.1, 0.2, 0.3], "axis": [0, 0, -1], "s0": 0.099, "r": 0} == d["sc"]
    assert d["lens"] is None and d["board"] is None and d["none"] is None
    assert d["bodies"] == [d["ins"], d["fix"]]


@needs_node
def test_from_prepared_bodies():
    """fromPrepared with the reference's BLAS list: triangle ranges become triBody, the caller BVH is dropped (every
    group's BVH is built at upload), the bodies keep their keys and kinds, and the light rides on the scope."""
    S = ("{nodes: new Float32Array(16), tris: new Float32Array(16 * 5), verts: new Float32Array(24), "
         "meshTable: [{mat: 0, flags: 1}], materials: [{metallic: 0, rho: [0.5, 0.4, 0.3], alpha: 0.25, roughness: 0.5}], "
         "blas: [{body: 'arm', kind: 'body', triOff: 1, triCount: 2, owner: 'CAM2', rhoEff: 0.3, occCapsule: {a: [0, 0, 0], "
         "b: [0, 0, 0.1], r: 0.01}}, {body: 'scope', kind: 'fixed', triOff: 4, triCount: 1}], emitter: {tris: [{v0: [0, 0, 0.1], e1: [0.01, 0, 0], "
         "e2: [0, 0.01, 0], n: [0, 0, 1], cdf: 1, area: 5e-5}], area: 5e-5, L0: [1, 1, 1], shadowSkipsScope: true, axis: [0, 0, -1], "
         "theta: [0, 0.5], p: [1, 0]}}")
    out = _node("import('%s').then(m => { const u = m.fromPrepared(%s); console.log(JSON.stringify({tb: Array.from(u.triBody), "
                "nodes: u.nodes === undefined, bodies: u.bodies, shadowSkip: u.shadowSkip, profiles: u.profiles, "
                "emitBody: u.emitBody})); })" % (MODULE, S))
    d = json.loads(out)
    assert d["tb"] == [0xFFFFFFFF, 0, 0, 0xFFFFFFFF, 1] and d["nodes"]
    assert d["bodies"] == [{"key": "arm", "kind": "body", "owner": "CAM2", "rhoEff": 0.3,
                            "occ": {"a": [0, 0, 0], "b": [0, 0, 0.1], "r": 0.01}},
                           {"key": "scope", "kind": "fixed", "owner": None, "rhoEff": 0, "occ": None}]
    assert d["shadowSkip"] == 2          # the scope's bit: shadow rays leave through its housing
    assert d["emitBody"] == 1            # the outlet's records are in the scope's frame: the light rides on it
    assert d["profiles"] == [{"light": 0, "axis": [0, 0, -1], "theta": [0, 0.5], "p": [1, 0]}]


def test_from_prepared_texture():
    """A printed material's ink goes to words 8-11 (rho_d, roughInk^2: gpu/renderer.js:293) and the textured
    meshes' vertices give their texture coordinates from the spare words (u in pos.w, v in normal.w)."""
    S = ("{nodes: new Float32Array(16), tris: T, verts: V, meshTable: [{mat: 0, flags: 1}, {mat: 1, flags: 136}], "
         "materials: [{metallic: 0, rho: [0.5, 0.5, 0.5], alpha: 0.25, roughness: 0.5}, {metallic: 0, rho: [0.86, 0.86, 0.85], "
         "alpha: 0.36, roughness: 0.6, texture: {ink: [0.045, 0.04, 0.05], roughInk: 0.4}}]}")
    setup = ("const T = new Float32Array(32), V = new Float32Array(48), tu = new Uint32Array(T.buffer); "
             "tu[3] = 0; tu[7] = 1; tu[11] = 2; tu[12] = 0; tu[19] = 3; tu[23] = 4; tu[27] = 5; tu[28] = 1; "
             "for (let v = 0; v < 6; v++) { V[8 * v + 3] = 0.1 * v; V[8 * v + 7] = 1 - 0.1 * v; } ")
    out = _node("%simport('%s').then(m => { const u = m.fromPrepared(%s); console.log(JSON.stringify({uvs: Array.from(u.uvs), "
                "mat: Array.from(u.materials)})); })" % (setup, MODULE, S))
    d = json.loads(out)
    uv = np.array(d["uvs"], np.float32).reshape(6, 2)
    assert np.allclose(uv[:3], 0) and np.allclose(uv[3:], [[0.3, 0.7], [0.4, 0.6], [0.5, 0.5]])
    m = np.array(d["mat"], np.float32).reshape(2, 16)
    assert np.allclose(m[1, :4], [0.86, 0.86, 0.85, 0.36]) and np.allclose(m[1, 8:12], [0.045, 0.04, 0.05, 0.16])
    assert np.allclose(m[0, 8:12], [1, 1, 1, 0])


@pytest.mark.gpu
@needs_node
def test_js_textured_frame_equals_c_abi_frame(tmp_path):
    """The printed boards through TwinRenderer (uvs in the scene, boardTexture at create) against the same frame
    through the C ABI."""
    sc = scenes.textured_scene()
    cam = scenes.posed_camera(96, 72, spp=4)
    for k in ("tris", "verts", "materials", "lights", "emit", "uvs"):
        np.ascontiguousarray(sc[k], np.float32).tofile(tmp_path / (k + ".bin"))
    np.ascontiguousarray(sc["meshes"], np.uint32).tofile(tmp_path / "meshes.bin")
    np.ascontiguousarray(sc["tri_body"], np.uint32).tofile(tmp_path / "triBody.bin")
    np.ascontiguousarray(sc["board_mask"], np.uint8).tofile(tmp_path / "mask.bin")
    poses = [dict(q0=(0.0, 0.0, 0.0, 1.0), p0=(-0.55, -1.45, 1.0)), dict(q0=(0.0, 0.0, 0.0, 1.0), p0=(0.9, 0.6, 0.6))]

    def T(t):
        M = np.eye(4)
        M[:3, 3] = t
        return [float(x) for x in M.T.reshape(-1)]
    sp = scenes.default_sensor()
    man = dict(module=MODULE, emitArea=sc["emit_area"], spp=4, maxDepth=cam["max_depth"], seed=cam["seed"],
               bodies=["box", "can"], poses0={"bodies": {"box": T(poses[0]["p0"]), "can": T(poses[1]["p0"])}},
               uvs=True, boardTexture=[sc["board_mask"].shape[1], sc["board_mask"].shape[0]],
               cameras=[dict(name="CAM1", lens=[float(v) for v in cam["lens"]], R=[float(v) for v in cam["R"]],
                             pos=[float(v) for v in cam["pos"]], size=[cam["width"], cam["height"]])],
               sensor=_js_sensor(sp), exposure=sp["exposure"])
    (tmp_path / "manifest.json").write_text(json.dumps(man))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js_render.mjs"), str(tmp_path)], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    js_lin = np.fromfile(tmp_path / "CAM1.linear.bin", np.float32).reshape(cam["height"], cam["width"], 4)
    ctx = lib.Context(0)
    ctx.upload_scene(sc)
    ctx.set_instances(poses)
    cam["cam_index"] = 0
    ctx.set_camera(0, cam)
    ctx.render([0])
    lin = ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"])
    ctx.close()
    a, b = js_lin[:, :, :3].astype(np.float64), lin[:, :, :3].astype(np.float64)
    assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)


@pytest.mark.gpu
@needs_node
def test_js_posed_frame_equals_c_abi_frame(tmp_path):
    """setPoses with 4x4 body poses (rotations from Rodrigues' formula, one body moving over the exposure) through
    TwinRenderer, against the same frame through the C ABI with the poses as quaternions."""
    parts = scenes.posed_parts()
    sc = scenes.posed_scene(*parts)
    cam = scenes.posed_camera(96, 72, spp=4)
    cam["skip_bodies"] = 1        # the camera owns body 0 ("box"): its own rays ignore it
    rng = np.random.default_rng(3)
    pts = np.column_stack([rng.uniform(-1.6, 1.6, 200), rng.uniform(-2.2, 1.2, 200), rng.uniform(0.0, 1.2, 200)])
    for k in ("tris", "verts", "materials", "lights", "emit"):
        np.ascontiguousarray(sc[k], np.float32).tofile(tmp_path / (k + ".bin"))
    np.ascontiguousarray(sc["meshes"], np.uint32).tofile(tmp_path / "meshes.bin")
    np.ascontiguousarray(sc["tri_body"], np.uint32).tofile(tmp_path / "triBody.bin")

    def T(R, t):
        M = np.eye(4)
        M[:3, :3], M[:3, 3] = R, t
        return [float(x) for x in M.T.reshape(-1)]
    z, x = (0.0, 0.0, 1.0), (1.0, 0.0, 0.0)
    poses0 = {"bodies": {"box": T(scenes.rot_axis(z, 0.35), (0.75, -1.45, 0.3)),
                         "can": T(scenes.rot_axis(x, 0.4), (-0.75, -1.35, 0.45))}}
    poses1 = {"bodies": {"box": T(scenes.rot_axis(z, 1.25), (0.95, -1.2, 0.35)),
                         "can": T(scenes.rot_axis(x, 0.4), (-0.75, -1.35, 0.45))}}
    sp = scenes.default_sensor()
    man = dict(module=MODULE, emitArea=sc["emit_area"], spp=4, maxDepth=cam["max_depth"], seed=cam["seed"],
               bodies=[{"key": "box", "kind": "body", "owner": "CAM1"}, "can"], poses0=poses0, poses1=poses1,
               shutter=[0.0, 1.0], visiblePoints=[[float(v) for v in p] for p in pts],
               cameras=[dict(name="CAM1", lens=[float(v) for v in cam["lens"]], R=[float(v) for v in cam["R"]],
                             pos=[float(v) for v in cam["pos"]], size=[cam["width"], cam["height"]])],
               sensor=_js_sensor(sp), exposure=sp["exposure"])
    (tmp_path / "manifest.json").write_text(json.dumps(man))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js_render.mjs"), str(tmp_path)], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    js_lin = np.fromfile(tmp_path / "CAM1.linear.bin", np.float32).reshape(cam["height"], cam["width"], 4)
    ctx = lib.Context(0)
    ctx.upload_scene(sc)
    ctx.set_instances([dict(q0=scenes.quat_axis(z, 0.35), p0=(0.75, -1.45, 0.3), q1=scenes.quat_axis(z, 1.25),
                            p1=(0.95, -1.2, 0.35)), dict(q0=scenes.quat_axis(x, 0.4), p0=(-0.75, -1.35, 0.45))])
    cam["cam_index"] = 0
    ctx.set_camera(0, cam)
    ctx.render([0])
    lin = ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"])
    vis = ctx.visible(0, pts, tol=2e-4, time=0.5)
    ctx.close()
    a, b = js_lin[:, :, :3].astype(np.float64), lin[:, :, :3].astype(np.float64)
    assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)
    js_vis = np.array(json.loads((tmp_path / "visible.json").read_text()))
    assert np.array_equal(js_vis, vis) and 0 < vis.sum() < vis.size


@pytest.mark.gpu
@needs_node
def test_js_reference_estimator_frame_equals_c_abi(tmp_path):
    """estimator 'reference' through TwinRenderer (node cache built on first use, direct light + 96 Jacobi steps,
    shadow culling on, the refinement pass of 6 sub-samples x 1 light sample with shadow pixels included, the fixed
    camera's first-hit and occlusion caches, as the reference's DEFAULTS) against the same frame through the C ABI."""
    sc, cams = scenes.c1()
    cam = dict(cams[0], spp=4)
    for k in ("nodes", "tris", "verts", "materials", "lights", "emit"):
        np.ascontiguousarray(sc[k], np.float32).tofile(tmp_path / (k + ".bin"))
    np.ascontiguousarray(sc["meshes"], np.uint32).tofile(tmp_path / "meshes.bin")
    sp = scenes.default_sensor()
    man = dict(module=MODULE, emitArea=sc["emit_area"], spp=4, maxDepth=1, seed=cam["seed"], estimator="reference",
               cameras=[dict(name="CAM1", lens=[float(v) for v in cam["lens"]], R=[float(v) for v in cam["R"]],
                             pos=[float(v) for v in cam["pos"]], size=[cam["width"], cam["height"]])],
               sensor=_js_sensor(sp), exposure=sp["exposure"])
    (tmp_path / "manifest.json").write_text(json.dumps(man))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js_render.mjs"), str(tmp_path)], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    js_lin = np.fromfile(tmp_path / "CAM1.linear.bin", np.float32).reshape(cam["height"], cam["width"], 4)
    summ = json.loads((tmp_path / "summary.json").read_text())
    ctx = lib.Context(0)
    ctx.upload_scene(sc)
    info = ctx.radiosity_build()
    ctx.interior_bvh(True)
    ctx.radiosity_update(True, 96)
    cam.update(cam_index=0, max_depth=1, refine=6, refine_light_samples=1,
               flags=lib.VT_OPT_REFERENCE | lib.VT_OPT_CULL_SHADOWS | lib.VT_OPT_REFINE_SHADOWS | lib.VT_OPT_HIT_CACHE
               | lib.VT_OPT_SHADOW_CACHE)
    ctx.set_camera(0, cam)
    ctx.render([0])
    lin = ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"])
    _, e = ctx.radiosity_state()
    ctx.close()
    assert summ["active"] == info["active"] and np.allclose(summ["eAvg"], e, rtol=1e-6)
    assert np.array_equal(js_lin, lin)
