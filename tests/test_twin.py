"""The real twin scene through the camera path (vtwin/twin.py; the reference's gpu/scene_gpu.js prepareScene and
control/kinematics.js posesFromState).

Inputs: twin_assets/ (the reference's assets/scene.json, its 44 meshes and configs/twin/physics.json, copied by
tests/golden/fetch_twin_assets.py; tests that need it skip without it). Golden vectors from the reference's own
scene_gpu.js and kinematics.js cannot be made here: they use the ?? and ?. operators, which Node 12 (the only
JavaScript runtime in the image) does not parse. They are pinned instead by:
  * weld + subdivide (vt_weld_subdivide): closed-form counts on a hand-made triangle; on the real meshes, no edge
    longer than h, the area unchanged (1e-12), every new vertex the exact midpoint of two older ones, no T-junction
    (boundary length unchanged), no coincident vertices left after welding;
  * the scene's own numbers: raw static triangles + lens fronts = scene.json totals.static_triangles, the dynamic
    total, the light outlet's area = scene.scope.emitter_area_m2;
  * kinematics: the scope's lens point and view direction at the default state = scene.scope.default origin_D and
    view_dir_D (exported by the reference's Python twin); orthonormal poses, jaws opened by the jaw angle;
  * the bench: its triangles tile the extent exactly, fine cells no longer than h.
GPU: the twin's three cameras (1280 x 720, KB4) rendered with the bodies at the default state against the oracle
on the same samples.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from vtwin import lib, scenes, twin

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASSETS = os.path.join(ROOT, "twin_assets")
needs_assets = pytest.mark.skipif(not os.path.exists(os.path.join(ASSETS, "scene.json")),
                                  reason="twin_assets/ missing (python tests/golden/fetch_twin_assets.py)")


@pytest.fixture(scope="module")
def loaded():
    return twin.load(ASSETS)


@pytest.mark.skipif(not os.path.exists(os.path.join(ASSETS, "lite", "lite_nodes.bin")),
                    reason="twin_assets/lite missing (python tests/golden/fetch_twin_assets.py)")
def test_baked_node_table_loads():
    """The lite renderer's baked node table read as a gather table: its miss fraction equals the one its generator
    recorded (lite.json stats.gatherMissFrac), every hit names a node of the table, normals are unit vectors and
    areas positive."""
    L = twin.lite_nodes(ASSETS)
    n, K, g = L["count"], L["K"], L["gtable"]
    assert g.size == n * K and K == 256
    miss = g == 0xFFFFFFFF
    assert miss.mean() == pytest.approx(L["stats"]["gatherMissFrac"], abs=1e-15)
    assert g[~miss].max() < n
    assert np.allclose(np.linalg.norm(L["nrm"], axis=1), 1.0, atol=1e-6) and (L["area"] > 0).all()


@pytest.fixture(scope="module")
def prep(loaded):
    scene, meshes, physics = loaded
    return twin.prepare(scene, meshes, physics)


def _edges(I):
    e = np.concatenate([I[:, [0, 1]], I[:, [1, 2]], I[:, [2, 0]]])
    return np.sort(e, axis=1)


def _area(P, I):
    P = np.asarray(P, np.float64)
    return 0.5 * np.linalg.norm(np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]]), axis=1)


def test_subdivide_closed_forms():
    """An equilateral triangle of side s splits into 4^k equal ones, k = ceil(log2(s / h)), on the triangular
    lattice's (2^k + 1)(2^k + 2)/2 nodes; a right triangle with legs under h and a longer hypotenuse splits once into
    two; two triangles with their own copies of a quad's shared corners weld to four vertices."""
    h = 1.0
    for s, k in [(0.9, 0), (1.5, 1), (3.0, 2), (7.9, 3)]:
        P = np.array([[0, 0, 0], [s, 0, 0], [s / 2, s * math.sqrt(3) / 2, 0]], np.float32)
        N = np.tile(np.array([0, 0, 1], np.float32), (3, 1))
        P2, N2, I2 = lib.weld_subdivide(P, N, np.array([[0, 1, 2]], np.uint32), h)
        assert I2.shape[0] == 4 ** k
        assert P2.shape[0] == (2 ** k + 1) * (2 ** k + 2) // 2          # the triangular lattice's nodes
        assert np.allclose(N2, [0, 0, 1])
    P = np.array([[0, 0, 0], [0.9, 0, 0], [0, 0.9, 0]], np.float32)
    P2, _, I2 = lib.weld_subdivide(P, np.tile(np.float32([0, 0, 1]), (3, 1)), np.array([[0, 1, 2]], np.uint32), 1.0)
    assert I2.shape[0] == 2 and np.allclose(P2[3], [0.45, 0.45, 0])
    # welding: a quad given as two triangles with their own copies of the shared corners
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    N = np.tile(np.float32([0, 0, 1]), (6, 1))
    P2, _, I2 = lib.weld_subdivide(P, N, np.arange(6, dtype=np.uint32).reshape(2, 3), 10.0)
    assert P2.shape[0] == 4 and I2.tolist() == [[0, 1, 2], [0, 2, 3]]


@needs_assets
@pytest.mark.parametrize("name", ["membrane_L", "lid_spine_1_crest", "static_shell_left", "hatch_apex"])
def test_weld_subdivide_invariants_on_twin_meshes(loaded, name):
    scene, meshes, _ = loaded
    P, N, I, _ = meshes[name]
    h = 0.003
    P2, N2, I2 = lib.weld_subdivide(P, N, I, h)
    e = _edges(I2)
    L = np.linalg.norm(P2[e[:, 0]] - P2[e[:, 1]], axis=1)
    assert L.max() <= h
    assert _area(P2, I2).sum() == pytest.approx(_area(P, I).sum(), rel=1e-12)
    # welded vertices first (unique keys), then midpoints of two earlier vertices, exactly
    key = np.concatenate([np.floor(P2 * 1e7 + 0.5), np.floor(N2 * 1e4 + 0.5)], axis=1)
    Pw, _, Iw = lib.weld_subdivide(P, N, I, 1e9)
    nw = Pw.shape[0]
    assert np.unique(key[:nw], axis=0).shape[0] == nw
    assert np.array_equal(P2[:nw], Pw)
    ue, cnt = np.unique(e, axis=0, return_counts=True)
    assert cnt.max() <= 2
    # no T-junctions: the boundary (edges of one triangle) keeps its length
    ew = _edges(Iw)
    uw, cw = np.unique(ew, axis=0, return_counts=True)
    bl = lambda Pp, ee: float(np.linalg.norm(Pp[ee[:, 0]] - Pp[ee[:, 1]], axis=1).sum())
    assert bl(P2, ue[cnt == 1]) == pytest.approx(bl(Pw, uw[cw == 1]), rel=1e-12)
    # midpoints: each new vertex is the mean of the two ends of an edge that contains it
    rng = np.random.default_rng(3)
    for v in rng.choice(np.arange(nw, P2.shape[0]), size=min(200, P2.shape[0] - nw), replace=False):
        tri = I2[np.any(I2 == v, axis=1)]
        nb = np.unique(tri[tri != v])
        d = np.linalg.norm(P2[nb][:, None, :] + P2[nb][None, :, :] - 2 * P2[v], axis=2)
        assert d.min() == 0.0


@needs_assets
def test_twin_counts_and_outlet(loaded, prep):
    scene, meshes, _ = loaded
    lens_fronts = sum(m["triangle_count"] for m in scene["meshes"] if m.get("owner_camera"))
    assert prep["stats"]["raw_static_tris"] + lens_fronts == scene["totals"]["static_triangles"]
    tb = np.asarray(prep["tri_body"])
    body_tris = int(np.sum(tb != twin.STATIC_BODY))
    outlet = meshes[scene["scope"]["emitter_mesh"]][2].shape[0]
    assert body_tris - lens_fronts == scene["totals"]["dynamic_triangles"]
    assert prep["emit_area"] == pytest.approx(scene["scope"]["emitter_area_m2"], rel=1e-6)
    assert prep["emit"].size // 16 == outlet
    # the outlet rides on the scope: its triangles are the scope body's
    scope = [b["key"] for b in prep["bodies"]].index("scope")
    assert prep["emit_body"] == scope
    em = (np.asarray(prep["meshes"]).reshape(-1, 4)[np.asarray(prep["tris"]).reshape(-1, 16).view(np.uint32)[:, 12], 1]
          & twin.MESH_FLAG["EMITTER"]) != 0
    assert em.sum() == outlet and np.all(tb[em] == scope)
    assert len(prep["bodies"]) == 14 and prep["body_count"] == 14
    kinds = sorted(b["kind"] for b in prep["bodies"])
    assert kinds.count("instrument") == 6 and kinds.count("fixed") == 3 and kinds.count("fixture") == 3
    # every material the meshes name is in physics (no grey fallback)
    assert not any(p.get("missing") for p in prep["materials_prepared"] if p["name"] != "emitter")


@needs_assets
def test_twin_cavity_cuts(loaded, prep):
    """cavityCapsule's cuts (gpu/scene_gpu.js:486-503) from the scene's ports: each cannula keeps s >= 0 with its
    sheath radius, each instrument part and the scope s >= its cannula's toe, lens fronts and boards carry no
    capsule; the cut plane is the port's pivot and inward axis."""
    scene, _, _ = loaded
    ports = {p["id"]: p for p in scene["ports"]}
    inst = twin.instances(prep, twin.poses_from_state(scene, twin.default_state(scene)))
    kinds = {}
    for body, rec in zip(prep["bodies"], inst):
        if rec is None:
            continue
        cav = rec.get("cavity")
        kinds.setdefault(body["kind"], []).append(cav is not None)
        if body["kind"] in ("fixed", "board"):
            assert cav is None and rec["occ_r"] == 0.0
            continue
        port = ports.get(body.get("port"))
        if port is None:
            assert cav is None
            continue
        assert np.allclose(cav["origin"], port["pivot"]) and np.allclose(cav["axis"], -np.asarray(port["axis_out"]))
        if body["kind"] == "fixture":
            assert cav["s0"] == 0.0 and cav["r"] == pytest.approx(port["cannula_outer_radius"])
        else:
            assert cav["s0"] == pytest.approx(port["cannula_toe_insertion"]) and cav["r"] == 0.0
    print("bodies with a cavity cut, by kind: %s" % {k: sum(v) for k, v in kinds.items()})
    assert sum(kinds.get("instrument", [])) >= 4 and sum(kinds.get("fixture", [])) == 3 and all(kinds.get("scope", [False]))


@needs_assets
def test_kinematics_against_the_scene_export(loaded):
    scene, _, _ = loaded
    state = twin.default_state(scene)
    poses = twin.poses_from_state(scene, state)
    sd = scene["scope"]["default"]
    T = poses["scope"]
    assert np.allclose(T[:3, 3], sd["origin_D"], atol=1e-12)
    dov = scene["scope"]["direction_of_view_deg"] * math.pi / 180
    assert np.allclose(T[:3, :3] @ [0, math.sin(dov), math.cos(dov)], sd["view_dir_D"], atol=1e-12)
    for key, T in poses.items():
        if T is None:
            assert key.startswith("board/")
            continue
        R = T[:3, :3]
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12) and np.linalg.det(R) == pytest.approx(1.0)
    for q in state["instruments"]:
        a, b, s = poses[q["id"] + "/jaw_a"], poses[q["id"] + "/jaw_b"], poses[q["id"] + "/shaft"]
        ang = math.acos(np.clip(a[:3, 2] @ b[:3, 2], -1, 1))
        assert ang == pytest.approx(q["jaw"], abs=1e-12)
        pv = np.array(next(p for p in scene["ports"] if p["id"] == q["id"])["pivot"])
        assert np.allclose(s[:3, 3], pv + q["insertion"] * s[:3, 2], atol=1e-12)


def test_bench_tiles_its_extent():
    ext = [[-0.3, -0.2], [0.25, 0.28]]
    fine = [[-0.1, -0.05], [0.12, 0.1]]
    h = 0.003
    P, N, I = twin.bench_mesh(0.004, ext, fine, h)
    assert np.all(P[:, 2] == 0.004)
    assert _area(P, I).sum() == pytest.approx(0.55 * 0.48, rel=1e-12)
    n = np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]])
    assert np.all(n[:, 2] > 0)                      # counter-clockwise seen from +z
    fine_tris = I[: (len(I) - 8)]
    e = _edges(fine_tris)
    L = np.linalg.norm(P[e[:, 0]] - P[e[:, 1]], axis=1)
    axis_aligned = (np.abs(P[e[:, 0], 0] - P[e[:, 1], 0]) < 1e-12) | (np.abs(P[e[:, 0], 1] - P[e[:, 1], 1]) < 1e-12)
    assert L[axis_aligned].max() <= h


# ---------------------------------------------------------------------------- GPU

def _twin_frame(prep, scene, spp, depth):
    cams = twin.cameras(scene, spp=spp, depth=depth, seed=7, prep=prep)
    poses = twin.poses_from_state(scene, twin.default_state(scene))
    return cams, poses


@pytest.mark.gpu
@needs_assets
def test_gpu_twin_frame_matches_oracle(loaded, prep):
    """CAM3 (the scope, a cannula, both instruments and another camera's lens front in view) at 4 spp, depth 3,
    bodies at the default state: a tile around the instruments against the oracle on the same samples."""
    scene, _, _ = loaded
    cams, poses = _twin_frame(prep, scene, 4, 3)
    inst = twin.instances(prep, poses)
    print("twin frame: upload %d triangles, %d bodies" % (prep["tris"].size // 16, len(prep["bodies"])), flush=True)
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(prep)
        ctx.upload_profiles(prep)
        ctx.set_instances(inst)
        for k, c in enumerate(cams):
            ctx.set_camera(k, c)
        ctx.render([0, 1, 2], mode=lib.VT_RENDER_BATCHED)
        imgs = [ctx.readback(k, lib.VT_KIND_LINEAR, c["width"], c["height"]) for k, c in enumerate(cams)]
        st = ctx.stats()
    finally:
        ctx.close()
    assert st["stack_overflows"] == 0
    for im in imgs:
        assert np.isfinite(im).all() and im[..., :3].mean() > 0
    cam = cams[2]
    x0, y0, w, h = 600, 300, 48, 32
    sc = dict(prep)
    sc["poses"] = inst
    ref, _, _ = oracle.render_tile(sc, cam, x0, y0, w, h)
    a = imgs[2][y0:y0 + h, x0:x0 + w, :3].astype(np.float64)
    b = ref[:, :, :3].astype(np.float64)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    print("twin CAM3 tile: relative RMSE %.3g (GPU mean %.4g, oracle %.4g)" % (rel, a.mean(), b.mean()))
    assert rel < 1e-3


@pytest.mark.gpu
@needs_assets
def test_gpu_twin_reference_estimator_matches_oracle(loaded, prep):
    """The reference's own camera estimator on the twin (shadeHit over the node radiosity cache of 445,968 static
    nodes, bodies at the default state with their occlusion capsules and rho_eff from the albedo tables): CAM3 at
    the judge preset's 4 light samples, against the oracle fed the GPU's node state."""
    scene, _, _ = loaded
    cams, poses = _twin_frame(prep, scene, 4, 1)
    print("twin reference estimator: build the node cache", flush=True)
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(prep)
        ctx.upload_profiles(prep)
        ctx.set_instances(twin.instances(prep, poses))
        info = ctx.radiosity_build()
        n_mat = np.asarray(prep["materials"]).size // 16
        hh = ctx.radiosity_tables(6, (n_mat, 4), np.float32)
        alb = ctx.radiosity_tables(5, (n_mat, 32, 8), np.float32)
        twin.rho_eff(prep, hh[:, :3])
        inst = twin.instances(prep, poses)
        ctx.set_instances(inst)
        ctx.radiosity_update(direct=True, iters=96)
        nodes = np.asarray(prep["verts"]).size // 8
        st, e = ctx.radiosity_state(nodes)
        cam = dict(cams[2], flags=lib.VT_OPT_REFERENCE | lib.VT_OPT_CULL_SHADOWS, light_samples=4)
        ctx.set_camera(0, cam)
        ctx.render([0])
        img = ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"])
    finally:
        ctx.close()
    print("nodes %d, active %d, eAvg %s" % (info["nodes"], info["active"], e))
    assert info["static_nodes"] == prep["stats"]["static_nodes"]
    assert np.isfinite(img).all() and img[..., :3].mean() > 0
    x0, y0, w, h = 600, 300, 48, 32
    sc2 = dict(prep, poses=inst, node_state=st, alb=alb, hh=hh, slot=1, e_avg=e)
    ref, _, _ = oracle.render_tile(sc2, cam, x0, y0, w, h)
    a = img[y0:y0 + h, x0:x0 + w, :3].astype(np.float64)
    b = ref[:, :, :3].astype(np.float64)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    close = np.mean(np.all(np.abs(a - b) <= 1e-3 * np.maximum(b, 1e-2), axis=2))
    print("twin reference estimator CAM3 tile: relative RMSE %.3g, close %.4f" % (rel, close))
    assert rel < 1e-3 and close > 0.99


def _project(cam, X):
    """KB4 projection of a world point into a vt_camera record (lens.wgsl.js's forward model): pixel or None."""
    R = np.asarray(cam["R"], float).reshape(3, 3).T          # rows: the camera axes in the world
    xc = R @ (np.asarray(X, float) - np.asarray(cam["pos"], float))
    if xc[2] <= 0:
        return None
    fx, fy, cx, cy, k1, k2, k3, k4, al = [float(v) for v in cam["lens"]]
    r = math.hypot(xc[0], xc[1])
    th = math.atan2(r, xc[2])
    t2 = th * th
    thd = th * (1 + t2 * (k1 + t2 * (k2 + t2 * (k3 + t2 * k4))))
    xd, yd = (xc[0] / r * thd, xc[1] / r * thd) if r > 0 else (0.0, 0.0)
    return fx * (xd + al * yd) + cx, fy * yd + cy


@pytest.mark.gpu
@needs_assets
def test_gpu_twin_autoplay_frames_match_oracle(loaded, prep):
    """End to end from the pose source: the seeded task set (vtwin/autoplay.py) in the middle of its reach, crossing
    and scope re-aim phases, each frame exposed over 1/30 s (poses at t and t + 1/30 s, motion-blurred bodies and
    the light moving with the scope), on the device against the oracle on a tile around an instrument's hinge in
    the camera that sees it best."""
    from vtwin import autoplay
    scene, _, _ = loaded
    script = autoplay.make_script("task_set", scene, seed=11)
    cams = twin.cameras(scene, spp=4, depth=3, seed=5, prep=prep)
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(prep)
        ctx.upload_profiles(prep)
        for phase in ("reach", "cross", "scope_reaim"):
            t0, t1 = next((a, b) for a, b, n in script.labels if n == phase)
            t = 0.5 * (t0 + t1)
            p0 = twin.poses_from_state(scene, script.state_at(t))
            p1 = twin.poses_from_state(scene, script.state_at(t + 1.0 / 30.0))
            inst = twin.instances(prep, p0, p1)
            ctx.set_instances(inst)
            st = script.state_at(t)
            q = st["instruments"][0]
            hinge = autoplay.hinge_of(autoplay.port_of(scene, scene["instruments"][0]["port"]), q)
            best = None
            for k, cam in enumerate(cams):
                uv = _project(cam, hinge)
                if uv is None:
                    continue
                m = min(uv[0], cam["width"] - uv[0], uv[1], cam["height"] - uv[1])
                if best is None or m > best[0]:
                    best = (m, k, uv)
            assert best is not None and best[0] > 40
            _, k, (u, v) = best
            cam = cams[k]
            w, h = 48, 32
            x0 = int(min(max(u - w / 2, 0), cam["width"] - w))
            y0 = int(min(max(v - h / 2, 0), cam["height"] - h))
            print("autoplay %s at t = %.2f s: camera %d, tile (%d, %d)" % (phase, t, k, x0, y0), flush=True)
            ctx.set_camera(0, cam)
            ctx.render([0], mode=lib.VT_RENDER_BATCHED)
            img = ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"])
            ref, _, _ = oracle.render_tile(dict(prep, poses=inst), cam, x0, y0, w, h)
            a = img[y0:y0 + h, x0:x0 + w, :3].astype(np.float64)
            b = ref[:, :, :3].astype(np.float64)
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            close = np.mean(np.all(np.abs(a - b) <= 1e-3 * np.maximum(b, 1e-2), axis=2))
            print("  relative RMSE %.3g, pixels within 1e-3 %.4f, largest pixel difference %.3g (mean %.4g)" % (
                rel, close, np.abs(a - b).max(), b.mean()))
            # moving jaws: a sample whose FP32 hit on a posed, moving triangle lands on the other side of an edge
            # than the FP64 one changes its pixel wholly; the bar allows a few such pixels per tile
            assert b.mean() > 0 and rel < 5e-3 and close > 0.97
    finally:
        ctx.close()


@pytest.mark.gpu
@needs_assets
def test_gpu_twin_node_cache_matches_oracle(loaded, prep):
    """The twin's node cache on the device against the oracle, stage by stage (each oracle stage fed the device's
    previous one), bodies and the light on the scope at the default state: the gather table rows and miss counts
    of 3,000 seeded static nodes (the oracle traces only those), the direct light of 6,000 seeded active nodes, and
    8 Jacobi steps over every active node."""
    scene, _, _ = loaded
    cams, poses = _twin_frame(prep, scene, 1, 1)
    K = 64
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(prep)
        ctx.upload_profiles(prep)
        inst = twin.instances(prep, poses)
        ctx.set_instances(inst)
        print("twin node cache: build", flush=True)
        info = ctx.radiosity_build(K)
        V, A = info["nodes"], info["active"]
        n_mat = np.asarray(prep["materials"]).size // 16
        g = ctx.radiosity_tables(0, V * K, np.uint32)
        miss = ctx.radiosity_tables(1, V, np.float32)
        act = ctx.radiosity_tables(2, A, np.uint32)
        nb = ctx.radiosity_tables(3, (A, 16), np.uint32)
        nw = ctx.radiosity_tables(4, (A, 16), np.float32)
        alb = ctx.radiosity_tables(5, (n_mat, 32, 8), np.float32)
        hh = ctx.radiosity_tables(6, (n_mat, 4), np.float32)
        ctx.radiosity_update(direct=True, iters=0)
        st_gpu, _ = ctx.radiosity_state(V)
        ctx.radiosity_update(direct=False, iters=8)
        st_gpu8, e_gpu = ctx.radiosity_state(V)
    finally:
        ctx.close()
    sc = dict(prep, poses=inst)
    ni = oracle.node_inputs(sc)
    rng = np.random.default_rng(77)
    stat = np.nonzero(ni["static"])[0]
    sub = np.sort(rng.choice(stat, 3000, replace=False))
    npos = ni["npos"].copy()
    keep = np.zeros(V, bool)
    keep[sub] = True
    npos[~keep, 3] = 0.0                     # the oracle traces the sampled nodes' gather rays only
    nis = dict(ni, npos=npos)
    g0, miss0 = oracle.gather_table(sc, nis, K, nthreads=16)
    g1 = oracle.gather_remap(sc, nis, K, g0, miss, nthreads=16)      # hit nodes' activity from the device
    rows = g.reshape(V, K)[sub], g1.reshape(V, K)[sub]
    same_miss, same_g = np.mean(miss[sub] == miss0[sub]), np.mean(rows[0] == rows[1])
    print("gather: %d nodes, miss counts equal %.4f, entries equal %.4f" % (sub.size, same_miss, same_g))
    assert same_miss > 0.99 and same_g > 0.995
    asub = np.sort(rng.choice(act, 6000, replace=False))
    so = np.zeros_like(st_gpu)
    oracle.node_direct(sc, ni, asub, 4, alb, so, nthreads=16)
    B, Bo = st_gpu[asub, 12:15].astype(np.float64), so[asub, 12:15].astype(np.float64)
    rel = np.linalg.norm(B - Bo) / np.linalg.norm(Bo)
    codes = np.mean(st_gpu[asub, 15] == so[asub, 15])
    print("node direct: %d nodes, B_dir rel err %.2e, visibility codes equal %.4f" % (asub.size, rel, codes))
    assert np.linalg.norm(Bo) > 0 and rel <= 1e-3 and codes > 0.995
    st2 = st_gpu.copy()
    slot = oracle.jacobi_steps(sc, ni, act, g, K, nb, nw, alb, hh, st2, 1, 8)
    for k in (0, slot):
        x, y = st_gpu8[act, 4 * k:4 * k + 3].astype(np.float64), st2[act, 4 * k:4 * k + 3].astype(np.float64)
        r = np.linalg.norm(x - y) / np.linalg.norm(y)
        print("Jacobi 8 steps, state word %d: rel err %.2e over %d active nodes" % (k, r, act.size))
        assert r <= 1e-5
    assert np.allclose(e_gpu, oracle.e_avg(ni, act, st2, hh), rtol=1e-5)


@pytest.mark.gpu
@needs_assets
@pytest.mark.parametrize("caches", [False, True])
def test_gpu_twin_judge_preset_matches_oracle(loaded, prep, caches):
    """The reference's judge preset on the twin (1 sub-sample x 4 light samples, then 6 refinement sub-samples on
    edges, posed bodies and shadow pixels): CAM3 around the instruments against the oracle's two passes; with the
    fixed camera's caches (the reference's default: the still-light shading and its keys), a second frame from the
    kept caches equal to the first."""
    scene, _, _ = loaded
    cams, poses = _twin_frame(prep, scene, 1, 1)
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(prep)
        ctx.upload_profiles(prep)
        ctx.set_instances(twin.instances(prep, poses))
        ctx.radiosity_build()
        n_mat = np.asarray(prep["materials"]).size // 16
        hh = ctx.radiosity_tables(6, (n_mat, 4), np.float32)
        alb = ctx.radiosity_tables(5, (n_mat, 32, 8), np.float32)
        twin.rho_eff(prep, hh[:, :3])
        inst = twin.instances(prep, poses)
        ctx.set_instances(inst)
        ctx.radiosity_update(direct=True, iters=96)
        st, e = ctx.radiosity_state(np.asarray(prep["verts"]).size // 8)
        flags = lib.VT_OPT_REFERENCE | lib.VT_OPT_CULL_SHADOWS | lib.VT_OPT_REFINE_SHADOWS
        if caches:
            flags |= lib.VT_OPT_HIT_CACHE | lib.VT_OPT_SHADOW_CACHE
        cam = dict(cams[2], flags=flags, light_samples=4, refine=6, refine_light_samples=1)
        print("twin judge preset, caches %s" % caches, flush=True)
        ctx.set_camera(0, cam)
        ctx.render([0])
        img = ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"])
        samples = ctx.stats()["samples"]
        if caches:
            r1 = ctx.stats()["rays"]
            ctx.render([0])
            assert np.array_equal(ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"]), img)
            print("rays: %d from empty caches, %d from kept ones" % (r1, ctx.stats()["rays"]))
    finally:
        ctx.close()
    px = cam["width"] * cam["height"]
    print("judge preset: %.1f %% of the pixels refined" % (100.0 * (samples - px) / 6 / px))
    x0, y0, w, h = 600, 300, 48, 32
    sc2 = dict(prep, poses=inst, node_state=st, alb=alb, hh=hh, slot=1, e_avg=e)
    ref, _, _ = oracle.render_tile(sc2, cam, x0, y0, w, h)
    a = img[y0:y0 + h, x0:x0 + w, :3].astype(np.float64)
    b = ref[:, :, :3].astype(np.float64)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    close = np.mean(np.all(np.abs(a - b) <= 1e-3 * np.maximum(b, 1e-2), axis=2))
    print("twin judge preset CAM3 tile: relative RMSE %.3g, close %.4f" % (rel, close))
    assert rel < 1e-3 and close > 0.99



@pytest.mark.gpu
@needs_assets
def test_gpu_twin_interior_tree(loaded, prep):
    """The interior tree (_buildInteriorBVH, gpu/renderer.js:533-571): it keeps exactly the static triangles on the
    bench or with a node whose gather rays escape in fewer than 90 % of cases (counted here from the miss table),
    dropping the shell's outer skin; CAM3, inside the shell, sees the same frame through it as through the whole
    scene but for the few rays that leave through an opening and would meet the skin (under 0.1 % of the pixels,
    the reference's own frame), no slower; rays from outside the scene meet the skin in the whole tree and a
    triangle no nearer in the interior tree."""
    scene, _, _ = loaded
    cams, poses = _twin_frame(prep, scene, 1, 1)
    tw = np.asarray(prep["tris"], np.float32).view(np.uint32).reshape(-1, 16)
    static = np.asarray(prep["tri_body"]) == 0xFFFFFFFF
    vid, mesh = tw[static][:, [3, 7, 11]], tw[static][:, 12]
    bench = (np.asarray(prep["meshes"]).reshape(-1, 4)[mesh, 1] & 64) != 0
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(prep)
        ctx.upload_profiles(prep)
        inst = twin.instances(prep, poses)
        ctx.set_instances(inst)
        info = ctx.radiosity_build()
        miss = ctx.radiosity_tables(1, info["nodes"], np.float32)
        skin = np.all(miss[vid] >= 0.9 * 64, axis=1) & ~bench
        n_int = ctx.interior_bvh(True)
        print("interior tree: %d of %d static triangles (%d skin)" % (n_int, static.sum(), skin.sum()), flush=True)
        assert n_int == static.sum() - skin.sum() and 0 < skin.sum()
        ctx.radiosity_update(direct=True, iters=32)
        cam = dict(cams[2], flags=lib.VT_OPT_REFERENCE | lib.VT_OPT_CULL_SHADOWS, light_samples=4)
        ctx.set_camera(0, cam)
        img, times = [None, None], [[], []]
        # alternate the two trees (the clocks and caches warm up for both alike); the fastest frame of each
        for rep in range(4):
            for k, enable in enumerate((True, False)):
                assert ctx.interior_bvh(enable) == (n_int if enable else static.sum())
                ctx.render([0])
                if rep == 0:
                    img[k] = ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"])
                times[k].append(ctx.stats()["ms_render"])
        print("frame times (ms), interior tree %s, whole tree %s" % (np.round(times[0], 3), np.round(times[1], 3)))
        ms = [min(t[1:]) for t in times]
        # rays from outside the static scene's box toward points inside it: the interior tree holds a subset of the
        # whole tree's triangles (same geometry), so its closest hit is never nearer, and the skin the whole tree
        # meets first is gone from it
        pos = np.asarray(prep["tris"], np.float32).reshape(-1, 16)[static][:, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(-1, 3)
        lo, hi = pos.min(axis=0), pos.max(axis=0)
        ctr, rad = 0.5 * (lo + hi), 0.75 * float(np.linalg.norm(hi - lo))
        rng = np.random.default_rng(11)
        n = 20000
        u = rng.normal(size=(n, 3))
        org = ctr + rad * u / np.linalg.norm(u, axis=1, keepdims=True)
        tgt = lo + (hi - lo) * (0.25 + 0.5 * rng.random((n, 3)))
        d = (tgt - org) / np.linalg.norm(tgt - org, axis=1, keepdims=True)
        rays = np.concatenate([org, np.zeros((n, 1)), d, np.full((n, 1), np.inf)], axis=1)
        hits = []
        for enable in (True, False):
            ctx.interior_bvh(enable)
            hits.append(ctx.trace(rays))
    finally:
        ctx.close()
    t_int, t_all = hits[0][:, 0].astype(np.float64), hits[1][:, 0].astype(np.float64)
    hit_all = np.isfinite(t_all)
    farther = np.mean(t_int[hit_all] > t_all[hit_all] * (1 + 1e-6))
    print("rays from outside: %d hit the whole tree, %.1f %% of them meet a farther triangle in the interior tree"
          % (hit_all.sum(), 100 * farther))
    assert np.all(t_int[hit_all] >= t_all[hit_all] * (1 - 1e-6)) and hit_all.mean() > 0.5 and farther > 0.1
    a, b = img[0][..., :3].astype(np.float64), img[1][..., :3].astype(np.float64)
    diff = np.any(img[0] != img[1], axis=2).mean()
    print("interior vs whole tree: %.4f %% of the pixels differ, render %.3f vs %.3f ms" % (100 * diff, ms[0], ms[1]))
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    print("relative difference %.3g" % rel)
    # the frame's walk is no slower through the smaller tree (the two times within clock noise of each other)
    assert diff < 1e-3 and rel < 1e-2 and ms[0] < ms[1] * 1.05
