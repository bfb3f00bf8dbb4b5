"""The node radiosity cache and the reference camera estimator that reads it (gpu/renderer.js:456-530, 707-852;
gpu/wgsl/kernels.wgsl.js:271-461; gpu/wgsl/shade.wgsl.js:57-334).

Pins (CPU): the smoothing neighbours against brute force; the whole oracle pipeline (gather table, node direct
light, Jacobi, reference estimator) inside a closed Lambertian sphere with an emitting cap against the
infinite-bounce closed form rho Le f / (1 - rho (1 - f)) that a converged node cache must reproduce.
GPU parity, stage by stage against the oracle fed the GPU's own inputs: albedo tables, gather table and activity,
node direct light, Jacobi steps, the reference estimator's image (static scene; posed bodies with occlusion
capsules; option bits; shadow culling); and the sphere's closed form through the GPU pipeline.
"""
import math

import numpy as np
import pytest

import oracle
from vtwin import lib, scenes

OBVH = lambda p, i: (lambda b: (b["nodes"], b["order"]))(oracle.build_bvh(p, i))


def _brute_neighbours(pos, nrm, k, R, sg):
    n = pos.shape[0]
    idx = np.zeros((n, k), np.uint32)
    w = np.zeros((n, k), np.float32)
    for a in range(n):
        d2 = np.sum((pos.astype(np.float64) - pos[a]) ** 2, axis=1)
        ok = (d2 <= R * R) & (nrm.astype(np.float64) @ nrm[a].astype(np.float64) >= 0.9)
        cand = np.nonzero(ok)[0]
        ww = np.exp(-0.5 * d2[cand] / (sg * sg))
        order = np.lexsort((cand, -ww))[:k]
        idx[a, : order.size] = cand[order]
        w[a, : order.size] = ww[order]
    return idx, w


def test_neighbours_match_brute_force():
    rng = np.random.default_rng(4)
    pos = rng.uniform(0, 0.02, size=(600, 3)).astype(np.float32)
    pos[:300, 2] = 0.0                                  # a plane and a cloud
    nrm = np.zeros((600, 3), np.float32)
    nrm[:, 2] = 1.0
    nrm[::7] = [0.0, 1.0, 0.0]                          # some face away
    idx, w = lib.node_neighbours(pos, nrm, 16, 0.006, 0.003)
    bi, bw = _brute_neighbours(pos, nrm, 16, 0.006, 0.003)
    assert np.array_equal(idx[bw > 0], bi[bw > 0])
    assert np.allclose(w, bw, rtol=1e-6, atol=0)
    i0, w0 = lib.node_neighbours(pos, nrm, 16, 0.006, 0.0)      # no smoothing: every node alone
    assert np.array_equal(i0[:, 0], np.arange(600)) and np.all(w0[:, 0] == 1) and np.all(w0[:, 1:] == 0)


def _sphere(rho=0.6, Le=4.0, n_theta=24, n_phi=48, cap_rows=5):
    from test_oracle_pins import furnace_sphere
    sc, cam, _ = furnace_sphere(rho=rho, Le=Le, n_theta=n_theta, n_phi=n_phi, cap_rows=cap_rows, depth=1)
    P, I, M = sc["pos"], sc["idx"], sc["tri_mesh"]
    a, b, c = P[I[:, 0]], P[I[:, 1]], P[I[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    f = area[M == 1].sum() / area.sum()
    cam = dict(cam, flags=lib.VT_OPT_REFERENCE, spp=64)
    return sc, cam, rho * Le * f / (1 - rho * (1 - f))


def _oracle_cache(sc, K, samples, nbr_k=16, radius=0.006, sigma=0.0, iters=0):
    ni = oracle.node_inputs(sc)
    alb, hh = oracle.albedo_tables(sc)
    g, miss = oracle.gather_table(sc, ni, K)
    active = np.nonzero(ni["static"] & (miss <= 0.5 * K))[0].astype(np.uint32)
    g = oracle.gather_remap(sc, ni, K, g, miss)
    pos = ni["npos"][active, :3]
    nrm = ni["nnorm"][active, :3]
    nb, nw = lib.node_neighbours(pos, nrm, nbr_k, radius, sigma)
    st = np.zeros((ni["npos"].shape[0], 16), np.float32)
    oracle.node_direct(sc, ni, active, samples, alb, st)
    slot = oracle.jacobi_steps(sc, ni, active, g, K, nb, nw, alb, hh, st, 1, iters)
    return dict(ni=ni, alb=alb, hh=hh, g=g, miss=miss, active=active, nb=nb, nw=nw, state=st, slot=slot)


def test_oracle_reference_estimator_sphere_closed_form():
    """Converged, the cache carries every bounce: the reference estimator's pixels (NEE + E_free A / pi) equal the
    infinite-bounce furnace value (gather rays K = 256, 64 emitter samples per node, no smoothing)."""
    sc, cam, want = _sphere()
    cc = _oracle_cache(sc, 256, 64, iters=60)
    sc2 = dict(sc, node_state=cc["state"], alb=cc["alb"], hh=cc["hh"], slot=cc["slot"])
    img, _, _ = oracle.render_tile(sc2, cam, 16, 16, 32, 32)
    got = img[:, :, 0].mean()
    print("reference estimator in the sphere: %.5f, closed form %.5f" % (got, want))
    assert got == pytest.approx(want, rel=0.02)
    # the path tracer of the same scene to depth 32 agrees with both
    pt, _, _ = oracle.render_tile(sc, dict(cam, flags=0, max_depth=32, spp=64), 16, 16, 32, 32)
    assert pt[:, :, 0].mean() == pytest.approx(want, rel=0.02)


def _direct_inputs(sc, poses=None):
    """Estimator inputs for direct-light frames (VT_OPT_DIRECT_ONLY): the node cache is never read."""
    alb, hh = oracle.albedo_tables(sc)
    V = np.asarray(sc["verts"]).size // 8
    out = dict(sc, node_state=np.zeros((V, 16), np.float32), alb=alb, hh=hh, slot=1, e_avg=np.zeros(3, np.float32))
    if poses is not None:
        out["poses"] = poses
    return out


def test_oracle_still_light_pins():
    """The still-light shading of static surfaces (shade.wgsl.js:272-301) tests each light sample against the static
    scene, then against the bodies: the visibility of one shadow ray against both, so a frame without the refinement
    pass equals the traced-shadow estimator's exactly, posed bodies included. Its refinement keys mark body shadows,
    penumbrae and light samples whose luminances spread by more than 20 %: with one light sample and no bodies none
    can occur, and refining 'shadow' pixels adds nothing to the pixels on edges (same image, same rays); with four
    light samples it refines more pixels than edges alone and fewer than every pixel that traced a shadow ray."""
    from test_bodies import _poses
    R = lib.VT_OPT_REFERENCE | lib.VT_OPT_DIRECT_ONLY
    parts = scenes.posed_parts()
    sc = _direct_inputs(scenes.posed_scene(*parts), _poses(True))
    cam = dict(scenes.posed_camera(96, 72, spp=4), light_samples=4)
    a, _, _ = oracle.render_tile(sc, dict(cam, flags=R), 0, 0, 96, 72)
    b, _, _ = oracle.render_tile(sc, dict(cam, flags=R | lib.VT_OPT_SHADOW_CACHE), 0, 0, 96, 72)
    assert np.array_equal(a, b) and a[..., :3].max() > 0
    sc1, cams = scenes.c1(bvh=OBVH)
    sc1 = _direct_inputs(sc1)
    for nl in (1, 4):
        base = dict(cams[0], spp=1, light_samples=nl, refine=4, refine_light_samples=1)
        run = {}
        for name, f in (("edges", R), ("traced", R | lib.VT_OPT_REFINE_SHADOWS),
                        ("still", R | lib.VT_OPT_REFINE_SHADOWS | lib.VT_OPT_SHADOW_CACHE)):
            run[name] = oracle.render_tile(sc1, dict(base, flags=f), 0, 0, 640, 480)[:2]
        print("light samples %d: rays edges %d, still light %d, traced shadows %d"
              % (nl, run["edges"][1], run["still"][1], run["traced"][1]))
        if nl == 1:
            assert np.array_equal(run["still"][0], run["edges"][0]) and run["still"][1] == run["edges"][1]
            assert run["traced"][1] > run["edges"][1]
        else:
            assert run["edges"][1] < run["still"][1] < run["traced"][1]


# ---------------------------------------------------------------------------- GPU parity

def _gpu_cache(ctx, sc, K=64, samples=4, nbr_k=16, radius=0.006, sigma=0.003, iters=0, poses=None):
    ctx.upload_scene(sc)
    if poses is not None:
        ctx.set_instances(poses)
    info = ctx.radiosity_build(K, samples, nbr_k, radius, sigma)
    V = np.asarray(sc["verts"]).size // 8
    A = info["active"]
    n_mat = np.asarray(sc["materials"]).size // 16
    t = dict(info=info,
             g=ctx.radiosity_tables(0, V * K, np.uint32), miss=ctx.radiosity_tables(1, V, np.float32),
             active=ctx.radiosity_tables(2, A, np.uint32), nb=ctx.radiosity_tables(3, (A, nbr_k), np.uint32),
             nw=ctx.radiosity_tables(4, (A, nbr_k), np.float32), alb=ctx.radiosity_tables(5, (n_mat, 32, 8), np.float32),
             hh=ctx.radiosity_tables(6, (n_mat, 4), np.float32))
    return t


@pytest.fixture(scope="module")
def ctx():
    c = lib.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
def test_gpu_albedo_tables_match_oracle(ctx):
    sc, _ = scenes.c1(bvh=OBVH)
    mats = [scenes.material_record(k, (0.7, 0.5, 0.3), roughness=r, ior=1.45)
            for k in (scenes.KIND_DIELECTRIC, scenes.KIND_CONDUCTOR, scenes.KIND_COATED, scenes.KIND_LAMBERT)
            for r in (0.05, 0.3, 0.8)]
    sc["materials"] = np.concatenate([np.asarray(sc["materials"]).reshape(-1, 16)] + [m[None] for m in mats]).reshape(-1)
    t = _gpu_cache(ctx, sc)
    alb, hh = oracle.albedo_tables(sc)
    err = np.abs(t["alb"] - alb).max()
    print("albedo tables: max |GPU - FP64| %.2e, hh %.2e" % (err, np.abs(t["hh"] - hh).max()))
    assert err < 2e-5 and np.abs(t["hh"] - hh).max() < 2e-5


@pytest.mark.gpu
def test_gpu_gather_table_and_cache_match_oracle(ctx):
    """c1: the gather table and activity, node direct light and 24 Jacobi steps, each stage against the oracle fed
    the GPU's previous stage."""
    sc, _ = scenes.c1(bvh=OBVH)
    K = 64
    t = _gpu_cache(ctx, sc, K=K)
    ni = oracle.node_inputs(sc)
    g0, miss0 = oracle.gather_table(sc, ni, K)
    # the GPU's table after the remap: compare the oracle's build + remap on the oracle's own counts
    act0 = np.nonzero(ni["static"] & (miss0 <= 0.5 * K))[0]
    g1 = oracle.gather_remap(sc, ni, K, g0, miss0)
    same = np.mean(t["g"] == g1)
    print("gather entries equal: %.5f; active %d vs %d; misses equal %.5f" % (
        same, t["active"].size, act0.size, np.mean(t["miss"] == miss0)))
    assert same > 0.995 and np.mean(t["miss"] == miss0) > 0.99
    assert abs(t["active"].size - act0.size) <= 0.002 * act0.size + 2
    # node direct light and Jacobi on the GPU's own tables
    ctx.radiosity_update(direct=True, iters=0)
    st_gpu, _ = ctx.radiosity_state(ni["npos"].shape[0])
    st = np.zeros_like(st_gpu)
    oracle.node_direct(sc, ni, t["active"], 4, t["alb"], st)
    a = t["active"]
    B, Bo = st_gpu[a, 12:15], st[a, 12:15]
    print("node B_dir: rel err %.2e, visibility codes equal %.5f" % (
        np.linalg.norm(B - Bo) / np.linalg.norm(Bo), np.mean(st_gpu[a, 15] == st[a, 15])))
    assert np.linalg.norm(B - Bo) <= 1e-3 * np.linalg.norm(Bo)
    assert np.mean(st_gpu[a, 15] == st[a, 15]) > 0.995
    st2 = st_gpu.copy()
    ctx.radiosity_update(direct=False, iters=24)
    st_gpu2, e_gpu = ctx.radiosity_state(ni["npos"].shape[0])
    slot = oracle.jacobi_steps(sc, ni, a, t["g"], K, t["nb"], t["nw"], t["alb"], t["hh"], st2, 1, 24)
    for k in (0, slot):
        x, y = st_gpu2[a, 4 * k:4 * k + 3], st2[a, 4 * k:4 * k + 3]
        print("state word %d: rel err %.2e" % (k, np.linalg.norm(x - y) / np.linalg.norm(y)))
        assert np.linalg.norm(x - y) <= 1e-5 * np.linalg.norm(y)
    e = oracle.e_avg(ni, a, st2, t["hh"])
    assert np.allclose(e_gpu, e, rtol=1e-5)


@pytest.mark.gpu
def test_gpu_node_cache_follows_material_and_light_updates(ctx):
    """After vt_update_materials the node cache's albedo tables and gather weights are the new materials' (the
    next Jacobi steps equal the oracle's on the new tables), and the next update relights every node even when
    asked for Jacobi steps only; after vt_update_lights the direct light scales with the radiance."""
    sc, _ = scenes.c1_mixed(bvh=OBVH)
    K = 64
    t = _gpu_cache(ctx, sc, K=K)
    ctx.radiosity_update(direct=True, iters=8)                # slot 1 -> 8 steps end in slot 1
    ni = oracle.node_inputs(sc)
    a = t["active"]
    mats = np.asarray(sc["materials"], np.float32).reshape(-1, 16).copy()
    r = scenes.Rng(77)
    mats[:16, 0:3] = r.uniform(0.2, 0.9, 48).reshape(16, 3)
    sc2 = dict(sc, materials=mats.reshape(-1))
    ctx.update_materials(mats)
    n_mat = mats.shape[0]
    alb = ctx.radiosity_tables(5, (n_mat, 32, 8), np.float32)
    hh = ctx.radiosity_tables(6, (n_mat, 4), np.float32)
    alb_o, hh_o = oracle.albedo_tables(sc2)
    assert np.abs(alb - alb_o).max() < 2e-5 and np.abs(hh - hh_o).max() < 2e-5
    ctx.radiosity_update(direct=False, iters=0)               # stale direct light: relit anyway
    st_gpu, _ = ctx.radiosity_state(ni["npos"].shape[0])
    st = np.zeros_like(st_gpu)
    oracle.node_direct(sc2, ni, a, 4, alb, st)
    B, Bo = st_gpu[a, 12:15], st[a, 12:15]
    print("after the material update: B_dir rel err %.2e" % (np.linalg.norm(B - Bo) / np.linalg.norm(Bo)))
    assert np.linalg.norm(B - Bo) <= 1e-3 * np.linalg.norm(Bo)
    st0 = st_gpu.copy()
    ctx.radiosity_update(direct=False, iters=6)
    st_gpu2, _ = ctx.radiosity_state(ni["npos"].shape[0])
    slot = oracle.jacobi_steps(sc2, ni, a, t["g"], K, t["nb"], t["nw"], alb, hh, st0, 1, 6)
    x, y = st_gpu2[a, 4 * slot:4 * slot + 3], st0[a, 4 * slot:4 * slot + 3]
    print("Jacobi on the new materials: rel err %.2e" % (np.linalg.norm(x - y) / np.linalg.norm(y)))
    assert np.linalg.norm(x - y) <= 1e-5 * np.linalg.norm(y)
    # lights x 1.5: the next update relights, B_dir scales exactly with the radiance
    lights = np.asarray(sc["lights"], np.float32).reshape(-1, 4).copy()
    ctx.update_lights(lights * np.float32(1.5))
    ctx.radiosity_update(direct=False, iters=0)
    st_gpu3, _ = ctx.radiosity_state(ni["npos"].shape[0])
    assert np.allclose(st_gpu3[a, 12:15], 1.5 * st_gpu2[a, 12:15], rtol=2e-6, atol=1e-12)
    ctx.update_lights(lights)


def _quat_mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _seg_point_dist(p, q, x):
    """Distance from each segment p[i] -> q (q one point) to each point x[j]: (len(p), len(x))."""
    d = q - p
    t = np.clip(np.einsum("ik,ijk->ij", d, x[None, :, :] - p[:, None, :]) / np.einsum("ik,ik->i", d, d)[:, None], 0, 1)
    c = p[:, None, :] + t[..., None] * d[:, None, :]
    return np.linalg.norm(c - x[None], axis=2)


@pytest.mark.gpu
def test_gpu_incremental_node_direct(ctx):
    """The incremental node update after one body moved (_movedSinceSolve / _submitNodeDirect with a mask,
    gpu/renderer.js:756-795; nodeDirect's moved-body test, kernels.wgsl.js:357-366), under one small ceiling light.
    B_dir and the visibility codes equal a full pass at the new pose bit for bit at every active node; nodes kept
    from the last pass exist (their capsule visibility is the old pose's, as in the reference) and every one of them
    has its segment to the light's centre farther than the light's radius from the moved body's surface, then and
    now; the full pass with posed bodies matches the oracle; a stale cache falls back to a full pass, no moved body
    to none, and a mask past the body count is refused."""
    from test_bodies import _poses
    parts, meshes, mats, lights = scenes.posed_parts()
    P, N, I = scenes.quad((0.0, 0.0, 2.4), 0.3, (0.0, 0.0, -1.0))
    parts = parts[:4] + [(P, N, I, np.full(I.shape[0], 16), None)]
    sc = scenes.posed_scene(parts, meshes[:17], mats, [[40.0, 40.0, 40.0, 0.0]])
    poses = _poses(True)
    poses[1].update(occ_a=(0, 0, -0.3), occ_b=(0, 0, 0.3), occ_r=0.15, rho_eff=0.5)
    t = _gpu_cache(ctx, sc, poses=poses)
    a = t["active"]
    nodes = np.asarray(sc["verts"]).size // 8
    assert ctx.radiosity_update_moved(0b10, iters=0) == "full"        # nothing solved since the build
    stA, _ = ctx.radiosity_state(nodes)
    assert ctx.radiosity_update_moved(0, iters=0) == "none"
    assert np.array_equal(ctx.radiosity_state(nodes)[0].view(np.uint32), stA.view(np.uint32))
    posesB = [dict(p) for p in poses]
    posesB[1]["p0"] = tuple(np.asarray(poses[1]["p0"]) + (0.2, 0.15, 0.0))
    ctx.set_instances(posesB)
    assert ctx.radiosity_update_moved(0b10, iters=0) == "incremental"
    st_inc, _ = ctx.radiosity_state(nodes)
    ctx.radiosity_update(direct=True, iters=0)
    st_full, _ = ctx.radiosity_state(nodes)
    assert np.array_equal(st_inc[a, 12:16].view(np.uint32), st_full[a, 12:16].view(np.uint32))
    kept = (st_inc[a, 3] == stA[a, 3]) & (st_full[a, 3] != stA[a, 3])
    relit = (st_inc[a, 3] == st_full[a, 3]) & (st_full[a, 3] != stA[a, 3])
    print("active %d: capsule visibility changed at %d, kept %d, relit %d" % (
        a.size, int(np.sum(st_full[a, 3] != stA[a, 3])), int(kept.sum()), int(relit.sum())))
    assert kept.sum() > 0 and relit.sum() > 0 and np.all(kept | relit | (st_full[a, 3] == stA[a, 3]))
    # kept nodes: the segment to the light's centre clears the moved body's surface by the light's radius
    ni = oracle.node_inputs(sc)
    Pl = P.astype(np.float64)
    c = Pl.mean(axis=0)
    R = np.linalg.norm(Pl - c, axis=1).max()
    Pb, _, Ib, _, _ = parts[3]
    Pb = np.asarray(Pb, np.float64)
    surf = np.concatenate([Pb, Pb[np.asarray(Ib)].mean(axis=1)])
    for pose in (poses[1], posesB[1]):
        w = surf @ _quat_mat(pose["q0"]).T + np.asarray(pose["p0"])
        dmin = _seg_point_dist(ni["npos"][a[kept], :3].astype(np.float64), c, w).min()
        print("kept nodes: closest approach to the moved body %.4f m (light radius %.4f m)" % (dmin, R))
        assert dmin > R
    # the full pass with posed bodies against the oracle
    st = np.zeros_like(st_full)
    oracle.node_direct(dict(sc, poses=posesB), ni, a, 4, t["alb"], st)
    B, Bo = st_full[a, 12:15], st[a, 12:15]
    print("node B_dir with bodies: rel err %.2e, visibility codes equal %.5f" % (
        np.linalg.norm(B - Bo) / np.linalg.norm(Bo), np.mean(st_full[a, 15] == st[a, 15])))
    assert np.linalg.norm(B - Bo) <= 1e-3 * np.linalg.norm(Bo)
    assert np.mean(st_full[a, 15] == st[a, 15]) > 0.995
    # a material change makes the cache stale: the next incremental request relights everything
    ctx.update_materials(np.asarray(sc["materials"], np.float32))
    assert ctx.radiosity_update_moved(0b01, iters=0) == "full"
    with pytest.raises(lib.VTError):
        ctx.radiosity_update_moved(0b100, iters=0)


@pytest.mark.gpu
@pytest.mark.parametrize("opts", [0, lib.VT_OPT_CULL_SHADOWS, lib.VT_OPT_DIRECT_ONLY, lib.VT_OPT_NO_INDIRECT,
                                  lib.VT_OPT_NORMALS])
def test_gpu_reference_estimator_matches_oracle(ctx, opts):
    sc, cams = scenes.c1(bvh=OBVH)
    t = _gpu_cache(ctx, sc)
    ctx.radiosity_update(direct=True, iters=16)              # slot 1 -> 16 steps end in slot 1
    nodes = np.asarray(sc["verts"]).size // 8
    st, e = ctx.radiosity_state(nodes)                        # also refreshes eAvg for the render
    cam = dict(cams[0], spp=16, flags=lib.VT_OPT_REFERENCE | opts, light_samples=2)
    print("reference estimator, options %d" % opts, flush=True)
    ctx.set_camera(0, cam)
    ctx.render([0])
    img = ctx.readback(0, lib.VT_KIND_LINEAR, 640, 480)
    x0, y0, w, h = 256, 160, 64, 48
    sc2 = dict(sc, node_state=st, alb=t["alb"], hh=t["hh"], slot=1, e_avg=e)
    ref, _, _ = oracle.render_tile(sc2, cam, x0, y0, w, h)
    a = img[y0:y0 + h, x0:x0 + w, :3].astype(np.float64)
    b = ref[:, :, :3].astype(np.float64)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    close = np.mean(np.all(np.abs(a - b) <= 1e-3 * np.maximum(b, 1e-2), axis=2))
    print("GPU vs oracle reference estimator: relative RMSE %.3g, close %.4f" % (rel, close))
    assert rel < 1e-3 and close > 0.99


@pytest.mark.gpu
@pytest.mark.parametrize("refine_shadows", [False, True])
def test_gpu_refinement_pass_matches_oracle(ctx, refine_shadows):
    """camRefine: pixels on a mesh edge (or with traced shadow rays) take 6 more sub-samples of 1 light sample after a
    first pass of 1 sub-sample x 4 light samples (the judge preset); the image and the refined pixel count against
    the oracle's two passes."""
    sc, cams = scenes.c1(bvh=OBVH)
    t = _gpu_cache(ctx, sc)
    ctx.radiosity_update(direct=True, iters=16)
    nodes = np.asarray(sc["verts"]).size // 8
    st, e = ctx.radiosity_state(nodes)
    flags = lib.VT_OPT_REFERENCE | lib.VT_OPT_CULL_SHADOWS | (lib.VT_OPT_REFINE_SHADOWS if refine_shadows else 0)
    cam = dict(cams[0], spp=1, flags=flags, light_samples=4, refine=6, refine_light_samples=1)
    print("refinement pass, refine shadows %s" % refine_shadows, flush=True)
    ctx.set_camera(0, cam)
    ctx.render([0])
    img = ctx.readback(0, lib.VT_KIND_LINEAR, 640, 480)
    samples = ctx.stats()["samples"]
    ctx.set_camera(0, dict(cam, refine=0))
    ctx.render([0])
    plain = ctx.readback(0, lib.VT_KIND_LINEAR, 640, 480)
    refined = samples - 640 * 480
    print("refined pixels: %d (%.2f %%)" % (refined // 6, 100.0 * refined / 6 / (640 * 480)))
    assert refined % 6 == 0 and 0 < refined // 6 < 640 * 480
    changed = np.any(img != plain, axis=2)
    assert changed.sum() <= refined // 6 and changed.sum() > 0.5 * refined // 6
    x0, y0, w, h = 256, 160, 64, 48
    sc2 = dict(sc, node_state=st, alb=t["alb"], hh=t["hh"], slot=1, e_avg=e)
    ref, _, _ = oracle.render_tile(sc2, cam, x0, y0, w, h)
    a = img[y0:y0 + h, x0:x0 + w, :3].astype(np.float64)
    b = ref[:, :, :3].astype(np.float64)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    close = np.mean(np.all(np.abs(a - b) <= 1e-3 * np.maximum(b, 1e-2), axis=2))
    print("GPU vs oracle with the refinement pass: relative RMSE %.3g, close %.4f" % (rel, close))
    assert rel < 1e-3 and close > 0.99


@pytest.mark.gpu
def test_gpu_reference_estimator_with_bodies(ctx):
    """Posed bodies: one BSDF bounce into the cache, the bodies' occlusion capsules dim the static surfaces'
    cached light, moving over the exposure."""
    from test_bodies import _poses
    parts = scenes.posed_parts()
    sc = scenes.posed_scene(*parts)
    poses = _poses(True)
    poses[0].update(occ_a=(-0.25, 0, 0), occ_b=(0.25, 0, 0), occ_r=0.15, rho_eff=0.3)
    poses[1].update(occ_a=(0, 0, -0.3), occ_b=(0, 0, 0.3), occ_r=0.15, rho_eff=0.5)
    t = _gpu_cache(ctx, sc, poses=poses)
    ctx.radiosity_update(direct=True, iters=16)
    nodes = np.asarray(sc["verts"]).size // 8
    st, e = ctx.radiosity_state(nodes)            # eAvg: what a body's bounce onto another body reads
    assert np.all(e > 0)
    for opts in (0, lib.VT_OPT_NO_CAPSULE, lib.VT_OPT_NO_MIS, lib.VT_OPT_CULL_SHADOWS):
        cam = dict(scenes.posed_camera(96, 72, spp=8), flags=lib.VT_OPT_REFERENCE | opts)
        print("reference estimator with bodies, options %d" % opts, flush=True)
        ctx.set_camera(0, cam)
        ctx.render([0])
        img = ctx.readback(0, lib.VT_KIND_LINEAR, 96, 72)
        sc2 = dict(sc, node_state=st, alb=t["alb"], hh=t["hh"], slot=1, e_avg=e, poses=poses)
        ref, _, _ = oracle.render_tile(sc2, cam, 0, 0, 96, 72)
        a, b = img[:, :, :3].astype(np.float64), ref[:, :, :3].astype(np.float64)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print("GPU vs oracle: relative RMSE %.3g" % rel)
        assert rel < 1e-3


@pytest.mark.gpu
def test_gpu_fixed_camera_caches_are_exact(ctx):
    """The fixed camera's caches (hitCache, shadowCache; gpu/renderer.js:940-957): frames that read the static first
    hits and static occlusion kept from earlier frames -- the bodies moved in between -- equal frames rendered from
    empty caches bit for bit, with and without the first-hit cache, refinement pass included; the cached frame traces
    fewer rays; the still-light frame against the oracle."""
    from test_bodies import _poses
    parts = scenes.posed_parts()
    sc = scenes.posed_scene(*parts)
    pa, pb = _poses(False), _poses(True)
    for p in (pa, pb):
        p[0].update(occ_a=(-0.25, 0, 0), occ_b=(0.25, 0, 0), occ_r=0.15, rho_eff=0.3)
        p[1].update(occ_a=(0, 0, -0.3), occ_b=(0, 0, 0.3), occ_r=0.15, rho_eff=0.5)
    t = _gpu_cache(ctx, sc, poses=pa)
    ctx.radiosity_update(direct=True, iters=16)
    nodes = np.asarray(sc["verts"]).size // 8
    st, e = ctx.radiosity_state(nodes)
    still = lib.VT_OPT_REFERENCE | lib.VT_OPT_REFINE_SHADOWS | lib.VT_OPT_SHADOW_CACHE
    cam = dict(scenes.posed_camera(96, 72, spp=2), light_samples=4, refine=4, refine_light_samples=2)
    for k, flags in enumerate((still | lib.VT_OPT_HIT_CACHE, still)):
        print("fixed-camera caches, options %d" % flags, flush=True)
        ctx.set_instances(pa)
        kept, fresh_slot = 2 * k, 2 * k + 1
        ctx.set_camera(kept, dict(cam, flags=flags))
        ctx.render([kept])
        first, r1 = ctx.readback(kept, lib.VT_KIND_LINEAR, 96, 72), ctx.stats()["rays"]
        ctx.render([kept])
        again, r2 = ctx.readback(kept, lib.VT_KIND_LINEAR, 96, 72), ctx.stats()["rays"]
        print("rays: %d from empty caches, %d from kept ones" % (r1, r2))
        assert np.array_equal(first, again) and r2 < r1
        ctx.set_instances(pb)
        ctx.render([kept])
        moved = ctx.readback(kept, lib.VT_KIND_LINEAR, 96, 72)
        ctx.set_camera(fresh_slot, dict(cam, flags=flags))      # a slot of its own: empty caches
        ctx.render([fresh_slot])
        fresh = ctx.readback(fresh_slot, lib.VT_KIND_LINEAR, 96, 72)
        assert np.array_equal(moved, fresh) and not np.array_equal(moved, first)
    sc2 = dict(sc, node_state=st, alb=t["alb"], hh=t["hh"], slot=1, e_avg=e, poses=pb)
    ref, _, _ = oracle.render_tile(sc2, dict(cam, flags=still), 0, 0, 96, 72)
    a, b = moved[:, :, :3].astype(np.float64), ref[:, :, :3].astype(np.float64)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    print("GPU still light vs oracle: relative RMSE %.3g" % rel)
    assert rel < 1e-3


@pytest.mark.gpu
def test_gpu_reference_estimator_sphere_closed_form(ctx):
    sc, cam, want = _sphere()
    _gpu_cache(ctx, sc, K=256, samples=64, sigma=0.0)
    ctx.radiosity_update(direct=True, iters=60)
    ctx.set_camera(0, cam)
    ctx.render([0])
    img = ctx.readback(0, lib.VT_KIND_LINEAR, 64, 64)
    got = img[16:48, 16:48, 0].mean()
    print("GPU reference estimator in the sphere: %.5f, closed form %.5f" % (got, want))
    assert got == pytest.approx(want, rel=0.02)


@pytest.mark.gpu
def test_gpu_reference_estimator_rejects_misuse(ctx):
    sc, cams = scenes.c1(bvh=OBVH)
    ctx.upload_scene(sc)
    cam = dict(cams[0], spp=1, flags=lib.VT_OPT_REFERENCE)
    ctx.set_camera(0, cam)
    with pytest.raises(lib.VTError):
        ctx.render([0])                                    # no cache built
    ctx.radiosity_build()
    with pytest.raises(lib.VTError):
        ctx.render([0], mode=lib.VT_RENDER_WAVEFRONT)      # the wavefront form runs the path tracer
    ctx.set_camera(1, dict(cams[0], spp=1, flags=0, cam_index=1))
    with pytest.raises(lib.VTError):
        ctx.render([0, 1])                                 # one estimator per frame


def _capsule_cases(n=4000, seed=12):
    rng = np.random.default_rng(seed)
    c = np.zeros((n, 16), np.float32)
    a = rng.uniform(-0.5, 0.5, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    L = rng.uniform(0.05, 0.8, n)
    r = rng.uniform(0.005, 0.2, n)
    p = a + d * L[:, None] * rng.uniform(-0.5, 1.5, (n, 1)) + rng.normal(size=(n, 3)) * rng.uniform(0.05, 1.0, (n, 1))
    nn = rng.normal(size=(n, 3))
    nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    c[:, 0:3], c[:, 4:7], c[:, 8:11], c[:, 11], c[:, 12:15] = p, nn, a, r, a + d * L[:, None]
    return c


def _cavity_vis(poses, pts, nrm):
    """or_capsule_vis of the posed_parts bodies at the given poses over points and normals."""
    sc = dict(scenes.posed_scene(*scenes.posed_parts()), poses=poses)
    S = oracle.scene_struct(sc)
    return np.array([oracle.lib().or_capsule_vis(oracle.C.byref(S), oracle.ptr(np.ascontiguousarray(p, np.float64)),
                                                 oracle.ptr(np.ascontiguousarray(n, np.float64)))
                     for p, n in zip(pts, nrm)])


def test_oracle_cavity_cut_is_the_shortened_capsule():
    """cavityCapsule (gpu/scene_gpu.js:486-503): body 1's capsule (0, 0, -0.3)-(0, 0, 0.3), r 0.15, posed at the
    identity, cut at s0 along an axis from an origin: the capsule visibility everywhere equals that of the capsule
    shortened by hand to the cut point (axis along +z, along -z, and tilted); a cut below both ends changes nothing,
    one above both removes the capsule, and a cavity radius replaces the radius."""
    rng = np.random.default_rng(31)
    pts = rng.uniform(-0.6, 0.6, (300, 3))
    pts[:, 2] = rng.uniform(-0.5, 0.5, 300)
    nrm = rng.normal(size=(300, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    base = dict(q0=(0.0, 0.0, 0.0, 1.0), p0=(0.0, 0.0, 0.0), occ_a=(0, 0, -0.3), occ_b=(0, 0, 0.3), occ_r=0.15,
                rho_eff=0.2)
    far = dict(q0=(0.0, 0.0, 0.0, 1.0), p0=(9.0, 9.0, 9.0))
    vis = lambda b1: _cavity_vis([far, b1], pts, nrm)
    v_full = vis(base)
    assert np.mean(v_full < 0.999) > 0.3
    tilt = np.array([0.3, -0.2, 1.0]) / np.linalg.norm([0.3, -0.2, 1.0])
    o = np.array([0.05, 0.02, -0.04])
    # the point of the segment where (X - o) . tilt = 0.12, by hand
    A, B = np.array([0, 0, -0.3]), np.array([0, 0, 0.3])
    t = (0.12 - (A - o) @ tilt) / ((B - A) @ tilt)
    X = A + (B - A) * t
    for cav, a, b in ((dict(origin=(0, 0, 0), axis=(0, 0, 1), s0=0.1), (0, 0, 0.1), (0, 0, 0.3)),
                      (dict(origin=(0, 0, 0), axis=(0, 0, -2.0), s0=0.1), (0, 0, -0.3), (0, 0, -0.1)),
                      (dict(origin=tuple(o), axis=tuple(tilt * 3.0), s0=0.12), tuple(X), (0, 0, 0.3))):
        got = vis(dict(base, cavity=cav))
        want = vis(dict(base, occ_a=a, occ_b=b))
        print("cut %s: max |cut - shortened| %.2e, changed vs uncut %.3f" % (
            cav["axis"], np.abs(got - want).max(), np.mean(got != v_full)))
        # the hand-cut ends are stored as float32 in the body record, the library's cut point stays in FP64
        assert np.abs(got - want).max() < 1e-7 and np.mean(got != v_full) > 0.1
    assert np.array_equal(vis(dict(base, cavity=dict(origin=(0, 0, 0), axis=(0, 0, 1), s0=-0.5))), v_full)
    assert np.all(vis(dict(base, cavity=dict(origin=(0, 0, 0), axis=(0, 0, 1), s0=0.5))) == 1.0)
    assert np.array_equal(vis(dict(base, cavity=dict(origin=(0, 0, 0), axis=(0, 0, 1), s0=-0.5, r=0.05))),
                          vis(dict(base, occ_r=0.05)))


@pytest.mark.gpu
def test_gpu_cavity_cut_matches_oracle(ctx):
    """The reference estimator with both bodies' occlusion capsules cut to a cavity (a moving box cut across its
    long axis, the cylinder cut with a sheath radius) against the oracle on the same samples; the cut changes the
    image."""
    from test_bodies import _poses
    parts = scenes.posed_parts()
    sc = scenes.posed_scene(*parts)
    poses = _poses(True)
    poses[0].update(occ_a=(-0.25, 0, 0), occ_b=(0.25, 0, 0), occ_r=0.15, rho_eff=0.3,
                    cavity=dict(origin=(0.75, -1.45, 0.3), axis=(1.0, 0.2, 0.0), s0=0.02))
    poses[1].update(occ_a=(0, 0, -0.3), occ_b=(0, 0, 0.3), occ_r=0.15, rho_eff=0.5,
                    cavity=dict(origin=(-0.75, -1.35, 0.45), axis=(0.0, -0.5, 1.0), s0=-0.05, r=0.1))
    t = _gpu_cache(ctx, sc, poses=poses)
    ctx.radiosity_update(direct=True, iters=16)
    st, e = ctx.radiosity_state(np.asarray(sc["verts"]).size // 8)
    cam = dict(scenes.posed_camera(96, 72, spp=8), flags=lib.VT_OPT_REFERENCE)
    ctx.set_camera(0, cam)
    ctx.render([0])
    img = ctx.readback(0, lib.VT_KIND_LINEAR, 96, 72)
    sc2 = dict(sc, node_state=st, alb=t["alb"], hh=t["hh"], slot=1, e_avg=e, poses=poses)
    ref, _, _ = oracle.render_tile(sc2, cam, 0, 0, 96, 72)
    a, b = img[:, :, :3].astype(np.float64), ref[:, :, :3].astype(np.float64)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    uncut = [{k: v for k, v in p.items() if k != "cavity"} for p in poses]
    ctx.set_instances(uncut)
    ctx.render([0])
    img2 = ctx.readback(0, lib.VT_KIND_LINEAR, 96, 72)
    changed = np.mean(np.any(img2 != img, axis=2))
    print("cavity cut: GPU vs oracle relative RMSE %.3g; %.3f of the pixels change without the cut" % (rel, changed))
    assert rel < 1e-3 and changed > 0.01
    with pytest.raises(lib.VTError):
        bad = [dict(p) for p in poses]
        bad[0]["cavity"] = dict(origin=(0, 0, 0), axis=(0, 0, 0), s0=0.0)
        ctx.set_instances(bad)


@pytest.mark.gpu
def test_gpu_capsule_occlusion_matches_oracle(ctx):
    """capsuleOcc on random points, normals and capsules (in and out of the horizon, near and far) against the
    FP64 restatement pinned by the reference's golden values."""
    c = _capsule_cases()
    got = ctx.capsule_test(c)
    cd = c.astype(np.float64)
    want = np.array([oracle.lib().or_capsule_occ(oracle.ptr(np.ascontiguousarray(x[0:3])), oracle.ptr(np.ascontiguousarray(x[4:7])),
                                                 oracle.ptr(np.ascontiguousarray(x[8:11])), oracle.ptr(np.ascontiguousarray(x[12:15])), x[11])
                     for x in cd])
    err = np.abs(got - want)
    k = int(np.argmax(err))
    print("capsule occlusion: max |GPU - FP64| %.2e (case %d: %.6f vs %.6f), mean %.2e, %d nonzero" % (
        err.max(), k, got[k], want[k], err.mean(), np.sum(want > 0)))
    assert np.sum(want > 1e-3) > 500
    assert err.max() < 1e-4
