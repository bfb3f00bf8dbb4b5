"""Posed bodies (the reference's BLAS instances, gpu/scene_gpu.js:313-356; setPoses gpu/renderer.js:619-633;
traceSceneSkip gpu/wgsl/trace.wgsl.js:144-164): every body has a BVH of its own in its own frame and is placed at
each sub-sample's exposure time by slerp / lerp of its shutter-open and shutter-close poses.

Oracle pins (CPU): a body posed at (R, t) renders as the same geometry moved into the world by (R, t) and traced
as static triangles, where R comes from Rodrigues' formula (not from a quaternion); a moving body seen at one
exposure time (a shutter interval of zero length) renders as the scene placed at the rotation angle and position
interpolated by hand; a hidden body renders as the scene without it.
GPU parity: the HIP path against the oracle on the same samples, for a static pose, a moving body over the whole
exposure and a hidden body, in the persistent and the wavefront forms; the material adjoint with bodies.
"""
import math

import numpy as np
import pytest

import oracle
from vtwin import scenes

Z, X = (0.0, 0.0, 1.0), (1.0, 0.0, 0.0)
A0, A1 = 0.35, 1.25            # body 0 turns about z from A0 to A1 over the exposure
P0, P1 = (0.75, -1.45, 0.3), (0.95, -1.2, 0.35)
B1 = (X, 0.4, (-0.75, -1.35, 0.45))


@pytest.fixture(scope="module")
def parts():
    return scenes.posed_parts()


def _poses(moving):
    q0 = scenes.quat_axis(Z, A0)
    p = [dict(q0=q0, p0=P0)]
    if moving:
        p[0].update(q1=scenes.quat_axis(Z, A1), p1=P1)
    p.append(dict(q0=scenes.quat_axis(B1[0], B1[1]), p0=B1[2]))
    return p


def _placed(parts, t=0.0, moving=False, drop=None):
    """The scene with both bodies moved into the world at exposure time t (by hand: angle and position lerp)."""
    ang = A0 + (A1 - A0) * t if moving else A0
    pos = np.asarray(P0) + (np.asarray(P1) - np.asarray(P0)) * t if moving else np.asarray(P0)
    place = {0: (scenes.rot_axis(Z, ang), pos), 1: (scenes.rot_axis(B1[0], B1[1]), np.asarray(B1[2]))}
    kept = [p for p in parts[0] if drop is None or p[4] != drop]
    return scenes.posed_scene(kept, *parts[1:], place=place)


def _posed(parts, poses):
    sc = scenes.posed_scene(*parts)
    sc["poses"] = poses
    return sc


def _compare(a, b, rel_bar, close_bar=0.99):
    a = a[:, :, 0:3].astype(np.float64)
    b = b[:, :, 0:3].astype(np.float64)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    close = np.mean(np.all(np.abs(a - b) <= 1e-3 * np.maximum(b, 1e-2), axis=2))
    return rel, close


# ---------------------------------------------------------------------------- oracle pins (CPU)

def test_oracle_static_pose_is_the_placed_geometry(parts):
    cam = scenes.posed_camera(48, 36, spp=4)
    a, _, _ = oracle.render_tile(_posed(parts, _poses(False)), cam, 0, 0, 48, 36)
    b, _, _ = oracle.render_tile(_placed(parts), cam, 0, 0, 48, 36)
    rel, close = _compare(a, b, 1e-6)
    print("static pose vs placed: rel %.3g, close %.4f" % (rel, close))
    assert rel < 1e-6 and close > 0.999


@pytest.mark.parametrize("t", [0.25, 0.5, 1.0])
def test_oracle_moving_body_at_one_time(parts, t):
    """shutter [t, t]: every sub-sample sees slerp(q0, q1, t), lerp(p0, p1, t): the scene placed at the angle
    A0 + (A1 - A0) t about z (both ends share the axis, so slerp turns linearly in angle)."""
    cam = scenes.posed_camera(48, 36, spp=4, shutter=(t, t))
    a, _, _ = oracle.render_tile(_posed(parts, _poses(True)), cam, 0, 0, 48, 36)
    b, _, _ = oracle.render_tile(_placed(parts, t, moving=True), cam, 0, 0, 48, 36)
    rel, close = _compare(a, b, 1e-6)
    print("t=%.2f moving vs placed: rel %.3g, close %.4f" % (t, rel, close))
    assert rel < 1e-6 and close > 0.999


def test_oracle_hidden_body_is_absent(parts):
    cam = scenes.posed_camera(48, 36, spp=4)
    p = _poses(False)
    p[1] = None
    a, _, _ = oracle.render_tile(_posed(parts, p), cam, 0, 0, 48, 36)
    b, _, _ = oracle.render_tile(_placed(parts, drop=1), cam, 0, 0, 48, 36)
    rel, close = _compare(a, b, 1e-6)
    assert rel < 1e-6 and close > 0.999
    # both bodies are in view: taking either away changes more than 3 % of the pixels
    c, _, _ = oracle.render_tile(_placed(parts), cam, 0, 0, 48, 36)
    for drop in (0, 1):
        d, _, _ = oracle.render_tile(_placed(parts, drop=drop), cam, 0, 0, 48, 36)
        assert 1.0 - _compare(c, d, 0)[1] > 0.03


def test_oracle_motion_blur_lies_between_the_ends(parts):
    """Over the whole exposure the image is the time average: each pixel's blurred value sits within the range
    of the sharp images at the sampled times, in the mean over 4x4 blocks (a sanity bound, not a closed form)."""
    cam = scenes.posed_camera(32, 24, spp=16)
    blur, _, _ = oracle.render_tile(_posed(parts, _poses(True)), cam, 0, 0, 32, 24)
    ends = [oracle.render_tile(_placed(parts, t, moving=True), dict(cam, shutter=(t, t)), 0, 0, 32, 24)[0]
            for t in (0.0, 0.5, 1.0)]
    blk = lambda im: im[:, :, 0:3].astype(np.float64).reshape(6, 4, 8, 4, 3).mean(axis=(1, 3))
    lo = np.minimum.reduce([blk(e) for e in ends])
    hi = np.maximum.reduce([blk(e) for e in ends])
    b = blk(blur)
    assert np.mean((b >= lo - 0.05 * hi - 1e-3) & (b <= hi * 1.05 + 1e-3)) > 0.97
    assert np.linalg.norm(b - blk(ends[0])) > 0   # the body does move


# ---------------------------------------------------------------------------- GPU parity

@pytest.fixture(scope="module")
def ctx():
    from vtwin import lib
    c = lib.Context(0)
    yield c
    c.close()


def _gpu(ctx, sc, cam, poses, **kw):
    from vtwin import lib
    ctx.upload_scene(sc)
    ctx.set_instances(poses)
    ctx.set_camera(0, cam)
    ctx.render([0], **kw)
    img = ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"])
    st = ctx.stats()
    assert st["stack_overflows"] == 0
    return img, st


CASES = {
    "static": (False, None, (0.0, 1.0)),
    "moving": (True, None, (0.0, 1.0)),
    "moving_half": (True, None, (0.2, 0.7)),
    "hidden": (True, 1, (0.0, 1.0)),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("mode", ["persistent", "wavefront"])
def test_gpu_posed_bodies_match_oracle(ctx, parts, case, mode):
    from vtwin import lib
    moving, hide, shutter = CASES[case]
    poses = _poses(moving)
    if hide is not None:
        poses[hide] = None
    sc = _posed(parts, poses)
    cam = scenes.posed_camera(96, 72, spp=8, shutter=shutter)
    print("posed bodies: case %s, mode %s" % (case, mode), flush=True)
    m = lib.VT_RENDER_WAVEFRONT if mode == "wavefront" else lib.VT_RENDER_BATCHED
    img, st = _gpu(ctx, sc, cam, poses, mode=m)
    assert st["samples"] == 96 * 72 * 8
    ref, _, _ = oracle.render_tile(sc, cam, 0, 0, 96, 72)
    rel, close = _compare(img, ref, 1e-3)
    print("GPU vs oracle: relative RMSE %.3g, close %.4f" % (rel, close))
    assert rel < 1e-3 and close > 0.99


@pytest.mark.gpu
def test_gpu_set_instances_validates(ctx, parts):
    from vtwin import lib
    sc = _posed(parts, _poses(False))
    ctx.upload_scene(sc)
    with pytest.raises(lib.VTError):
        ctx.set_instances(_poses(False)[:1])          # one pose for two bodies
    bad = _poses(False)
    bad[0] = dict(q0=(0.0, 0.0, 0.0, 0.0), p0=(0, 0, 0))
    with pytest.raises(lib.VTError):
        ctx.set_instances(bad)
    cam = scenes.posed_camera(16, 12, spp=1, shutter=(0.0, 1.5))
    with pytest.raises(lib.VTError):
        ctx.set_camera(0, cam)


@pytest.mark.gpu
def test_gpu_material_adjoint_with_bodies(ctx, parts):
    """The path-replay adjoint replays the same posed paths: material and light gradients against the oracle's
    adjoint of the same samples (the oracle adjoint is pinned by finite differences elsewhere)."""
    from vtwin import lib
    poses = _poses(True)
    sc = _posed(parts, poses)
    cam = scenes.posed_camera(48, 36, spp=4)
    _gpu(ctx, sc, cam, poses)
    rng = np.random.default_rng(5)
    adj = rng.uniform(-1.0, 1.0, size=(36, 48, 4)).astype(np.float32)
    n_mat, n_light = sc["materials"].size // 16, sc["lights"].size // 4
    gm, gl = ctx.backward([0], [adj], n_mat, n_light)
    om, ol = oracle.backward_tile(sc, cam, 0, 0, 48, 36, adj)
    print("posed adjoint: materials max rel %.2e, lights max rel %.2e" % (
        np.max(np.abs(gm - om)) / np.max(np.abs(om)), np.max(np.abs(gl - ol)) / np.max(np.abs(ol))))
    assert np.allclose(gl, ol, rtol=1e-3, atol=1e-4 * np.max(np.abs(ol)))
    assert np.allclose(gm, om, rtol=1e-3, atol=1e-4 * np.max(np.abs(om)))


# ---------------------------------------------------------------------------- skip masks and visible()

def _brute_blocked(sc, c, pts, tol):
    """FP64 Moller-Trumbore over every triangle of a static scene: is the segment from c to tol short of each point
    blocked? (the definition visible() answers, with no BVH)."""
    T = np.asarray(sc["tris"], np.float32).reshape(-1, 16).astype(np.float64)
    v0, e1, e2 = T[:, 0:3], T[:, 4:7], T[:, 8:11]
    out = []
    for p in np.asarray(pts, np.float64):
        v = p - c
        dist = np.linalg.norm(v)
        if not dist > tol:
            out.append(False)
            continue
        d = v / dist
        pv = np.cross(d, e2)
        det = np.einsum("ij,ij->i", e1, pv)
        ok = np.abs(det) > 1e-30
        inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
        s = c - v0
        u = np.einsum("ij,ij->i", s, pv) * inv
        q = np.cross(s, e1)
        w = (q @ d) * inv
        t = np.einsum("ij,ij->i", e2, q) * inv
        hit = ok & (u >= 0) & (w >= 0) & (u + w <= 1) & (t > 0) & (t < dist - tol)
        out.append(bool(hit.any()))
    return np.array(out)


def _probe_points(n=400, seed=7):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(-1.6, 1.6, n), rng.uniform(-2.2, 1.2, n), rng.uniform(0.0, 1.2, n)])


@pytest.mark.parametrize("skip", [0, 1, 2])
def test_oracle_visible_is_brute_force(parts, skip):
    """visible() with moving bodies at time 0.5 and the camera's skip mask, against brute force over the scene placed
    at that time with the skipped bodies left out."""
    cam = dict(scenes.posed_camera(), skip_bodies=skip)
    pts = _probe_points()
    got = oracle.visible(_posed(parts, _poses(True)), cam, pts, tol=2e-4, time=0.5)
    drop = {0: None, 1: 0, 2: 1}[skip]
    ref = ~_brute_blocked(_placed(parts, 0.5, moving=True, drop=drop), np.asarray(cam["pos"], np.float64), pts, 2e-4)
    print("skip %d: %d of %d points visible, %d disagree" % (skip, got.sum(), got.size, np.sum(got != ref)))
    assert 0.2 < got.mean() < 0.95
    assert np.sum(got != ref) <= 2       # float32 scene data on grazing segments


def test_oracle_skip_masks_remove_the_body_from_those_rays(parts):
    """Depth 1 (camera rays and their shadow rays only): a body skipped by the camera and by the shadow rays is gone
    from the image."""
    cam = dict(scenes.posed_camera(48, 36, spp=4, depth=1), skip_bodies=1)
    sc = _posed(parts, _poses(False))
    sc["shadow_skip"] = 1
    a, _, _ = oracle.render_tile(sc, cam, 0, 0, 48, 36)
    b, _, _ = oracle.render_tile(_placed(parts, drop=0), cam, 0, 0, 48, 36)
    rel, close = _compare(a, b, 0)
    assert rel < 1e-6 and close > 0.999


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["persistent", "wavefront"])
def test_gpu_skip_masks_match_oracle(ctx, parts, mode):
    from vtwin import lib
    poses = _poses(True)
    sc = _posed(parts, poses)
    sc["shadow_skip"] = 2
    cam = dict(scenes.posed_camera(96, 72, spp=8), skip_bodies=1)
    print("skip masks: mode %s" % mode, flush=True)
    img, _ = _gpu(ctx, sc, cam, poses, mode=lib.VT_RENDER_WAVEFRONT if mode == "wavefront" else lib.VT_RENDER_BATCHED)
    ref, _, _ = oracle.render_tile(sc, cam, 0, 0, 96, 72)
    rel, close = _compare(img, ref, 1e-3)
    print("GPU vs oracle with skip masks: relative RMSE %.3g, close %.4f" % (rel, close))
    assert rel < 1e-3 and close > 0.99


@pytest.mark.gpu
@pytest.mark.parametrize("skip", [0, 1])
def test_gpu_visible_matches_oracle(ctx, parts, skip):
    poses = _poses(True)
    sc = _posed(parts, poses)
    cam = dict(scenes.posed_camera(), skip_bodies=skip)
    ctx.upload_scene(sc)
    ctx.set_instances(poses)
    ctx.set_camera(0, cam)
    pts = _probe_points(2000, seed=9)
    for t in (0.0, 0.5, 1.0):
        got = ctx.visible(0, pts, tol=2e-4, time=t)
        ref = oracle.visible(sc, cam, pts, tol=2e-4, time=t)
        print("visible t=%.1f skip %d: %d visible, %d disagree" % (t, skip, ref.sum(), np.sum(got != ref)))
        assert np.sum(got != ref) <= 4


# ---------------------------------------------------------------------------- a light that rides on a body
# The reference's light outlet rides on the scope: its emitter records and emission axis are in the scope's frame
# and every sample is posed with the scope at its exposure time (shade.wgsl.js:31-47, 55). Here body 0 carries a
# light with an angular profile about its +x axis (scenes.posed_light_parts).

@pytest.fixture(scope="module")
def light_parts():
    return scenes.posed_light_parts()


def _light_posed(light_parts, poses):
    sc = scenes.posed_scene(*light_parts, emit_body=0)
    sc["poses"] = poses
    sc["profiles"] = {0: scenes.BODY_LIGHT_PROFILE}
    return sc


def _light_placed(light_parts, t, moving=True):
    """The scene with both bodies, the light included, moved into the world at exposure time t, and the light's
    axis turned with body 0 by hand."""
    ang = A0 + (A1 - A0) * t if moving else A0
    pos = np.asarray(P0) + (np.asarray(P1) - np.asarray(P0)) * t if moving else np.asarray(P0)
    R0 = scenes.rot_axis(Z, ang)
    place = {0: (R0, pos), 1: (scenes.rot_axis(B1[0], B1[1]), np.asarray(B1[2]))}
    sc = scenes.posed_scene(*light_parts, place=place)
    axis, th, p = scenes.BODY_LIGHT_PROFILE
    sc["profiles"] = {0: (tuple(R0 @ np.asarray(axis)), th, p)}
    return sc


@pytest.mark.parametrize("t", [0.0, 0.6, 1.0])
def test_oracle_light_on_a_body_is_the_placed_light(light_parts, t):
    """shutter [t, t]: the light's samples, normals and profile axis posed with body 0 at time t render as the
    light placed in the world by hand at that time (static emitter records, rotated axis)."""
    cam = scenes.posed_camera(48, 36, spp=4, shutter=(t, t))
    a, _, _ = oracle.render_tile(_light_posed(light_parts, _poses(True)), cam, 0, 0, 48, 36)
    b, _, _ = oracle.render_tile(_light_placed(light_parts, t), cam, 0, 0, 48, 36)
    rel, close = _compare(a, b, 1e-6)
    print("t=%.2f light on a body vs placed: rel %.3g, close %.4f, mean %.4g" % (t, rel, close, b[..., :3].mean()))
    assert b[..., :3].mean() > 1e-3 and rel < 1e-6 and close > 0.999


def test_oracle_light_on_a_body_moves_the_lighting(light_parts):
    """The light turns with body 0 over the exposure: the images at the two ends differ, and the whole exposure
    averages them (block means between the ends)."""
    cam = scenes.posed_camera(32, 24, spp=16)
    sc = _light_posed(light_parts, _poses(True))
    blur, _, _ = oracle.render_tile(sc, cam, 0, 0, 32, 24)
    ends = [oracle.render_tile(sc, dict(cam, shutter=(t, t)), 0, 0, 32, 24)[0] for t in (0.0, 0.5, 1.0)]
    blk = lambda im: im[:, :, 0:3].astype(np.float64).reshape(6, 4, 8, 4, 3).mean(axis=(1, 3))
    lo = np.minimum.reduce([blk(e) for e in ends])
    hi = np.maximum.reduce([blk(e) for e in ends])
    assert np.linalg.norm(blk(ends[0]) - blk(ends[2])) > 0.1 * np.linalg.norm(blk(ends[0]))
    assert np.mean((blk(blur) >= lo - 0.05 * hi - 1e-3) & (blk(blur) <= hi * 1.05 + 1e-3)) > 0.97


def test_oracle_node_direct_with_a_light_on_a_body(light_parts):
    """nodeDirect samples the light at mid-exposure: the static nodes' B_dir under the posed light equals that
    under the light placed at t = 0.5."""
    sc = _light_posed(light_parts, _poses(True))
    sp = _light_placed(light_parts, 0.5)
    ni, nj = oracle.node_inputs(sc), oracle.node_inputs(sp)
    act = np.nonzero(ni["static"])[0][::7].astype(np.uint32)
    assert np.array_equal(ni["npos"][act], nj["npos"][act])
    alb, _ = oracle.albedo_tables(sc)
    st_a = np.zeros((ni["npos"].shape[0], 16), np.float32)
    st_b = np.zeros_like(st_a)
    oracle.node_direct(sc, ni, act, 4, alb, st_a)
    oracle.node_direct(sp, nj, act, 4, alb, st_b)
    B, Bo = st_a[act, 12:15].astype(np.float64), st_b[act, 12:15].astype(np.float64)
    print("node B_dir, light on a body vs placed: rel %.3g" % (np.linalg.norm(B - Bo) / np.linalg.norm(Bo)))
    assert np.linalg.norm(Bo) > 0 and np.linalg.norm(B - Bo) <= 1e-5 * np.linalg.norm(Bo)
    assert np.mean(st_a[act, 15] == st_b[act, 15]) > 0.999


def _gpu_light(ctx, sc, poses):
    ctx.upload_scene(sc)
    ctx.set_instances(poses)
    axis, th, p = sc["profiles"][0]
    ctx.set_light_profile(0, axis, th, p)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["persistent", "wavefront"])
@pytest.mark.parametrize("shutter", [(0.0, 1.0), (0.6, 0.6)])
def test_gpu_light_on_a_body_matches_oracle(ctx, light_parts, mode, shutter):
    from vtwin import lib
    poses = _poses(True)
    sc = _light_posed(light_parts, poses)
    cam = scenes.posed_camera(96, 72, spp=8, shutter=shutter)
    print("light on a body: mode %s, shutter %s" % (mode, shutter), flush=True)
    _gpu_light(ctx, sc, poses)
    ctx.set_camera(0, cam)
    ctx.render([0], mode=lib.VT_RENDER_WAVEFRONT if mode == "wavefront" else lib.VT_RENDER_BATCHED)
    img = ctx.readback(0, lib.VT_KIND_LINEAR, 96, 72)
    ref, _, _ = oracle.render_tile(sc, cam, 0, 0, 96, 72)
    rel, close = _compare(img, ref, 1e-3)
    print("GPU vs oracle: relative RMSE %.3g, close %.4f" % (rel, close))
    assert rel < 1e-3 and close > 0.99


@pytest.mark.gpu
def test_gpu_light_on_a_body_reference_estimator(ctx, light_parts):
    """The node cache and the reference estimator under a light on a moving body: node direct light against the
    oracle; the estimator (shadow culling toward the light's posed centre, fixed-camera caches) against the oracle
    fed the GPU's cache; a new pose of the light's body clears the occlusion cache (the next frame equals an
    uncached one) and makes the incremental node update a full one, while another body's move stays incremental."""
    from vtwin import lib
    poses = _poses(True)
    sc = _light_posed(light_parts, poses)
    _gpu_light(ctx, sc, poses)
    info = ctx.radiosity_build(64, 4, 16, 0.006, 0.003)
    V = np.asarray(sc["verts"]).size // 8
    act = ctx.radiosity_tables(2, info["active"], np.uint32)
    alb = ctx.radiosity_tables(5, (17, 32, 8), np.float32)
    hh = ctx.radiosity_tables(6, (17, 4), np.float32)
    assert ctx.radiosity_update_moved(0b01, iters=16) == "full"
    st, e = ctx.radiosity_state(V)
    ni = oracle.node_inputs(sc)
    so = np.zeros_like(st)
    oracle.node_direct(sc, ni, act, 4, alb, so)
    B, Bo = st[act, 12:15].astype(np.float64), so[act, 12:15].astype(np.float64)
    print("node B_dir under the light on a body: rel %.3g, codes equal %.5f" % (
        np.linalg.norm(B - Bo) / np.linalg.norm(Bo), np.mean(st[act, 15] == so[act, 15])))
    assert np.linalg.norm(B - Bo) <= 1e-3 * np.linalg.norm(Bo) and np.mean(st[act, 15] == so[act, 15]) > 0.995
    R = lib.VT_OPT_REFERENCE
    cam = scenes.posed_camera(96, 72, spp=4)
    sc2 = dict(sc, node_state=st, alb=alb, hh=hh, slot=1, e_avg=e)
    for opts in (R | lib.VT_OPT_CULL_SHADOWS, R | lib.VT_OPT_HIT_CACHE | lib.VT_OPT_SHADOW_CACHE):
        c = dict(cam, flags=opts)
        ctx.set_camera(0, c)
        ctx.render([0])
        img = ctx.readback(0, lib.VT_KIND_LINEAR, 96, 72)
        ref, _, _ = oracle.render_tile(sc2, c, 0, 0, 96, 72)
        rel, close = _compare(img, ref, 1e-3)
        print("reference estimator, options 0x%x: relative RMSE %.3g, close %.4f" % (opts, rel, close))
        assert rel < 1e-3 and close > 0.99
    # camera 0 keeps its caches: the next frame reads them; then the light's body turns further, so the kept
    # occlusion answers are stale and the frame must equal one from empty caches (a camera slot of its own)
    first = img
    ctx.render([0])
    assert np.array_equal(ctx.readback(0, lib.VT_KIND_LINEAR, 96, 72), first)
    moved = [dict(p) for p in poses]
    moved[0]["q0"] = scenes.quat_axis(Z, A0 + 0.4)
    moved[0]["q1"] = scenes.quat_axis(Z, A1 + 0.4)
    ctx.set_instances(moved)
    ctx.render([0])
    cached = ctx.readback(0, lib.VT_KIND_LINEAR, 96, 72)
    ctx.set_camera(1, dict(cam, flags=opts))
    ctx.render([1])
    fresh = ctx.readback(1, lib.VT_KIND_LINEAR, 96, 72)
    assert np.array_equal(cached, fresh) and not np.array_equal(cached, first)
    assert ctx.radiosity_update_moved(0b01, iters=0) == "full"
    moved[1]["p0"] = tuple(np.asarray(B1[2]) + (0.1, 0.0, 0.0))
    ctx.set_instances(moved)
    assert ctx.radiosity_update_moved(0b10, iters=0) == "incremental"


@pytest.mark.gpu
def test_gpu_light_on_a_body_validates(ctx, light_parts):
    """A light that rides on a body needs every emitter triangle in that body; a body light left static is
    refused, as is a light named on a body past the count."""
    from vtwin import lib
    sc = scenes.posed_scene(*light_parts)                    # emitter triangles in body 0, light static
    with pytest.raises(lib.VTError):
        ctx.upload_scene(sc)
    with pytest.raises(lib.VTError):
        ctx.upload_scene(dict(sc, emit_body=1))              # the light's triangles lie in body 0
    with pytest.raises(lib.VTError):
        ctx.upload_scene(dict(sc, emit_body=5))
    ctx.upload_scene(dict(sc, emit_body=0))
