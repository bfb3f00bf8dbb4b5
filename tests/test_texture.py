"""The printed board's texture (the reference's charuco_print: gpu/scene_gpu.js:180-192 and 324-325,
gpu/renderer.js:293 and 395-412, gpu/wgsl/shade.wgsl.js:258-265): textured meshes mix rho_d and the GGX alpha between
the ink (material words 8-11) and the paper by a print mask sampled bilinearly at the hit's texture coordinates.

The oracle's sampler (or_board_mask) is pinned against scipy's map_coordinates (order 1, edge-clamped), which fixes
the texel-centre convention, the axis order and the clamping independently; the mix is pinned by its ends (an all-
paper mask renders as the paper material, an all-ink mask as the ink) and the coordinates' orientation by a mask
whose top half is ink. Parity unpinned: the reference's filter weights come from the browser's WebGPU sampler,
whose sub-texel precision is implementation-defined (often 8 bits); this build samples with binary32 weights on
host and device alike.
"""
import os

import numpy as np
import pytest

import oracle
from vtwin import lib, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASSETS = os.path.join(ROOT, "twin_assets")


def test_board_mask_matches_scipy_bilinear():
    from scipy.ndimage import map_coordinates
    m = scenes.board_mask_pattern(37, 23, seed=5)
    r = np.random.default_rng(3)
    u = r.uniform(-0.2, 1.2, 400).astype(np.float32)
    v = r.uniform(-0.2, 1.2, 400).astype(np.float32)
    got = oracle.board_mask(m, u, v)
    # texel (i, j) centred at ((i + 0.5) / w, (j + 0.5) / h): array coordinates (row v h - 0.5, column u w - 0.5)
    rows = v.astype(np.float64) * m.shape[0] - 0.5
    cols = u.astype(np.float64) * m.shape[1] - 0.5
    ref = map_coordinates(m.astype(np.float64) / 255.0, [rows, cols], order=1, mode="nearest")
    assert np.max(np.abs(got - ref)) < 2e-6


def test_board_mask_closed_forms():
    w, h = 16, 8
    flat = np.full((h, w), 77, np.uint8)
    assert np.allclose(oracle.board_mask(flat, [-3.0, 0.0, 0.31, 1.0, 5.0], [0.5, -1.0, 0.9, 2.0, 0.1]), 77 / 255, atol=1e-7)
    ramp = np.tile((np.arange(w) * 9).astype(np.uint8), (h, 1))
    u = np.linspace(0.5 / w, 1 - 0.5 / w, 50).astype(np.float32)
    lin = (u.astype(np.float64) * w - 0.5) * 9 / 255
    assert np.max(np.abs(oracle.board_mask(ramp, u, np.full_like(u, 0.3)) - lin)) < 2e-6
    # clamp to the edge texels outside the outermost centres
    assert oracle.board_mask(ramp, [0.01], [0.5])[0] == pytest.approx(0.0, abs=1e-7)
    assert oracle.board_mask(ramp, [0.999], [0.5])[0] == pytest.approx(135 / 255, abs=1e-6)
    # texel centres give the texels themselves
    i, j = np.meshgrid(np.arange(w), np.arange(h))
    rnd = np.random.default_rng(1).integers(0, 256, (h, w)).astype(np.uint8)
    got = oracle.board_mask(rnd, ((i + 0.5) / w).astype(np.float32).ravel(), ((j + 0.5) / h).astype(np.float32).ravel())
    assert np.max(np.abs(got - rnd.ravel() / 255.0)) < 2e-6


def _poses():
    # body 0 (its board on the -y face toward the camera) beside the table, body 1 out of the way
    return [dict(q0=(0.0, 0.0, 0.0, 1.0), p0=(-0.55, -1.45, 1.0)), dict(q0=(0.0, 0.0, 0.0, 1.0), p0=(0.9, 0.6, 0.6))]


def _oracle_scene(sc, poses):
    s = dict(sc)
    s["poses"] = poses
    return s


def _untextured(sc, rec):
    """The textured scene with its printed meshes flat on material record `rec`."""
    s = dict(sc)
    meshes = sc["meshes"].reshape(-1, 4).copy()
    mats = sc["materials"].reshape(-1, 16).copy()
    for k in np.nonzero(meshes[:, 1] & scenes.MESH_TEXTURED)[0]:
        mats[meshes[k, 0]] = rec
        meshes[k, 1] &= ~np.uint32(scenes.MESH_TEXTURED)
    s["meshes"], s["materials"] = meshes.reshape(-1), mats.reshape(-1)
    tris = sc["tris"].reshape(-1, 16).copy()
    tu = tris.view(np.uint32)
    tu[:, 13] = meshes[tu[:, 12], 1]
    s["tris"] = tris.reshape(-1)
    return s


def test_oracle_mix_ends_and_orientation():
    """An all-paper mask renders the boards as the paper material, an all-ink mask as the ink (rho_d and alpha of
    words 8-11), a mask with an inked top half darkens the upper half of the static board."""
    cam = scenes.posed_camera(64, 48, spp=2, depth=2)
    poses = _poses()
    base = scenes.textured_scene()
    rec = scenes.printed_material()
    paper = rec.copy()
    ink = rec.copy()
    ink[0:3], ink[3] = rec[8:11], rec[11]
    for fill, flat in ((255, paper), (0, ink)):
        sc = scenes.textured_scene(mask=np.full((9, 13), fill, np.uint8))
        a, _, _ = oracle.render_tile(_oracle_scene(sc, poses), cam, 0, 0, 64, 48)
        b, _, _ = oracle.render_tile(_oracle_scene(_untextured(sc, flat), poses), cam, 0, 0, 64, 48)
        assert np.max(np.abs(a[..., :3] - b[..., :3])) <= 1e-6 * max(1.0, float(np.max(b[..., :3])))
    # orientation: v = 0 is the mask's top row
    top_ink = np.full((8, 8), 255, np.uint8)
    top_ink[:4] = 0
    sc = scenes.textured_scene(mask=top_ink)
    img, _, _ = oracle.render_tile(_oracle_scene(sc, poses), cam, 0, 0, 64, 48)
    lum = img[..., :3].mean(axis=2)
    upper, lower = (lum[_pixel(cam, p)] for p in ((0.35, -1.12, 1.15), (0.35, -1.12, 0.75)))
    assert upper < lower
    assert base["board_mask"].shape == (72, 96)


def _pixel(cam, p):
    """(row, column) of world point p in camera cam (the oracle's KB4 projection, gpu/lens.js:43-51)."""
    import ctypes as C
    R = np.asarray(cam["R"], np.float64).reshape(3, 3)           # rows: the camera's right, down, forward axes
    X = (R @ (np.asarray(p, np.float64) - np.asarray(cam["pos"], np.float64))).astype(np.float64)
    L = cam["lens"]
    lz = oracle.lens(L[0], L[1], L[2], L[3], L[4:8], L[8])
    uv = np.zeros(2)
    oracle.lib().or_project(C.byref(lz), oracle.ptr(X), oracle.ptr(uv))
    return int(uv[1]), int(uv[0])


def _gpu_frame(sc, cam, poses, mode):
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(sc)
        ctx.set_instances(poses)
        ctx.set_camera(0, cam)
        ctx.render([0], mode=mode)
        img = ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"])
        st = ctx.stats()
    finally:
        ctx.close()
    return img, st


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["wavefront", "batched"])
def test_gpu_textured_boards_match_oracle(mode):
    """Both printed boards (static, and on a posed body) under the test mask: GPU frame against the oracle on the
    same samples, in both forms."""
    sc = scenes.textured_scene()
    cam = scenes.posed_camera(96, 72, spp=8, depth=3)
    poses = _poses()
    print("textured boards: mode %s" % mode, flush=True)
    img, st = _gpu_frame(sc, cam, poses, lib.VT_RENDER_WAVEFRONT if mode == "wavefront" else lib.VT_RENDER_BATCHED)
    assert st["stack_overflows"] == 0
    ref, _, _ = oracle.render_tile(_oracle_scene(sc, poses), cam, 0, 0, 96, 72)
    a, b = img[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    close = np.mean(np.abs(a - b) <= 1e-3 * np.maximum(np.abs(b), 1e-3))
    print("GPU vs oracle: relative RMSE %.3g, close %.4f" % (rel, close))
    assert rel < 1e-3 and close > 0.99
    # the mask shows: the same frame with an all-paper mask differs on the boards
    flat = scenes.textured_scene(mask=np.full((4, 4), 255, np.uint8))
    img2, _ = _gpu_frame(flat, cam, poses, lib.VT_RENDER_BATCHED)
    assert np.mean(np.abs(img2[..., :3] - img[..., :3]) > 1e-3) > 0.01


@pytest.mark.gpu
def test_gpu_texture_entry_points_validate():
    sc = scenes.textured_scene()
    ctx = lib.Context(0)
    try:
        with pytest.raises(lib.VTError):
            ctx.set_board_texture(np.zeros((4, 4), np.uint8))      # no scene yet
        ctx.upload_scene(sc)
        with pytest.raises(lib.VTError):
            ctx.set_board_texture(np.zeros((0, 4), np.uint8))
        bad = dict(sc)
        bad["uvs"] = None
        with pytest.raises(lib.VTError):
            ctx.upload_scene(bad)                                  # a textured mesh without coordinates
    finally:
        ctx.close()


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(os.path.join(ASSETS, "textures", "charuco_b.png")),
                    reason="twin_assets/textures missing (python tests/golden/fetch_twin_assets.py)")
def test_gpu_twin_charuco_board_matches_oracle():
    """The real ChArUco board (its mesh's texture coordinates, the printed mask, physics' charuco_print paper and
    ink) posed 0.14 m in front of CAM3, facing it: a tile at the image centre against the oracle, where ink and
    paper both show."""
    from vtwin import twin
    scene, meshes, physics = twin.load(ASSETS)
    prep = twin.prepare(scene, meshes, physics)
    assert prep.get("board_mask") is not None and prep["board_mask"].shape == (1440, 1920)
    cams = twin.cameras(scene, spp=4, depth=3, seed=7, prep=prep)
    cam = cams[2]
    poses = twin.poses_from_state(scene, twin.default_state(scene))
    R = np.asarray(cam["R"], np.float64).reshape(3, 3).T          # columns: camera right, down, forward (world)
    b = scene["boards"][0]
    ctr = np.array([(b["size_m"][0]) / 2 - b["margin_m"], b["size_m"][1] / 2 - b["margin_m"], 0.0])
    T = np.eye(4)
    T[:3, :3] = R                                                 # board z (into the board) along the view
    T[:3, 3] = np.asarray(cam["pos"], np.float64) + 0.14 * R[:, 2] - R @ ctr
    poses["board/" + b["name"]] = T
    inst = twin.instances(prep, poses)
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(prep)
        ctx.upload_profiles(prep)
        ctx.set_instances(inst)
        ctx.set_camera(0, cam)
        ctx.render([0], mode=lib.VT_RENDER_BATCHED)
        img = ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"])
    finally:
        ctx.close()
    x0, y0, w, h = int(cam["lens"][2]) - 24, int(cam["lens"][3]) - 16, 48, 32
    sc = dict(prep)
    sc["poses"] = inst
    ref, _, _ = oracle.render_tile(sc, cam, x0, y0, w, h)
    a = img[y0:y0 + h, x0:x0 + w, :3].astype(np.float64)
    r = ref[:, :, :3].astype(np.float64)
    rel = np.linalg.norm(a - r) / np.linalg.norm(r)
    lum = r.mean(axis=2)
    print("twin board tile: relative RMSE %.3g, luminance %.3g..%.3g" % (rel, lum.min(), lum.max()), flush=True)
    assert rel < 1e-3
    assert lum.max() > 4 * max(lum.min(), 1e-9)                   # ink and paper both in the tile
