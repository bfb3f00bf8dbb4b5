"""The path-replay adjoint: the FP64 oracle's gradients against central finite differences of its own forward
render (common random numbers), and the GPU's gradients against the oracle's on the same samples.

Loss = sum_p adj_p . I_p with a fixed adjoint image adj (the gradient of any image loss at the current image);
parameters: material colours (rho_d / F0, rgb), light radiances (rgb), and each camera's intrinsics and pose
(fx fy cx cy k1..k4, rotation, centre; interior term: image points, random numbers and triangles held fixed).
Bar: within 1e-3 relative of finite differences.
"""
import numpy as np
import pytest

import oracle
from vtwin import lib, scenes

TILE = (272, 176, 48, 32)      # x0, y0, w, h on the c1 camera (table top, dome, light region)


def _setup(spp=8):
    sc, cams = scenes.c1(bvh=lambda p, i: (lambda b: (b["nodes"], b["order"]))(oracle.build_bvh(p, i)))
    cam = dict(cams[0])
    cam["spp"] = spp
    r = scenes.Rng(99)
    x0, y0, w, h = TILE
    adj = np.zeros((h, w, 4), np.float32)
    adj[:, :, 0:3] = r.uniform(-1, 1, w * h * 3).reshape(h, w, 3)
    return sc, cam, adj


def _loss(sc, cam, adj, tile=TILE):
    x0, y0, w, h = tile
    img, _, _ = oracle.render_tile(sc, cam, x0, y0, w, h)
    return float(np.sum(img[:, :, 0:3].astype(np.float64) * adj[:, :, 0:3]))


def _setup_mixed(spp=8):
    """C1 with Lambert, GGX dielectric, GGX conductor and coated materials (scenes.c1_mixed), the oracle's BVH."""
    sc, cams = scenes.c1_mixed(bvh=lambda p, i: (lambda b: (b["nodes"], b["order"]))(oracle.build_bvh(p, i)))
    cam = dict(cams[0])
    cam["spp"] = spp
    r = scenes.Rng(99)
    x0, y0, w, h = TILE
    adj = np.zeros((h, w, 4), np.float32)
    adj[:, :, 0:3] = r.uniform(-1, 1, w * h * 3).reshape(h, w, 3)
    return sc, cam, adj


def _fd_sampling_fixed(sc, cam, adj, m, ch, hstep=1e-2, tile=TILE):
    """Central difference in material m's colour channel ch with the path tracer's sampling (lobe choice, BSDF
    pdf, MIS weights) held at the unperturbed materials: common random numbers take the same lobes and
    directions, so the estimate is a polynomial in the colour and the difference is the detached derivative."""
    base = np.asarray(sc["materials"], np.float32)
    out = []
    for sgn in (1, -1):
        a = base.reshape(-1, 16).copy()
        a[m, ch] += sgn * hstep
        out.append(_loss(dict(sc, materials=a.reshape(-1), sample_materials=base), cam, adj, tile))
    return (out[0] - out[1]) / (2 * hstep)


def _with(sc, key, idx, ch, delta):
    s2 = dict(sc)
    if key == "lights":
        a = sc["lights"].reshape(-1, 4).copy()
        a[idx, ch] += delta
    else:
        a = sc["materials"].reshape(-1, 16).copy()
        a[idx, ch] += delta
    s2[key] = a.reshape(-1)
    return s2


def test_oracle_adjoint_matches_finite_differences():
    sc, cam, adj = _setup()
    x0, y0, w, h = TILE
    gm, gl = oracle.backward_tile(sc, cam, x0, y0, w, h, adj)
    # lights: the image is linear in the radiance, so the central difference is exact up to rounding
    for li in range(4):
        for ch in range(3):
            hstep = 0.5
            fd = (_loss(_with(sc, "lights", li, ch, hstep), cam, adj) - _loss(_with(sc, "lights", li, ch, -hstep), cam, adj)) / (2 * hstep)
            assert gl[li, ch] == pytest.approx(fd, rel=1e-3, abs=1e-6 * max(1.0, abs(fd)))
    # Lambertian albedos: sampling does not depend on them, the estimate is a polynomial in each albedo
    order = np.argsort(-np.abs(gm[:16, 0]))
    for m in order[:6]:
        for ch in range(3):
            hstep = 1e-3
            fd = (_loss(_with(sc, "materials", m, ch, hstep), cam, adj) - _loss(_with(sc, "materials", m, ch, -hstep), cam, adj)) / (2 * hstep)
            assert gm[m, ch] == pytest.approx(fd, rel=1e-3, abs=2e-4 * np.max(np.abs(gm)))


KIND_NAMES = {scenes.KIND_LAMBERT: "lambert", scenes.KIND_DIELECTRIC: "dielectric rho_d",
              scenes.KIND_CONDUCTOR: "conductor F0", scenes.KIND_COATED: "coated F0"}


def test_oracle_ggx_adjoint_matches_finite_differences():
    """GGX dielectric (rho_d, whose specular probability moves with it, gpu/bsdf.js:145-148), conductor and coated
    (F0 inside Schlick) colour gradients of the FP64 path replay against central differences with the sampling
    held fixed: the two largest components of every kind within 1e-3 relative."""
    sc, cam, adj = _setup_mixed()
    x0, y0, w, h = TILE
    gm, _ = oracle.backward_tile(sc, cam, x0, y0, w, h, adj)
    kinds = np.asarray(sc["materials"], np.float32).reshape(-1, 16)[:16, 5].astype(int)
    gmax = np.max(np.abs(gm))
    for kind, name in KIND_NAMES.items():
        rows = np.nonzero(kinds == kind)[0]
        flat = [(abs(gm[m, ch]), m, ch) for m in rows for ch in range(3)]
        for _, m, ch in sorted(flat, reverse=True)[:2]:
            fd = _fd_sampling_fixed(sc, cam, adj, m, ch)
            print("%s: material %d channel %d adjoint %.6g, finite difference %.6g" % (name, m, ch, gm[m, ch], fd))
            assert gm[m, ch] == pytest.approx(fd, rel=1e-3, abs=2e-5 * gmax)


def test_oracle_sampling_table_changes_only_the_sampling():
    """The sampling table equal to the materials leaves the image bit-identical; a different one leaves the
    estimate unbiased (same mean within noise) but changes the samples."""
    sc, cam, adj = _setup_mixed(spp=4)
    x0, y0, w, h = TILE
    a, _, _ = oracle.render_tile(sc, cam, x0, y0, w, h)
    b, _, _ = oracle.render_tile(dict(sc, sample_materials=sc["materials"]), cam, x0, y0, w, h)
    assert np.array_equal(a, b)
    other = np.asarray(sc["materials"], np.float32).reshape(-1, 16).copy()
    other[:16, 0:3] = 0.5
    c, _, _ = oracle.render_tile(dict(sc, sample_materials=other.reshape(-1)), cam, x0, y0, w, h)
    assert not np.array_equal(a, c)


@pytest.mark.gpu
def test_gpu_ggx_adjoint_matches_oracle_and_fd():
    """Every material kind on the GPU: the FP32 path replay against the FP64 oracle's on the same samples, and the
    largest GGX components against the sampling-fixed finite differences."""
    sc, cam, adj = _setup_mixed()
    x0, y0, w, h = TILE
    full = np.zeros((cam["height"], cam["width"], 4), np.float32)
    full[y0:y0 + h, x0:x0 + w] = adj
    n_mat, n_light = sc["materials"].size // 16, sc["lights"].size // 4
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(sc)
        ctx.set_camera(0, cam)
        gm, gl = ctx.backward([0], [full], n_mat, n_light)
        st = ctx.stats()
    finally:
        ctx.close()
    om, ol = oracle.backward_tile(sc, cam, x0, y0, w, h, adj)
    print("mixed materials, GPU vs FP64 oracle: materials max rel %.2e, lights max rel %.2e" % (
        np.max(np.abs(gm - om)) / np.max(np.abs(om)), np.max(np.abs(gl - ol)) / np.max(np.abs(ol))))
    assert np.allclose(gm, om, rtol=1e-3, atol=1e-4 * np.max(np.abs(om)))
    assert np.allclose(gl, ol, rtol=1e-3, atol=1e-4 * np.max(np.abs(ol)))
    assert st["stack_overflows"] == 0
    kinds = np.asarray(sc["materials"], np.float32).reshape(-1, 16)[:16, 5].astype(int)
    for kind in (scenes.KIND_DIELECTRIC, scenes.KIND_CONDUCTOR, scenes.KIND_COATED):
        rows = np.nonzero(kinds == kind)[0]
        m, ch = max(((m, ch) for m in rows for ch in range(3)), key=lambda q: abs(om[q[0], q[1]]))
        fd = _fd_sampling_fixed(sc, cam, adj, m, ch)
        assert gm[m, ch] == pytest.approx(fd, rel=1e-3, abs=2e-5 * np.max(np.abs(om))), KIND_NAMES[kind]


def _gpu_tile_loss(ctx, cam, adj, tile, materials=None, lights=None):
    """sum adj . I over a tile of the GPU's forward frame (persistent form) after replacing the materials or
    lights; the sample set is the camera's, so two calls differ only by the parameter change"""
    x0, y0, w, h = tile
    if materials is not None:
        ctx.update_materials(materials)
    if lights is not None:
        ctx.update_lights(lights)
    ctx.render([0])
    img = ctx.readback(0, lib.VT_KIND_LINEAR, cam["width"], cam["height"])
    return float(np.sum(img[y0:y0 + h, x0:x0 + w, :3].astype(np.float64) * adj[:, :, :3]))


@pytest.mark.gpu
def test_gpu_adjoint_c5_tile():
    """The differentiation configuration's scene and cameras (2,013,432 triangles, 256 Lambert / GGX dielectric /
    GGX conductor materials, 16 lights, KB4 1280x960, 32 spp, depth 5), one 48x32 tile with a random-sign adjoint.

    * GPU adjoint against central differences of the GPU's own forward frames (same FP32 paths, common random
      numbers): the largest Lambert albedo and conductor F0 components (their sampling does not depend on them)
      and the largest light radiances, within 1e-3.
    * The oracle's FP64 adjoint against its sampling-fixed central differences for the largest GGX dielectric
      components, within 1e-3 (rho_d moves the lobe choice, so the GPU's forward cannot hold it).
    * GPU against the oracle on the same samples: material and light gradients within 5e-3 of their norm, the 14
      camera parameters within 5e-2 (the oracle's detached-sampling central differences, no pair left out). The
      difference is paths that meet other triangles in FP32 than in FP64 on this scene's 2 M small triangles over
      5 vertices: with random-sign adjoints a few such paths are what the gradient sums show (2e-3 of the largest
      material component measured, 2.8e-2 of the camera gradient's norm, whose per-sample terms are heavy-tailed
      where deep vertices graze)."""
    sc, cams = scenes.c5()
    cam = cams[0]
    tile = x0, y0, w, h = 608, 448, 48, 32
    r = scenes.Rng(98)
    adj = np.zeros((h, w, 4), np.float32)
    adj[:, :, 0:3] = r.uniform(-1, 1, w * h * 3).reshape(h, w, 3)
    full = np.zeros((cam["height"], cam["width"], 4), np.float32)
    full[y0:y0 + h, x0:x0 + w] = adj
    mats = np.asarray(sc["materials"], np.float32).reshape(-1, 16)
    lights = np.asarray(sc["lights"], np.float32).reshape(-1, 4)
    n_mat, n_light = mats.shape[0], lights.shape[0]
    kinds = mats[:256, 5].astype(int)
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(sc)
        ctx.set_camera(0, cam)
        gm, gl, gc = ctx.backward([0], [full], n_mat, n_light, cameras=True)
        assert ctx.stats()["stack_overflows"] == 0
        checks = []
        for kind in (scenes.KIND_LAMBERT, scenes.KIND_CONDUCTOR):
            rows = np.nonzero(kinds == kind)[0]
            flat = sorted(((abs(gm[m, ch]), m, ch) for m in rows for ch in range(3)), reverse=True)[:2]
            for _, m, ch in flat:
                hs = 1e-2
                mp, mm = mats.copy(), mats.copy()
                mp[m, ch] += hs
                mm[m, ch] -= hs
                fd = (_gpu_tile_loss(ctx, cam, adj, tile, materials=mp) - _gpu_tile_loss(ctx, cam, adj, tile, materials=mm)) / (2 * hs)
                checks.append((KIND_NAMES[kind], m, ch, gm[m, ch], fd))
        ctx.update_materials(mats)
        for _, li, ch in sorted(((abs(gl[li, ch]), li, ch) for li in range(n_light) for ch in range(3)), reverse=True)[:2]:
            hs = 0.5
            lp, lm = lights.copy(), lights.copy()
            lp[li, ch] += hs
            lm[li, ch] -= hs
            fd = (_gpu_tile_loss(ctx, cam, adj, tile, lights=lp) - _gpu_tile_loss(ctx, cam, adj, tile, lights=lm)) / (2 * hs)
            checks.append(("light", li, ch, gl[li, ch], fd))
    finally:
        ctx.close()
    for name, m, ch, g, fd in checks:
        print("c5 GPU %s %d channel %d: adjoint %.6g, GPU finite difference %.6g" % (name, m, ch, g, fd))
        assert g == pytest.approx(fd, rel=1e-3, abs=1e-5 * np.max(np.abs(gm))), (name, m, ch)
    om, ol = oracle.backward_tile(sc, cam, x0, y0, w, h, adj, nthreads=16)
    rows = np.nonzero(kinds == scenes.KIND_DIELECTRIC)[0]
    for _, m, ch in sorted(((abs(om[m, ch]), m, ch) for m in rows for ch in range(3)), reverse=True)[:2]:
        fd = _fd_sampling_fixed(sc, cam, adj, m, ch, tile=tile)
        print("c5 oracle dielectric rho_d %d channel %d: adjoint %.6g, finite difference %.6g (GPU %.6g)" % (
            m, ch, om[m, ch], fd, gm[m, ch]))
        assert om[m, ch] == pytest.approx(fd, rel=1e-3, abs=1e-5 * np.max(np.abs(om)))
    rel_m = np.linalg.norm(gm - om) / np.linalg.norm(om)
    rel_l = np.linalg.norm(gl - ol) / np.linalg.norm(ol)
    print("c5 tile, GPU vs FP64 oracle: materials %.2e, lights %.2e of the norm; largest component %.2e of the max" % (
        rel_m, rel_l, np.max(np.abs(gm - om)) / np.max(np.abs(om))))
    assert rel_m < 5e-3 and rel_l < 5e-3
    # camera parameters: the oracle's per-sample central differences at a relative step of 1e-8 (converged to 1e-5
    # of the largest component between 1e-8 and 3e-9 on this tile; 1e-6 is not yet linear where paths graze)
    want, changed = oracle.camera_grad_tile(sc, cam, x0, y0, w, h, adj, rel_step=1e-8, nthreads=16)
    scale = np.array([cam["lens"][0]] * 4 + [1.0] * 10)
    a, b = gc[0] * scale, want * scale
    rel_c = np.linalg.norm(a - b) / np.linalg.norm(b)
    print("GPU   ", np.array2string(a, precision=5), "\noracle", np.array2string(b, precision=5), "\nchanged", changed,
          "camera gradient %.2e of the norm" % rel_c)
    assert changed == 0
    # FP32 and FP64 paths that meet other triangles at deep vertices carry their own (large, near-grazing) vertex
    # motions: 2.8e-2 of the norm measured; the C1 tile's camera gradients agree to 1e-3 (test above)
    assert rel_c < 5e-2


@pytest.mark.gpu
def test_gpu_adjoint_matches_oracle_and_fd():
    sc, cam, adj = _setup()
    x0, y0, w, h = TILE
    full = np.zeros((cam["height"], cam["width"], 4), np.float32)
    full[y0:y0 + h, x0:x0 + w] = adj
    ctx = lib.Context(0)
    ctx.upload_scene(sc)
    ctx.set_camera(0, cam)
    n_mat, n_light = sc["materials"].size // 16, sc["lights"].size // 4
    gm, gl = ctx.backward([0], [full], n_mat, n_light)
    st = ctx.stats()
    ctx.close()
    om, ol = oracle.backward_tile(sc, cam, x0, y0, w, h, adj)
    print("GPU vs FP64 oracle gradients: materials max rel %.2e, lights max rel %.2e, backward %.2f ms" % (
        np.max(np.abs(gm - om)) / np.max(np.abs(om)), np.max(np.abs(gl - ol)) / np.max(np.abs(ol)), st["ms_backward"]))
    assert np.allclose(gl, ol, rtol=1e-3, atol=1e-4 * np.max(np.abs(ol)))
    assert np.allclose(gm, om, rtol=1e-3, atol=1e-4 * np.max(np.abs(om)))
    assert st["stack_overflows"] == 0 and st["rays_backward"] > 0
    # and the oracle's finite differences on the two largest albedo components
    big = np.dstack(np.unravel_index(np.argsort(-np.abs(om[:16]).ravel())[:2], (16, 3)))[0]
    for m, ch in big:
        fd = (_loss(_with(sc, "materials", m, ch, 1e-3), cam, adj) - _loss(_with(sc, "materials", m, ch, -1e-3), cam, adj)) / 2e-3
        assert gm[m, ch] == pytest.approx(fd, rel=1e-3)


def _cam_grads_gpu(sc, cam, adj_full, n_mat, n_light):
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(sc)
        ctx.set_camera(0, cam)
        gm, gl, gc = ctx.backward([0], [adj_full], n_mat, n_light, cameras=True)
        gm2, gl2 = ctx.backward([0], [adj_full], n_mat, n_light)
        st = ctx.stats()
    finally:
        ctx.close()
    return gm, gl, gc[0], gm2, gl2, st


@pytest.mark.gpu
def test_gpu_camera_gradients_match_oracle():
    """Camera intrinsics and pose: the GPU's forward-mode tangents (FP32) against the oracle's FP64 per-sample
    central differences on the same samples (interior term), on natural scales (pixels, radians, metres)."""
    sc, cam, adj = _setup()
    x0, y0, w, h = TILE
    full = np.zeros((cam["height"], cam["width"], 4), np.float32)
    full[y0:y0 + h, x0:x0 + w] = adj
    n_mat, n_light = sc["materials"].size // 16, sc["lights"].size // 4
    gm, gl, gc, gm2, gl2, st = _cam_grads_gpu(sc, cam, full, n_mat, n_light)
    want, changed = oracle.camera_grad_tile(sc, cam, x0, y0, w, h, adj)
    scale = np.array([cam["lens"][0]] * 4 + [1.0] * 10)
    a, b = gc * scale, want * scale
    print("GPU   ", np.array2string(a, precision=5), "\noracle", np.array2string(b, precision=5), "\nchanged pairs", changed)
    assert np.all(np.abs(a - b) <= 1e-3 * np.abs(b) + 2e-4 * np.abs(b).max())
    # the material and light gradients of the same call stay those of the oracle
    om, ol = oracle.backward_tile(sc, cam, x0, y0, w, h, adj)
    for g_, o_ in ((gm, om), (gl, ol), (gm2, om), (gl2, ol)):
        assert np.allclose(g_, o_, rtol=1e-3, atol=1e-4 * np.max(np.abs(o_)))
    assert st["stack_overflows"] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 0])
def test_gpu_camera_gradients_closed_form(mode):
    """The plane under a square emitter: the GPU camera gradient against the closed form's derivative."""
    from test_oracle_pins import _plane_and_light, _plane_image
    sc, _ = _plane_and_light()
    pos = (0.15, 0.05, 0.5)
    if mode == 1:
        lens, tile = [400, 400, 31.5, 31.5, 0, 0, 0, 0, 0], (24, 28, 8, 8)
    else:
        lens, tile = [100, 100, 31.5, 31.5, 0.03, -0.02, 0.01, -0.005, 0], (50, 40, 8, 8)
    cam = scenes.camera(scenes.look_at(pos, (0.15, 0.05, 0.0), up=(0, 1, 0)), pos, lens, 64, 64, 256, 0, 1, mode=mode)
    x0, y0, w, h = tile
    full = np.zeros((64, 64, 4), np.float32)
    full[y0:y0 + h, x0:x0 + w, 0:3] = 1.0
    _, _, gc, _, _, _ = _cam_grads_gpu(sc, cam, full, 2, 1)
    verts = np.array([[-0.2, -0.2, 1.0], [0.2, -0.2, 1.0], [0.2, 0.2, 1.0], [-0.2, 0.2, 1.0]])
    ys, xs = np.mgrid[y0:y0 + h, x0:x0 + w].astype(float)
    xs, ys = xs.ravel(), ys.ravel()
    th0 = np.zeros(14)
    th0[0:8] = np.asarray(lens[0:8], float)
    th0[11:14] = pos
    want = np.zeros(14)
    for k in range(14):
        hk = 1e-6 * (lens[0] if k < 4 else 1.0)
        tp, tm = th0.copy(), th0.copy()
        tp[k] += hk
        tm[k] -= hk
        want[k] = 3 * (_plane_image(tp, cam, xs, ys, verts, 0.5, 10.0).sum() - _plane_image(tm, cam, xs, ys, verts, 0.5, 10.0).sum()) / (2 * hk)
    scale = np.array([lens[0]] * 4 + [1.0] * 10)
    a, b = gc * scale, want * scale
    print("GPU   ", np.array2string(a, precision=4), "\nclosed", np.array2string(b, precision=4))
    assert np.all(np.abs(a - b) <= 0.03 * np.abs(b) + 0.005 * np.abs(b).max())
