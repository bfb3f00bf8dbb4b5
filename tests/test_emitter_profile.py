"""Angular emission profiles (physics.yaml light.profile; gpu/wgsl/shade.wgsl.js:12-27 profile, :48-52 emitted):
L_e(w) = L0 p(theta), theta between the outgoing direction and the light's axis, p linear between knots and 0 past
the last one.

Oracle pins (CPU): the physics profile's projected solid angle (the value physics.yaml derives from the same
table); NEE under a profiled square emitter against a 2-D quadrature of L0 p(theta) cos^2 / r^2 with a tilted axis;
a constant profile c scales the whole image by c. GPU parity: images, the light/material adjoint and the camera
tangents against the oracle with the physics profile on the c1 lights.
"""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle
from vtwin import scenes

PHYS_DEG = [0, 5, 10, 15, 20, 25, 30, 35, 40, 45, 50, 55, 60]
PHYS_P = [1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.933, 0.75, 0.5, 0.25, 0.067, 0.0, 0.0]
PHYS_TH = [math.radians(t) for t in PHYS_DEG]


def _prof(th, knots=PHYS_TH, p=PHYS_P):
    kn = np.ascontiguousarray(knots, np.float64)
    pp = np.ascontiguousarray(p, np.float64)
    return oracle.lib().or_profile(float(th), oracle.ptr(kn), oracle.ptr(pp), len(kn))


def test_projected_solid_angle_of_the_physics_profile():
    """2 pi int p(theta) cos sin dtheta = 1.305712 sr (physics.yaml light.profile.projected_solid_angle)."""
    v = integrate.quad(lambda t: _prof(t) * math.cos(t) * math.sin(t), 0, math.pi / 2, points=PHYS_TH[1:12],
                       limit=200, epsabs=1e-13)[0]
    assert 2 * math.pi * v == pytest.approx(1.305712, abs=2e-6)


def _square_profiled_irradiance(Le, a, h, axis, knots, p, n=1200):
    """midpoint rule over the square (n x n cells) of Le p(theta) cos^2 / r^2, p by numpy's own linear
    interpolation (0 past the last knot)"""
    ax = np.asarray(axis, float) / np.linalg.norm(axis)
    c = (np.arange(n) + 0.5) / n * 2 * a - a
    x, y = np.meshgrid(c, c, indexing="ij")
    r2 = x * x + y * y + h * h
    w = np.stack([x, y, np.full_like(x, -h)], -1) / np.sqrt(r2)[..., None]
    th = np.arccos(np.clip(w @ ax, -1.0, 1.0))
    pr = np.interp(th, knots, p, right=0.0)
    return float(np.sum(Le * pr * h * h / r2 ** 2) * (2 * a / n) ** 2)


@pytest.mark.parametrize("axis", [(0.0, 0.0, -1.0), (0.12, -0.05, -1.0)])
def test_oracle_profiled_nee_against_quadrature(axis):
    from test_oracle_pins import _plane_and_light
    sc, cam = _plane_and_light()
    knots, p = [0.0, 0.08, 0.16, 0.24], [1.0, 0.8, 0.3, 0.0]
    sc["profiles"] = {0: (axis, knots, p)}
    img, _, _ = oracle.render_tile(sc, cam, 30, 30, 4, 4, 0, 2048)
    want = 0.5 / math.pi * _square_profiled_irradiance(10.0, 0.2, 1.0, axis, knots, p)
    flat = 0.5 / math.pi * _square_profiled_irradiance(10.0, 0.2, 1.0, axis, [0.0, math.pi], [1.0, 1.0])
    got = img[:, :, 0].mean()
    print("profiled NEE %.6f, quadrature %.6f (Lambertian %.6f)" % (got, want, flat))
    assert want < 0.8 * flat                     # the profile matters here
    assert got == pytest.approx(want, rel=0.005)


def test_oracle_constant_profile_scales_the_image():
    sc, cams = scenes.c1(bvh=lambda p_, i: (lambda b: (b["nodes"], b["order"]))(oracle.build_bvh(p_, i)))
    cam = dict(cams[0], spp=4)
    ref, _, _ = oracle.render_tile(sc, cam, 272, 176, 32, 24)
    sc2 = dict(sc, profiles={k: ((0.3, 0.2, -1.0), [0.0, math.pi], [0.37, 0.37]) for k in range(4)})
    img, _, _ = oracle.render_tile(sc2, cam, 272, 176, 32, 24)
    a, b = img[:, :, :3].astype(np.float64), 0.37 * ref[:, :, :3].astype(np.float64)
    assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)


# ---------------------------------------------------------------------------- GPU parity

def _c1_profiled(spp=8):
    sc, cams = scenes.c1(bvh=lambda p_, i: (lambda b: (b["nodes"], b["order"]))(oracle.build_bvh(p_, i)))
    sc["profiles"] = {k: ((0.15 * (k - 1.5), 0.1, -1.0), PHYS_TH, PHYS_P) for k in range(4)}
    return sc, dict(cams[0], spp=spp)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["persistent", "wavefront"])
def test_gpu_profiled_render_matches_oracle(mode):
    from vtwin import lib
    sc, cam = _c1_profiled()
    print("profiled render: mode %s" % mode, flush=True)
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(sc)
        ctx.upload_profiles(sc)
        ctx.set_camera(0, cam)
        ctx.render([0], mode=lib.VT_RENDER_WAVEFRONT if mode == "wavefront" else lib.VT_RENDER_BATCHED)
        img = ctx.readback(0, lib.VT_KIND_LINEAR, 640, 480)
        ctx.set_light_profile(0)                 # back to Lambertian: the image changes
        ctx.render([0])
        flat = ctx.readback(0, lib.VT_KIND_LINEAR, 640, 480)
    finally:
        ctx.close()
    x0, y0, w, h = 256, 160, 64, 48
    ref, _, _ = oracle.render_tile(sc, cam, x0, y0, w, h)
    a = img[y0:y0 + h, x0:x0 + w, :3].astype(np.float64)
    b = ref[:, :, :3].astype(np.float64)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    print("GPU vs oracle with profiles: relative RMSE %.3g" % rel)
    assert rel < 1e-3
    assert np.linalg.norm(flat - img) > 0.05 * np.linalg.norm(img)


@pytest.mark.gpu
def test_gpu_profiled_gradients_match_oracle():
    """Light and material adjoint (the profile scales each emission term) and camera tangents (the profile's slope
    in theta moves with the camera rays that see the emitters) against the oracle on the same samples."""
    from vtwin import lib
    from test_adjoint import TILE
    sc, cam = _c1_profiled()
    x0, y0, w, h = TILE
    r = scenes.Rng(99)
    adj = np.zeros((h, w, 4), np.float32)
    adj[:, :, 0:3] = r.uniform(-1, 1, w * h * 3).reshape(h, w, 3)
    full = np.zeros((cam["height"], cam["width"], 4), np.float32)
    full[y0:y0 + h, x0:x0 + w] = adj
    n_mat, n_light = sc["materials"].size // 16, sc["lights"].size // 4
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(sc)
        ctx.upload_profiles(sc)
        ctx.set_camera(0, cam)
        gm, gl, gc = ctx.backward([0], [full], n_mat, n_light, cameras=True)
    finally:
        ctx.close()
    om, ol = oracle.backward_tile(sc, cam, x0, y0, w, h, adj)
    assert np.allclose(gl, ol, rtol=1e-3, atol=1e-4 * np.max(np.abs(ol)))
    assert np.allclose(gm, om, rtol=1e-3, atol=1e-4 * np.max(np.abs(om)))
    want, changed = oracle.camera_grad_tile(sc, cam, x0, y0, w, h, adj)
    scale = np.array([cam["lens"][0]] * 4 + [1.0] * 10)
    a, b = gc[0] * scale, want * scale
    print("GPU   ", np.array2string(a, precision=5), "\noracle", np.array2string(b, precision=5))
    assert np.all(np.abs(a - b) <= 1e-3 * np.abs(b) + 2e-4 * np.abs(b).max())
