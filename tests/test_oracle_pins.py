"""Pins of the oracle by what the mathematics fixes (closed forms, invariants, textbook reductions, independent
library routines), chosen so a plausible slip (a dropped term, a wrong sign or index) fails one of them."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate, special, stats

import oracle
from vtwin import lib, scenes

L = oracle.lib()


def _lens(k, f=300.0, cx=319.5, cy=239.5, alpha=0.0):
    return oracle.lens(f, f, cx, cy, list(k), alpha)


def _unproject(Ls, uv):
    r = np.zeros(3)
    L.or_unproject(C.byref(Ls), oracle.ptr(np.asarray(uv, np.float64)), oracle.ptr(r))
    return r


def _project(Ls, X):
    uv = np.zeros(2)
    L.or_project(C.byref(Ls), oracle.ptr(np.asarray(X, np.float64)), oracle.ptr(uv))
    return uv


# ---------------------------------------------------------------------------- KB4 lens
def test_kb4_reduces_to_equidistant():
    Ls = _lens([0, 0, 0, 0])
    for uv in ([500.0, 100.0], [10.0, 470.0], [319.5, 239.5 + 200]):
        r = _unproject(Ls, uv)
        rad = math.hypot(uv[0] - 319.5, uv[1] - 239.5) / 300.0
        assert math.acos(r[2]) == pytest.approx(rad, abs=1e-12)
        # azimuth preserved, OpenCV axes (x right, y down)
        assert math.atan2(r[1], r[0]) == pytest.approx(math.atan2(uv[1] - 239.5, uv[0] - 319.5), abs=1e-12)


def test_kb4_round_trip_and_derivative():
    r = scenes.Rng(5)
    for _ in range(20):
        k = [r.uniform(-0.05, 0.05) for _ in range(4)]
        Ls = _lens(k, alpha=r.uniform(-0.01, 0.01))
        thd_max = L.or_distort(C.byref(Ls), L.or_max_monotone_theta(C.byref(Ls), math.pi))
        checked = 0
        while checked < 10:
            uv = [r.uniform(0, 639), r.uniform(0, 479)]
            yd = (uv[1] - 239.5) / 300.0
            if math.hypot((uv[0] - 319.5) / 300.0 - Ls.alpha * yd, yd) > 0.98 * thd_max:
                continue          # outside the monotone range the inverse is not unique (gpu/lens.js:65)
            back = _project(Ls, _unproject(Ls, uv))
            assert np.allclose(back, uv, atol=1e-9)
            checked += 1
        for th in (0.1, 0.7, 1.2):
            h = 1e-6
            num = (L.or_distort(C.byref(Ls), th + h) - L.or_distort(C.byref(Ls), th - h)) / (2 * h)
            assert L.or_distort_deriv(C.byref(Ls), th) == pytest.approx(num, rel=1e-8)


def _scene_cameras():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kb4_scene_cameras.json")
    return json.load(open(path))["cameras"]


def _scene_lens(c):
    K = c["K"]
    return oracle.lens(K[0][0], K[1][1], K[0][2], K[1][2], c["D"], K[0][1] / K[0][0])


def test_kb4_matches_the_scene_export():
    """The reference's scene export records, per tracking camera, numbers its own KB4 model produced: the field
    angle of the aim target pixel, the largest field angle of the image (corners at the pixel edges, maxImageTheta,
    gpu/lens.js:76-83) and the fold of the mapping. The oracle's unproject reproduces the aim angle and the corner
    angle of all three measured lenses (k1..k4 all non-zero) to rounding, so every power and coefficient of
    distort / undistort is pinned; the fold is the root of dtheta_d/dtheta = 1 + 3k1 t^2 + 5k2 t^4 + 7k3 t^6 +
    9k4 t^8 (or pi/2 for a lens monotone that far), within the exporter's 1e-5, and maxMonotoneTheta's 4096-step
    scan brackets it."""
    for c in _scene_cameras():
        Ls = _lens_from(c)
        r = _unproject(Ls, c["aim_target_px"])
        ang = math.degrees(math.acos(r[2] / np.linalg.norm(r)))
        assert ang == pytest.approx(c["aim_target_off_axis_deg"], abs=1e-10), c["name"]
        K = c["K"]
        assert math.hypot(c["aim_target_px"][0] - K[0][2], c["aim_target_px"][1] - K[1][2]) == pytest.approx(
            c["aim_target_off_principal_px"], abs=1e-9)
        w, h = c["size"]
        assert L.or_max_image_theta(C.byref(Ls), w, h) == pytest.approx(c["theta_corner_rad"], abs=1e-12), c["name"]
        k1, k2, k3, k4 = c["D"]
        roots = [z.real for z in np.roots([9 * k4, 7 * k3, 5 * k2, 3 * k1, 1.0]) if abs(z.imag) < 1e-12 and z.real > 0]
        fold = min([math.sqrt(z) for z in roots] + [math.pi / 2])
        assert fold == pytest.approx(c["theta_fold_rad"], abs=1e-5), c["name"]
        assert L.or_distort_deriv(C.byref(Ls), fold * (1 - 1e-9)) > 0
        if fold < math.pi / 2:
            assert L.or_distort_deriv(C.byref(Ls), fold * (1 + 1e-6)) < 0
            t = L.or_max_monotone_theta(C.byref(Ls), math.pi)
            assert t <= c["theta_fold_rad"] < t + math.pi / 4096
        else:
            assert L.or_max_monotone_theta(C.byref(Ls), math.pi / 2) == pytest.approx(math.pi / 2)


def _lens_from(c):
    return _scene_lens(c)


def test_ri_map_closed_forms_and_vignetting():
    """relativeIlluminationMap (gpu/sensor.js:103-124): for k = 0 the equidistant lens gives cos^q(t) sin(t) / t at
    field angle t = radius / f; the vignetting table is read linearly between its knots and held at its ends."""
    Ls = _lens([0, 0, 0, 0])
    w, h, f = 640, 480, 300.0
    vt, vv = [0.0, 20.0, 40.0, 60.0], [1.0, 0.9, 0.6, 0.3]
    for q in (1.0, 0.5):
        ri = oracle.ri_map(Ls, w, h, q, vt, vv)
        ri0 = oracle.ri_map(Ls, w, h, q)
        for (x, y) in ((319, 239), (500, 100), (10, 470), (639, 0), (330, 250)):
            t = math.hypot(x - 319.5, y - 239.5) / f
            want = math.cos(t) ** q * math.sin(t) / t
            assert ri0[y, x] == pytest.approx(want, rel=1e-6)
            deg = math.degrees(t)
            v = np.interp(deg, vt, vv)
            assert ri[y, x] == pytest.approx(want * v, rel=1e-6), (x, y, deg)
    # the ends are held: a table starting at 30 degrees gives its first value on the axis
    ri = oracle.ri_map(Ls, w, h, 1.0, [30.0, 50.0], [0.8, 0.5])
    assert ri[239, 319] == pytest.approx(0.8 * oracle.ri_map(Ls, w, h, 1.0)[239, 319], rel=1e-6)
    corner_deg = math.degrees(math.hypot(319.5, 239.5) / f)
    assert corner_deg > 50
    assert ri[0, 0] == pytest.approx(0.5 * oracle.ri_map(Ls, w, h, 1.0)[0, 0], rel=1e-6)


def test_max_monotone_theta_brackets_the_turning_point():
    Ls = _lens([0.0, 0.0, 0.0, -0.5])           # d(theta_d)/d(theta) = 1 - 4.5 theta^8 -> root at 4.5^(-1/8)
    t = L.or_max_monotone_theta(C.byref(Ls), math.pi)
    root = 4.5 ** (-1 / 8)
    assert t <= root < t + math.pi / 4096 + 1e-12


def test_fp32_unproject_close_to_fp64():
    r = scenes.Rng(8)
    k = [r.uniform(-0.05, 0.05) for _ in range(4)]
    Ls = _lens(k)
    lens9 = np.array([300, 300, 319.5, 239.5] + k + [0.0], np.float32)
    for _ in range(50):
        uv = np.array([r.uniform(0, 639), r.uniform(0, 479)], np.float32)
        a = np.zeros(3, np.float32)
        L.or_unproject_f32(oracle.ptr(lens9), oracle.ptr(uv), oracle.ptr(a))
        b = _unproject(Ls, uv.astype(np.float64))
        assert np.linalg.norm(a - b) < 3e-6


# ---------------------------------------------------------------------------- BSDF
MATS = {
    "lambert": scenes.material_record(scenes.KIND_LAMBERT, (0.6, 0.4, 0.3)),
    "dielectric": scenes.material_record(scenes.KIND_DIELECTRIC, (0.5, 0.6, 0.2), roughness=0.3),
    "conductor": scenes.material_record(scenes.KIND_CONDUCTOR, (0.9, 0.7, 0.5), roughness=0.2),
    # coat roughness 0.3 keeps the coat lobe wide enough for the quadrature grids below
    "coated": scenes.material_record(scenes.KIND_COATED, (0.8, 0.8, 0.9), roughness=0.4, tau=(0.9, 0.8, 0.7),
                                     coat_roughness=0.3),
}


def _f(M, n, l, v):
    out = np.zeros(3)
    L.or_eval_bsdf(C.byref(M), oracle.ptr(np.asarray(n, float)), oracle.ptr(np.asarray(l, float)),
                   oracle.ptr(np.asarray(v, float)), oracle.ptr(out))
    return out


def _pdf(M, n, v, l):
    return L.or_pdf_bsdf(C.byref(M), oracle.ptr(np.asarray(n, float)), oracle.ptr(np.asarray(v, float)),
                         oracle.ptr(np.asarray(l, float)))


def _hemi(nt=96, npz=192):
    """midpoint grid over the upper hemisphere in (mu = cos theta, phi): directions and solid-angle weights"""
    mu = (np.arange(nt) + 0.5) / nt
    ph = (np.arange(npz) + 0.5) * 2 * math.pi / npz
    M, P = np.meshgrid(mu, ph, indexing="ij")
    s = np.sqrt(1 - M * M)
    d = np.stack([s * np.cos(P), s * np.sin(P), M], -1).reshape(-1, 3)
    return d, np.full(d.shape[0], (1.0 / nt) * (2 * math.pi / npz))


@pytest.mark.parametrize("name", list(MATS))
def test_bsdf_reciprocity(name):
    M = oracle.mat_from_record(MATS[name])
    r = scenes.Rng(3)
    n = np.array([0.0, 0.0, 1.0])
    for _ in range(200):
        a = np.array([r.uniform(-1, 1), r.uniform(-1, 1), r.uniform(0.05, 1)])
        b = np.array([r.uniform(-1, 1), r.uniform(-1, 1), r.uniform(0.05, 1)])
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        assert np.allclose(_f(M, n, a, b), _f(M, n, b, a), rtol=1e-12, atol=1e-14)


def test_lambert_is_rho_over_pi_and_albedo_rho():
    M = oracle.mat_from_record(MATS["lambert"])
    n = np.array([0.0, 0.0, 1.0])
    v = np.array([0.3, 0.1, 0.9])
    v /= np.linalg.norm(v)
    d, w = _hemi()
    A = sum(_f(M, n, l, v) * l[2] * wi for l, wi in zip(d, w))
    assert np.allclose(A, [0.6, 0.4, 0.3], rtol=2e-3)
    assert np.allclose(_f(M, n, d[7], v), np.array([0.6, 0.4, 0.3]) / math.pi, rtol=1e-15)


@pytest.mark.parametrize("name", list(MATS))
def test_pdf_integrates_to_one_and_matches_sampling(name):
    M = oracle.mat_from_record(MATS[name])
    n = np.array([0.0, 0.0, 1.0])
    v = np.array([0.2, 0.0, 0.98])
    v /= np.linalg.norm(v)
    d, w = _hemi(160, 320)
    tot = sum(_pdf(M, n, v, l) * wi for l, wi in zip(d, w))
    # at most 1; VNDF reflections leaving below the horizon take a few percent for the rough lobes
    assert 0.95 < tot < 1.01
    # albedo by quadrature of f cos vs the sampler's estimator f cos / pdf
    A_quad = sum(_f(M, n, l, v) * l[2] * wi for l, wi in zip(d, w))
    r = scenes.Rng(11)
    acc = np.zeros(3)
    N = 20000
    for _ in range(N):
        u = np.array([r.u(), r.u(), r.u()])
        l = np.zeros(3)
        p = L.or_sample_bsdf(C.byref(M), oracle.ptr(n), oracle.ptr(v), oracle.ptr(u), oracle.ptr(l))
        if p > 0:
            acc += _f(M, n, l, v) * l[2] / p
    A_mc = acc / N
    assert np.allclose(A_mc, A_quad, rtol=0.03)


def test_fresnel_average_simpson_vs_quad():
    for eta in (1.33, 1.5, 2.0):
        want = 2 * integrate.quad(lambda mu: L.or_fresnel_dielectric(mu, eta) * mu, 0, 1, limit=200)[0]
        assert L.or_fresnel_average(eta, 4096) == pytest.approx(want, abs=1e-9)
    # normal incidence closed form ((eta - 1)/(eta + 1))^2
    assert L.or_fresnel_dielectric(1.0, 1.5) == pytest.approx(0.04, abs=1e-15)


def test_ggx_d_normalised():
    for a in (0.05, 0.3, 0.8):
        val = integrate.quad(lambda mu: L.or_ggx_d(mu, a) * mu * 2 * math.pi, 0, 1, limit=400, points=[1.0])[0]
        assert val == pytest.approx(1.0, abs=2e-4)


@pytest.mark.parametrize("rough", [0.3, 0.55, 0.8])
def test_ggx_weak_white_furnace(rough):
    """The visible-normal distribution integrates to one for every view (Heitz 2014, the weak white furnace):
    int G1(v) max(0, v.h) D(h) / (n.v) dw_h = 1. It ties the Smith masking G1 = 1 / (1 + Lambda) to D: a different
    Lambda (a wrong power of tan or of alpha) or an unnormalised D fails it."""
    a = rough * rough
    nt, npz = 6000, 256
    mu = (np.arange(nt) + 0.5) / nt
    ph = (np.arange(npz) + 0.5) * 2 * math.pi / npz
    Mu, Ph = np.meshgrid(mu, ph, indexing="ij")
    s = np.sqrt(1 - Mu * Mu)
    h = np.stack([s * np.cos(Ph), s * np.sin(Ph), Mu], -1).reshape(-1, 3)
    dw = (1.0 / nt) * (2 * math.pi / npz)
    D = np.array([L.or_ggx_d(float(m), a) for m in mu])
    Dh = np.repeat(D[:, None], npz, 1).reshape(-1)
    for tv in (0.0, 0.6, 1.2, 1.45):
        v = np.array([math.sin(tv), 0.0, math.cos(tv)])
        nv = v[2]
        g1 = L.or_ggx_g1(nv, a)
        val = g1 * np.sum(np.maximum(0.0, h @ v) * Dh) * dw / nv
        assert val == pytest.approx(1.0, abs=2e-3), (rough, tv)


def test_ggx_height_correlated_smith_and_schlick_limits():
    """G2 is the height-correlated form 1 / (1 + Lambda(l) + Lambda(v)) (gpu/bsdf.js:59), so at l = v it equals
    1 / (2 / G1 - 1), which the separable product G1^2 does not; Lambda vanishes at normal incidence. Schlick's
    Fresnel is F0 at normal incidence, 1 at grazing and (1 - c)^5 in between (gpu/bsdf.js:44)."""
    for a in (0.01, 0.09, 0.4):
        assert L.or_ggx_lambda(1.0, a) == pytest.approx(0.0, abs=1e-15)
        for c in (0.05, 0.3, 0.7, 0.95):
            g1 = L.or_ggx_g1(c, a)
            assert L.or_ggx_g2(c, c, a) == pytest.approx(1 / (2 / g1 - 1), rel=1e-12)
            if a >= 0.4 and c < 0.5:        # where the two forms differ by more than rounding shows
                assert abs(L.or_ggx_g2(c, c, a) - g1 * g1) > 1e-3
    for f0 in (0.04, 0.5, 0.95):
        assert L.or_schlick(f0, 1.0) == pytest.approx(f0, abs=1e-15)
        assert L.or_schlick(f0, 0.0) == pytest.approx(1.0, abs=1e-15)
        assert L.or_schlick(f0, 0.5) == pytest.approx(f0 + (1 - f0) * 0.5 ** 5, abs=1e-15)
        xs = [L.or_schlick(f0, c) for c in np.linspace(0, 1, 21)]
        assert all(x >= y for x, y in zip(xs, xs[1:]))
    # exact dielectric Fresnel: normal incidence ((eta - 1) / (eta + 1))^2, grazing 1
    for eta in (1.33, 1.5, 2.4):
        assert L.or_fresnel_dielectric(1.0, eta) == pytest.approx(((eta - 1) / (eta + 1)) ** 2, rel=1e-12)
        assert L.or_fresnel_dielectric(0.0, eta) == pytest.approx(1.0, abs=1e-12)


def test_jacobi_update_delta_field_on_a_grid():
    """jacobiUpdate's smoothing (kernels.wgsl.js:440-461) on nodes of a 3 mm square grid, sigma 3 mm, radius 6.5 mm:
    every interior node sees the same 13-node stencil (offsets 0, 3, 4.24 and 6 mm), so for aux = 1 at one node and 0
    elsewhere each interior neighbour of that node gets E_free = exp(-d^2 / (2 sigma^2)) / S with S the stencil's
    weight sum from the grid geometry, and the source L = (B_dir + E_free V A_hh) / (pi A_hh). A missing weight, a
    missing normalisation or a node id taken for an active index fails it."""
    g = 11
    xs, ys = np.meshgrid(np.arange(g) * 0.003, np.arange(g) * 0.003, indexing="ij")
    pos = np.stack([xs.ravel(), ys.ravel(), np.zeros(g * g)], 1).astype(np.float32)
    nrm = np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (g * g, 1))
    n_static = g * g + 5                      # node ids 0..4 are not active: active index != node id
    active = (np.arange(g * g) + 5).astype(np.uint32)
    nb, nw = lib.node_neighbours(pos, nrm, 16, 0.0065, 0.003)
    sig2 = 0.003 ** 2
    S = 1 + 4 * math.exp(-0.5) + 4 * math.exp(-1.0) + 4 * math.exp(-2.0)
    centre = (g // 2) * g + g // 2
    aux = np.zeros((g * g, 4), np.float32)
    aux[centre, 0:3] = (1.0, 2.0, 0.5)
    state = np.zeros((n_static, 16), np.float32)
    state[active, 3] = 0.8                    # V
    state[active, 12:15] = 0.1                # B_dir
    node_mat = np.zeros(n_static, np.uint32)
    hh = np.array([[0.5, 0.25, 0.8]], np.float32)
    L.or_jacobi_update(g * g, oracle.ptr(active), oracle.ptr(nb), oracle.ptr(nw), 16, oracle.ptr(aux),
                       oracle.ptr(node_mat), oracle.ptr(np.ascontiguousarray(hh.reshape(-1))), oracle.ptr(state), 2)
    cx, cy = pos[centre, 0], pos[centre, 1]
    checked = 0
    for k in range(g * g):
        ix, iy = divmod(k, g)
        if not (2 <= ix < g - 2 and 2 <= iy < g - 2):
            continue                           # the stencil is whole only two nodes in from the border
        d2 = float((pos[k, 0] - cx) ** 2 + (pos[k, 1] - cy) ** 2)
        w = math.exp(-0.5 * d2 / sig2) if d2 <= 0.0065 ** 2 else 0.0
        want = np.array([1.0, 2.0, 0.5]) * w / S
        assert np.allclose(state[active[k], 0:3], want, rtol=1e-5, atol=1e-9), k
        Lw = (0.1 + want * 0.8 * hh[0]) / (math.pi * hh[0])
        assert np.allclose(state[active[k], 8:11], Lw, rtol=1e-5), k
        assert state[active[k], 11] == 1.0
        checked += w > 0
    assert checked == 13


def test_dielectric_directional_albedo_at_most_one():
    M = oracle.mat_from_record(scenes.material_record(scenes.KIND_DIELECTRIC, (1.0, 1.0, 1.0), roughness=0.5))
    spec, diff, hh, hs = np.zeros(32 * 3), np.zeros(32 * 3), np.zeros(3), np.zeros(3)
    L.or_albedo_table(C.byref(M), 32, 1024, oracle.ptr(spec), oracle.ptr(diff), oracle.ptr(hh), oracle.ptr(hs))
    assert np.all(spec + diff <= 1.0 + 1e-9)
    # diffuse part analytic: rho (1 - F(mu)) (gpu/bsdf.js:240-241 with fbar = fbar_computed)
    mu = (np.arange(32) + 0.5) / 32
    F = np.array([L.or_fresnel_dielectric(m, 1.5) for m in mu])
    assert np.allclose(diff.reshape(32, 3)[:, 0], 1 - F, rtol=1e-12)


# ---------------------------------------------------------------------------- sensor model routines
def _ulps(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def test_own_exp_log_cos_within_two_ulp():
    xs = np.linspace(-87.0, 88.0, 20001).astype(np.float32)
    got = np.array([L.or_expf(float(x)) for x in xs], np.float32)
    assert _ulps(got, np.exp(xs.astype(np.float64)).astype(np.float32)).max() <= 2
    xs = np.exp(np.linspace(-60, 60, 20001)).astype(np.float32)
    got = np.array([L.or_logf(float(x)) for x in xs], np.float32)
    want = np.log(xs.astype(np.float64))
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-30)) < 2.5e-7 or _ulps(got, want.astype(np.float32)).max() <= 2
    xs = np.linspace(0, 2 * math.pi, 20001).astype(np.float32)
    got = np.array([L.or_cosf(float(x)) for x in xs], np.float32)
    assert np.max(np.abs(got - np.cos(xs.astype(np.float64)))) < 2.5e-7
    zs = np.linspace(-4, 4, 4001).astype(np.float32)
    got = np.array([L.or_erfa(float(z)) for z in zs])
    # A&S 7.1.26 bounds the formula by 1.5e-7; binary32 evaluation adds a few ulp of 1 (6e-8 each)
    assert np.max(np.abs(got - special.erf(zs.astype(np.float64)))) < 4e-7


def _flat(level, w=256, h=256):
    rad = np.zeros((h, w, 4), np.float32)
    rad[:, :, 0:3] = level
    return rad, np.ones((h, w), np.float32)


def test_sensor_noise_off_closed_form():
    p = scenes.default_sensor(noise=False, exposure=1e-3)
    for level in (0.01, 0.3, 2.0):
        rad, ri = _flat(level, 64, 48)
        dn = oracle.sensor_expose(p, rad, ri)
        E = level * p["optics_k"]
        for ch, (y, x) in enumerate(((0, 0), (0, 1), (1, 1))):        # RGGB: R at (0,0), G at (0,1), B at (1,1)
            e = min(p["m"][ch, ch] * E * p["exposure"], p["full_well"])
            want = min(max(math.floor(e * p["conv_gain"] * p["gain"] + p["black"] + 0.5), 0), 4095)
            assert abs(int(dn[y, x]) - want) <= 1
            assert np.all(dn[y::2, x::2] == dn[y, x])


@pytest.mark.parametrize("lam_target", [3.5, 30.0, 900.0])
def test_sensor_noise_statistics_match_raw_noise_model(lam_target):
    p = scenes.default_sensor(noise=True, exposure=1e-3, read_noise_dn=2.0)
    p["black"] = 64.0            # a pedestal keeps negative read-noise excursions out of the clip at 0
    g = p["m"][1, 1]
    level = lam_target / (g * p["optics_k"] * p["exposure"])
    rad, ri = _flat(level, 256, 256)
    dn = oracle.sensor_expose(p, rad, ri).astype(np.float64)
    green = np.concatenate([dn[0::2, 1::2].ravel(), dn[1::2, 0::2].ravel()])
    lam = g * level * p["optics_k"] * p["exposure"]
    mean_want = lam * p["conv_gain"] * p["gain"] + p["black"]
    sd_want = L.or_raw_noise_dn(lam, p["read_noise"], p["conv_gain"], p["gain"])
    n = green.size
    assert green.mean() == pytest.approx(mean_want, abs=4 * sd_want / math.sqrt(n) + 0.5)
    var_want = sd_want ** 2 + 1.0 / 12.0          # + uniform quantisation of the ADC
    assert green.var() == pytest.approx(var_want, rel=0.06, abs=0.05)


def test_poisson_inversion_distribution():
    """Read noise off: electrons are exactly Poisson(lambda); DN = e * cg with cg = 1 keeps the counts."""
    p = scenes.default_sensor(noise=True, exposure=1e-3)
    p["read_noise"] = 0.0
    p["conv_gain"] = 1.0
    g = p["m"][0, 0]
    lam = 3.7
    rad, ri = _flat(lam / (g * p["optics_k"] * p["exposure"]), 256, 256)
    dn = oracle.sensor_expose(p, rad, ri)
    red = dn[0::2, 0::2].ravel().astype(int)
    ks = np.arange(0, 13)
    obs = np.array([np.sum(red == k) for k in ks] + [np.sum(red > 12)])
    pmf = np.append(stats.poisson.pmf(ks, lam), stats.poisson.sf(12, lam))
    chi2 = np.sum((obs - red.size * pmf) ** 2 / (red.size * pmf))
    assert chi2 < stats.chi2.ppf(0.999, len(obs) - 1)


# ---------------------------------------------------------------------------- radiosity
def test_dense_jacobi_matches_direct_solve():
    r = scenes.Rng(21)
    n, k = 96, 5
    F = r.uniform(0, 1, n * n).reshape(n, n)
    F = (F / F.sum(1, keepdims=True) * 0.85).astype(np.float32)
    e = r.uniform(0, 10, n * k).reshape(n, k).astype(np.float32)
    x = np.zeros((n, k))
    res = C.c_double(0)
    it = L.or_jacobi_dense(oracle.ptr(F), oracle.ptr(e), n, k, 1e-12, 2000, oracle.ptr(x), C.byref(res))
    want = np.linalg.solve(np.eye(n) - F.astype(np.float64), e.astype(np.float64))
    assert np.allclose(x, want, rtol=1e-10)
    assert 0 < it < 2000 and res.value <= 1e-12


def test_dense_jacobi_uniform_box():
    n, rho = 64, 0.9
    F = np.full((n, n), rho / n, np.float32)
    e = np.ones((n, 1), np.float32)
    x = np.zeros((n, 1))
    res = C.c_double(0)
    L.or_jacobi_dense(oracle.ptr(F), oracle.ptr(e), n, 1, 1e-13, 5000, oracle.ptr(x), C.byref(res))
    assert np.allclose(x, 1 / (1 - np.float64(np.float32(rho / n)) * n), rtol=1e-9)


def _lambert_tables(rho, bins=32):
    alb = np.zeros((1, bins, 2, 3), np.float32)
    alb[0, :, 1, :] = rho
    return alb[:, :, 0, :].reshape(-1).copy(), alb[:, :, 1, :].reshape(-1).copy(), np.full(3, rho, np.float32)


def test_gather_uniform_field_identity():
    """kernels.wgsl.js:414-415: with Lambertian tables and every gather ray hitting a node of source Ls, the
    gathered irradiance equals pi * (rho Ls) exactly."""
    N, K, rho, Ls = 40, 64, 0.7, 2.5
    r = scenes.Rng(2)
    verts = np.zeros((N, 8), np.float32)
    nrm = np.stack([r.uniform(-1, 1, N), r.uniform(-1, 1, N), r.uniform(0.2, 1, N)], 1)
    verts[:, 4:7] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    gt = r.int(0, N - 1, n=N * K).astype(np.uint32)
    spec, diff, hh = _lambert_tables(rho)
    st = np.zeros((N, 16), np.float32)
    st[:, 4:7] = Ls
    active = np.arange(N, dtype=np.uint32)
    aux = np.zeros((N, 4), np.float32)
    nm = np.zeros(N, np.uint32)
    L.or_jacobi_gather(N, oracle.ptr(active), oracle.ptr(gt), K, oracle.ptr(verts), oracle.ptr(nm), oracle.ptr(spec),
                       oracle.ptr(diff), 32, oracle.ptr(hh), oracle.ptr(st), 1, oracle.ptr(aux))
    assert np.allclose(aux[:, 0:3], math.pi * rho * Ls, rtol=1e-6)


def test_gather_dirs_are_cosine_distributed():
    K = 4096
    n = np.array([0.3, -0.2, 0.93], np.float32)
    d = np.zeros((K, 3), np.float32)
    for j in range(K):
        L.or_gather_dir(7, j, K, oracle.ptr(n), oracle.ptr(d[j]))
    mu = d @ (n / np.linalg.norm(n))
    assert np.all(mu > -1e-6)
    # cosine-weighted: mu^2 is uniform on [0, 1]
    assert stats.kstest(mu ** 2, "uniform").pvalue > 0.01


# ---------------------------------------------------------------------------- the estimator
def _plane_and_light(rho=0.5, Le=10.0, a=0.2, height=1.0, depth=1):
    B = scenes.Builder()
    P = np.array([[-5, -5, 0], [5, -5, 0], [5, 5, 0], [-5, 5, 0]], float)
    B.add(P, np.tile([0, 0, 1.0], (4, 1)), np.array([[0, 1, 2], [0, 2, 3]]), 0)
    Pq, Nq, Iq = scenes.quad((0.0, 0.0, height), 2 * a, (0.0, 0.0, -1.0))
    B.add(Pq, Nq, Iq, 1)
    P, N, I, M = B.arrays()
    mats = [scenes.material_record(scenes.KIND_LAMBERT, (rho,) * 3), scenes.material_record(scenes.KIND_LAMBERT, (0, 0, 0))]
    meshes = [[0, scenes.MESH_STATIC, 0, 0], [1, scenes.MESH_EMITTER, 0, 0]]
    bvh = lambda p, i: (lambda b: (b["nodes"], b["order"]))(oracle.build_bvh(p, i))
    sc = scenes.pack_scene(P, N, I, M, meshes, mats, [[Le, Le, Le, 0]], bvh=bvh)
    cam = scenes.camera(scenes.look_at((0, 0.0, 0.5), (0, 0, 0), up=(0, 1, 0)), (0, 0.0, 0.5),
                        [400, 400, 31.5, 31.5, 0, 0, 0, 0, 0], 64, 64, 256, 0, depth, mode=1)
    return sc, cam


def _square_irradiance(Le, a, h):
    """irradiance at the origin under a square emitter [-a, a]^2 at height h (Lambertian, facing down)"""
    f = lambda y, x: Le * h * h / (x * x + y * y + h * h) ** 2
    return integrate.dblquad(f, -a, a, -a, a, epsabs=1e-12)[0]


def test_estimator_nee_against_closed_form_irradiance():
    sc, cam = _plane_and_light()
    img, rays, _ = oracle.render_tile(sc, cam, 30, 30, 4, 4, 0, 256)
    want = 0.5 / math.pi * _square_irradiance(10.0, 0.2, 1.0)
    got = img[:, :, 0].mean()
    assert got == pytest.approx(want, rel=0.01)


def test_estimator_mis_bounce_agrees_with_nee_only():
    """With depth 2 the plane's second vertex is the light itself (paths end there): the MIS-weighted BSDF hits
    and the NEE samples must sum to the same irradiance as NEE alone at depth 1."""
    sc, cam = _plane_and_light(depth=2)
    img2, _, _ = oracle.render_tile(sc, cam, 30, 30, 4, 4, 0, 1024)
    want = 0.5 / math.pi * _square_irradiance(10.0, 0.2, 1.0)
    assert img2[:, :, 0].mean() == pytest.approx(want, rel=0.015)


def furnace_sphere(rho=0.8, Le=5.0, n_theta=32, n_phi=64, cap_rows=6, depth=4):
    """Unit sphere seen from its centre: an emitting cap (radiance Le, inward) and Lambertian albedo rho elsewhere.
    Inside a sphere the form factor from any point to a patch is area / (4 pi R^2), so with f the cap's share
    of the area every non-emitting point has L_D = rho Le f sum_{k<D} (rho (1 - f))^k after D bounces."""
    B = scenes.Builder()
    P, N, I = scenes.dome(1.0, n_theta, n_phi, pole=True)
    fan = n_phi + 2 * n_phi * (cap_rows - 1)
    B.add(P, N, I[:fan], 1)
    B.add(P, N, I[fan:], 0)
    flip = np.array([1.0, 1.0, -1.0])
    B.add(P * flip, N * flip, I[:, ::-1].copy(), 0)
    P, N, I, M = B.arrays()
    mats = [scenes.material_record(scenes.KIND_LAMBERT, (rho,) * 3), scenes.material_record(scenes.KIND_LAMBERT, (0, 0, 0))]
    meshes = [[0, scenes.MESH_STATIC, 0, 0], [1, scenes.MESH_EMITTER, 0, 0]]
    bvh = lambda p, i: (lambda b: (b["nodes"], b["order"]))(oracle.build_bvh(p, i))
    sc = scenes.pack_scene(P, N, I, M, meshes, mats, [[Le, Le, Le, 0]], bvh=bvh)
    a, b, c = P[I[:, 0]], P[I[:, 1]], P[I[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    f = area[M == 1].sum() / area.sum()
    cam = scenes.camera(scenes.look_at((0, 0, 0), (0, 0, -1), up=(0, 1, 0)), (0, 0, 0),
                        [40, 40, 31.5, 31.5, 0, 0, 0, 0, 0], 64, 64, 256, 0, depth, mode=1)
    q = rho * (1 - f)
    return sc, cam, rho * Le * f * (1 - q ** depth) / (1 - q)


@pytest.mark.parametrize("depth", [1, 2, 4])
def test_estimator_white_furnace_sphere(depth):
    """NEE alone (depth 1), and NEE + MIS-weighted continuations (depth 2, 4) against the geometric series; the
    inscribed polyhedron departs from the sphere's uniform form factor by about 1e-3."""
    sc, cam, want = furnace_sphere(depth=depth)
    img, _, _ = oracle.render_tile(sc, cam, 16, 16, 32, 32, 0, 256)
    assert img[:, :, 0].mean() == pytest.approx(want, rel=4e-3)


# ---------------------------------------------------------------------------- camera-parameter gradient
def _polygon_irradiance(p, verts, Le):
    """Irradiance on the plane z = const (normal +z) under a Lambertian polygon of radiance Le: Lambert's
    contour formula E = Le/2 sum_i gamma_i (n . unit(r_i x r_i+1))."""
    E = np.zeros(p.shape[0])
    for i in range(len(verts)):
        a, b = verts[i] - p, verts[(i + 1) % len(verts)] - p
        an, bn = a / np.linalg.norm(a, axis=1, keepdims=True), b / np.linalg.norm(b, axis=1, keepdims=True)
        g = np.arccos(np.clip((an * bn).sum(1), -1, 1))
        c = np.cross(a, b)
        E += 0.5 * g * (c / np.linalg.norm(c, axis=1, keepdims=True))[:, 2]
    return Le * np.abs(E)


def _rodrigues(r):
    th = np.linalg.norm(r)
    if th == 0:
        return np.eye(3)
    k = r / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K


def _kb4_dir(u, v, th, alpha=0.0):
    """camera-frame direction of image point (u, v) under KB4 intrinsics th[0:8], Newton in FP64"""
    fx, fy, cx, cy, k1, k2, k3, k4 = th[0:8]
    yd = (v - cy) / fy
    xd = (u - cx) / fx - alpha * yd
    thd = np.hypot(xd, yd)
    t = thd.copy()
    for _ in range(50):
        t2 = t * t
        t = t - (t * (1 + t2 * (k1 + t2 * (k2 + t2 * (k3 + t2 * k4)))) - thd) / (1 + t2 * (3 * k1 + t2 * (5 * k2 + t2 * (7 * k3 + t2 * 9 * k4))))
    s = np.sin(t) / thd
    return np.stack([xd * s, yd * s, np.cos(t)], 1)


def _plane_image(th, cam, xs, ys, verts, rho, Le):
    """closed-form image of the Lambertian plane z = 0 under the square emitter (depth 1) at pixel centres"""
    if cam["mode"] == 1:
        dc = np.stack([(xs - th[2]) / th[0], (ys - th[3]) / th[1], np.ones_like(xs)], 1)
        dc /= np.linalg.norm(dc, axis=1, keepdims=True)
    else:
        dc = _kb4_dir(xs, ys, th)
    R = np.asarray(cam["R"], float).reshape(3, 3).T
    d = (_rodrigues(th[8:11]) @ (R @ dc.T)).T
    o = th[11:14]
    p = o[None, :] + d * (-o[2] / d[:, 2])[:, None]
    return rho / math.pi * _polygon_irradiance(p, verts, Le)


@pytest.mark.parametrize("mode", [1, 0])
def test_camera_gradient_against_closed_form(mode):
    """The oracle's per-sample central differences (interior term) against the derivative of the closed-form
    image of a Lambertian plane under a square emitter, for every camera parameter: intrinsics, the rotation
    about the world axes and the centre. Pinhole, and KB4 with a wide field so k1..k4 matter."""
    sc, _ = _plane_and_light()
    pos = (0.15, 0.05, 0.5)
    if mode == 1:
        lens, tile = [400, 400, 31.5, 31.5, 0, 0, 0, 0, 0], (24, 28, 8, 8)
    else:
        lens, tile = [100, 100, 31.5, 31.5, 0.03, -0.02, 0.01, -0.005, 0], (50, 40, 8, 8)
    cam = scenes.camera(scenes.look_at(pos, (0.15, 0.05, 0.0), up=(0, 1, 0)), pos, lens, 64, 64, 256, 0, 1, mode=mode)
    x0, y0, w, h = tile
    adj = np.ones((h, w, 4), np.float32)
    adj[:, :, 3] = 0
    g, changed = oracle.camera_grad_tile(sc, cam, x0, y0, w, h, adj)
    assert changed == 0
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
    # compare on natural scales (pixels for fx..cy, radians for k and r, metres for t); Monte Carlo noise ~1 %
    scale = np.array([lens[0]] * 4 + [1.0] * 10)
    gs, ws = g * scale, want * scale
    print("oracle", np.array2string(gs, precision=4), "\nclosed", np.array2string(ws, precision=4))
    assert np.all(np.abs(gs - ws) <= 0.03 * np.abs(ws) + 0.005 * np.abs(ws).max())
