"""The ISP after the raw DN (gpu/wgsl/sensor.wgsl.js:210-427) and veiling glare (:87-141).

Oracle pins (CPU): flat raw fields in every demosaic mode and CFA (borders included) against the closed form; a
linear ramp reproduced exactly inside; bilinear mode against the textbook Bayer convolution kernels; the edge-aware
green along an edge; grey gives Cb = Cr = 128; the BT.601 primaries give the published JFIF YCbCr values and their
4:2:0 reconstruction; the tone curves against FP64 closed forms; the bilateral filter against a plain 3x3 Gaussian
when its range term vanishes, and on a step edge; the unsharp mask leaves a ramp alone and overshoots a step by the
closed form; the glare level of a single bright block against the separable Gaussian and its total.

GPU parity: the 8-bit RGBA, the Y / Cb / Cr planes and the AE / AWB sums bit for bit against the oracle on the same
DN (odd sizes, every mode); the camera path (render -> vt_sensor -> vt_isp -> readback kinds); DN with glare on.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle
from vtwin import scenes

SENS = dict(cfa=0, max_dn=4095.0, black=64.0, read_noise=6.0, conv_gain=0.33, gain=1.0)
FULL = SENS["max_dn"] - SENS["black"]


def _lin(over=None, **kw):
    p = dict(tone=2, chroma420=0)
    p.update(kw)
    return p


@pytest.mark.parametrize("cfa", [0, 1, 2, 3])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_flat_field_every_mode_and_border(cfa, mode):
    s = dict(SENS, cfa=cfa)
    dn = np.full((9, 11), 1000, np.uint16)
    o = oracle.isp(s, _lin(demosaic=mode), dn)
    c = (1000.0 - 64.0) / FULL
    assert np.all(o["rgba"][..., :3] == np.rint(np.float32(c) * np.float32(255.0)))
    assert np.allclose(o["lin"][..., :3], c, rtol=1e-6)
    assert list(o["stats"][1:]) == [round(c * 1023) * dn.size] * 3


def _ramp(w, h, a=300.0, bx=8.0, by=4.0):
    x, y = np.meshgrid(np.arange(w), np.arange(h))
    return (a + bx * x + by * y).astype(np.uint16)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_linear_ramp_exact_inside(mode):
    """A grey linear ramp: every interpolation the demosaic makes is exact on a linear function (means over
    symmetric sites; Hamilton-Adams' second derivatives vanish)."""
    dn = _ramp(24, 18)
    o = oracle.isp(dict(SENS, cfa=1), _lin(demosaic=mode), dn)
    want = (dn.astype(np.float64) - 64.0) / FULL
    inner = (slice(3, -3), slice(3, -3))      # the colour differences read greens that read mirrored samples
    for k in range(3):
        assert np.allclose(o["lin"][..., k][inner], want[inner], rtol=1e-6, atol=1e-7)


def test_bilinear_is_the_textbook_bayer_convolution():
    r = scenes.Rng(4)
    h, w = 20, 26
    dn = r.uniform(100, 4000, w * h).reshape(h, w).astype(np.uint16)
    o = oracle.isp(dict(SENS, cfa=0), _lin(demosaic=0), dn)
    y, x = np.mgrid[0:h, 0:w]
    mR = ((y % 2 == 0) & (x % 2 == 0)).astype(float)
    mB = ((y % 2 == 1) & (x % 2 == 1)).astype(float)
    mG = 1.0 - mR - mB
    kG = np.array([[0, 1, 0], [1, 4, 1], [0, 1, 0]]) / 4.0
    kRB = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]]) / 4.0
    raw = dn.astype(np.float64)
    planes = [ndimage.convolve(mR * raw, kRB), ndimage.convolve(mG * raw, kG), ndimage.convolve(mB * raw, kRB)]
    inner = (slice(1, -1), slice(1, -1))
    for k in range(3):
        want = (planes[k] - 64.0) / FULL
        assert np.allclose(o["lin"][..., k][inner], want[inner], rtol=2e-6, atol=1e-7), k


def test_edge_aware_green_follows_the_edge():
    """A vertical edge: at R and B sites beside it the edge-aware green averages along the edge (up, down), the
    bilinear green across it."""
    dn = np.full((12, 12), 1000, np.uint16)
    dn[:, 6:] = 3000
    s = dict(SENS, cfa=0)
    ea = oracle.isp_green(s, dict(demosaic=1), dn)
    bl = oracle.isp_green(s, dict(demosaic=0), dn)
    # (x = 5, y = 5) is a B site of RGGB (odd, odd), left of the edge
    assert ea[5, 5] == 1000.0
    assert bl[5, 5] == pytest.approx(0.25 * (1000 + 3000 + 1000 + 1000))


@pytest.mark.parametrize("tone", [0, 1, 2, 3])
def test_grey_gives_neutral_chroma(tone):
    for level in [64, 300, 1000, 2500, 4095]:
        dn = np.full((6, 8), level, np.uint16)
        o = oracle.isp(SENS, dict(tone=tone, chroma420=1, s_curve_c=1.4, black_offset=0.01), dn)
        assert np.all(o["cb"] == 128) and np.all(o["cr"] == 128)
        v = o["rgba"][..., 0]
        assert np.all(o["rgba"][..., 1] == v) and np.all(o["rgba"][..., 2] == v) and np.all(o["y"] == v)


@pytest.mark.parametrize("prim,ycc,rgb420", [(0, (76, 85, 255), (254, 0, 0)), (1, (150, 44, 21), (0, 255, 1)),
                                             (2, (29, 255, 107), (0, 0, 254))])
def test_bt601_primaries(prim, ycc, rgb420):
    """Full-range BT.601 (JFIF) YCbCr of pure red, green and blue, and the RGB rebuilt from the 4:2:0 planes."""
    ccm = np.zeros((3, 3))
    ccm[prim, prim] = 1.0
    dn = np.full((4, 6), 4095, np.uint16)
    o = oracle.isp(dict(SENS, black=0.0), dict(tone=2, chroma420=1, ccm=ccm), dn)
    assert (int(o["y"][0, 0]), int(o["cb"][0, 0]), int(o["cr"][0, 0])) == ycc
    Y, Cb, Cr = ycc
    want = [Y + 1.402 * (Cr - 128), Y - 0.344136 * (Cb - 128) - 0.714136 * (Cr - 128), Y + 1.772 * (Cb - 128)]
    want = tuple(int(np.clip(np.rint(v), 0, 255)) for v in want)
    assert want == rgb420
    assert tuple(int(v) for v in o["rgba"][0, 0, :3]) == rgb420


def test_tone_curves_against_closed_forms():
    x = np.linspace(-0.05, 1.05, 2203).astype(np.float32)
    c = np.clip(x.astype(np.float64), 0, 1)
    srgb = np.where(c <= 0.0031308, 12.92 * c, 1.055 * c ** (1 / 2.4) - 0.055)
    assert np.max(np.abs(oracle.isp_tone(dict(tone=0), x) - srgb)) < 2e-6
    assert np.max(np.abs(oracle.isp_tone(dict(tone=1, gamma=2.2), x) - c ** (1 / 2.2))) < 2e-6
    assert np.array_equal(oracle.isp_tone(dict(tone=2), x), np.clip(x, 0, 1))
    b, cc = 0.02, 1.6
    v = np.clip((x.astype(np.float64) - b) / (1 - b), 0, 1)
    u = np.where(v <= 0.0031308, 12.92 * v, 1.055 * v ** (1 / 2.4) - 0.055)
    s = u ** cc / (u ** cc + (1 - u) ** cc)
    got = oracle.isp_tone(dict(tone=3, s_curve_c=cc, black_offset=b), x)
    assert np.max(np.abs(got - s)) < 3e-6
    assert got[np.argmin(np.abs(u - 0.5))] == pytest.approx(0.5, abs=2e-3)   # S_c(1/2) = 1/2


def test_bilateral_reduces_to_a_gaussian_and_keeps_edges():
    r = scenes.Rng(8)
    h, w = 16, 18
    dn = r.uniform(500, 3500, w * h).reshape(h, w).astype(np.uint16)
    base = oracle.isp(dict(SENS, read_noise=1e6), _lin(), dn)["lin"]
    sg = 0.8
    nr = oracle.isp(dict(SENS, read_noise=1e6), _lin(nr=1, nr_sigma_px=sg), dn)["lin"]
    k = np.exp(-0.5 * (np.arange(-1, 2)[:, None] ** 2 + np.arange(-1, 2)[None, :] ** 2) / sg ** 2)
    k /= k.sum()
    inner = (slice(1, -1), slice(1, -1))
    for c in range(3):
        want = ndimage.convolve(base[..., c].astype(np.float64), k)
        assert np.allclose(nr[..., c][inner], want[inner], rtol=1e-5, atol=1e-7)
    # with the modelled noise, a flat field stays put and a step far above the noise survives
    flat = np.full((10, 10), 1500, np.uint16)
    a = oracle.isp(SENS, _lin(nr=1, nr_sigma_px=sg), flat)["lin"]
    assert np.allclose(a[..., :3], (1500 - 64.0) / FULL, rtol=1e-6)
    step = np.full((10, 10), 800, np.uint16)
    step[:, 5:] = 3200
    e0 = oracle.isp(SENS, _lin(), step)["lin"]
    e1 = oracle.isp(SENS, _lin(nr=1, nr_sigma_px=sg), step)["lin"]
    blur = oracle.isp(dict(SENS, read_noise=1e6), _lin(nr=1, nr_sigma_px=sg), step)["lin"]
    assert np.max(np.abs(e1 - e0)[:, list(range(3)) + list(range(7, 10))]) < 1e-6
    assert np.max(np.abs(e1 - e0)) < 0.1 * np.max(np.abs(blur - e0))


def test_unsharp_mask_ramp_and_step():
    s = dict(SENS, black=0.0)
    ramp = _ramp(20, 16, a=400.0, bx=16.0, by=0.0)
    a = oracle.isp(s, _lin(), ramp)["rgba"]
    b = oracle.isp(s, _lin(sharpen_amount=0.8, sharpen_radius=2), ramp)["rgba"]
    inner = (slice(4, -4), slice(4, -4))
    assert np.array_equal(a[inner], b[inner])
    # anywhere: c + amt (Y - box mean of Y), the box clamped at the borders (scipy's uniform filter, 'nearest')
    r = scenes.Rng(21)
    dn = r.uniform(200, 3800, 14 * 17).reshape(14, 17).astype(np.uint16)
    lin = oracle.isp(s, _lin(), dn)["lin"][..., :3].astype(np.float64)
    for amt, rad in [(0.6, 1), (1.3, 3)]:
        o = oracle.isp(s, _lin(sharpen_amount=amt, sharpen_radius=rad), dn)["rgba"][..., :3].astype(np.float64)
        t = np.clip(lin, 0, 1)
        Y = t @ np.array([0.299, 0.587, 0.114])
        m = ndimage.uniform_filter(Y, size=2 * rad + 1, mode="nearest")
        want = np.clip(t + amt * (Y - m)[..., None], 0, 1) * 255
        assert np.max(np.abs(o - want)) <= 0.5 + 1e-3
        assert np.max(np.abs(o - np.rint(t * 255))) >= 10      # it did sharpen


def test_awb_and_luma_sums():
    r = scenes.Rng(12)
    h, w = 15, 21
    dn = r.uniform(0, 4095, w * h).reshape(h, w).astype(np.uint16)
    o = oracle.isp(SENS, dict(tone=0, chroma420=1), dn)
    lin = oracle.isp(SENS, _lin(), dn)["lin"]
    for k in range(3):
        assert o["stats"][1 + k] == np.sum(np.rint(np.clip(lin[..., k], 0, 1) * 1023).astype(np.int64))
    assert o["stats"][0] == int(o["y"].astype(np.int64).sum())


def _glare_sensor(**kw):
    p = scenes.default_sensor(noise=False)
    p.update(glare_fraction=0.05, glare_sigma=2.0, glare_level=16)
    p.update(kw)
    return p


def test_glare_level_of_one_bright_block():
    """One F x F block of radiance v (RI 1): the level is v k_opt at that block, then the separable Gaussian
    g(d) = exp(-d^2 / 2 s^2) / (s sqrt(2 pi)) in each direction (zero outside, normalised on the infinite line)."""
    F, gw, gh, s = 16, 21, 17, 2.0
    w, h = F * gw, F * gh
    rad = np.zeros((h, w, 4), np.float32)
    bx, by, v = 10, 8, 5.0
    rad[by * F:(by + 1) * F, bx * F:(bx + 1) * F, :3] = v
    ri = np.ones((h, w), np.float32)
    p = _glare_sensor(glare_sigma=s)
    lvl = oracle.glare_level(p, rad, ri)
    g = lambda d: np.exp(-0.5 * d * d / (s * s)) / (s * math.sqrt(2 * math.pi))
    want = v * p["optics_k"] * g(np.arange(gh)[:, None] - by) * g(np.arange(gw)[None, :] - bx)
    for c in range(3):
        assert np.allclose(lvl[..., c], want, rtol=1e-5, atol=1e-9 * want.max())
    assert lvl[..., 0].sum() == pytest.approx(want.sum(), rel=1e-5)
    assert lvl[..., 0].sum() == pytest.approx(v * p["optics_k"], rel=1e-4)     # 4 sigma from the nearest edge


# ---------------------------------------------------------------------------- GPU parity

ISP_CASES = [
    dict(demosaic=0, tone=0, chroma420=1),
    dict(demosaic=1, tone=1, gamma=2.2, chroma420=0, wb=(1.9, 1.0, 1.6)),
    dict(demosaic=2, tone=3, s_curve_c=1.5, black_offset=0.02, nr=1, nr_sigma_px=0.6, sharpen_amount=0.7,
         sharpen_radius=2, chroma420=1, ccm=((1.6, -0.4, -0.2), (-0.3, 1.5, -0.2), (0.0, -0.5, 1.5))),
    dict(demosaic=2, tone=2, chroma420=0, nr=1, nr_sigma_px=0.4, sharpen_amount=0.3, sharpen_radius=4),
]


def _isp_dn(w, h, seed):
    """raw DN of a scene with edges, gradients and noise (the sensor model on a synthetic radiance image)"""
    r = scenes.Rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    rad = np.zeros((h, w, 4), np.float32)
    rad[..., 0] = 2.0 + 1.5 * np.sin(x / 7.0) + (x > w // 2) * 3.0
    rad[..., 1] = 1.0 + y / h * 4.0
    rad[..., 2] = 3.0 * ((x // 6 + y // 5) % 2)
    rad[..., :3] += r.uniform(0, 0.3, w * h * 3).reshape(h, w, 3).astype(np.float32)
    ri = np.ones((h, w), np.float32)
    p = scenes.default_sensor(exposure=4e-3, frame=3, seed=77)
    return oracle.sensor_expose(p, rad, ri)


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(ISP_CASES)))
@pytest.mark.parametrize("size", [(64, 48), (37, 23)])
def test_gpu_isp_bit_exact(case, size):
    from vtwin import lib
    w, h = size
    q = ISP_CASES[case]
    for cfa in (0, 1, 3):
        print("isp case %d cfa %d size %dx%d" % (case, cfa, w, h), flush=True)
        dn = _isp_dn(w, h, 10 + case)
        s = dict(SENS, cfa=cfa)
        want = oracle.isp(s, q, dn)
        ctx = lib.Context(0)
        try:
            got = ctx.isp_test(s, q, dn)
        finally:
            ctx.close()
        d = np.count_nonzero(got["rgba"] != want["rgba"])
        assert d == 0, "%d of %d bytes differ (max %d)" % (
            d, got["rgba"].size, np.abs(got["rgba"].astype(int) - want["rgba"]).max())
        assert np.array_equal(got["stats"], want["stats"])
        if q.get("chroma420", 1):
            for k in ("y", "cb", "cr"):
                assert np.array_equal(got[k], want[k]), k


@pytest.mark.gpu
def test_gpu_camera_isp_and_readback_kinds():
    """render -> vt_sensor -> vt_isp on the c1 camera; every readback kind against the oracle's ISP on the GPU's
    own raw DN."""
    from vtwin import lib
    sc, cams = scenes.c1()
    cam = dict(cams[0], spp=2)
    p = scenes.default_sensor(exposure=3e-3, frame=1, seed=5)
    q = ISP_CASES[2]
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(sc)
        ctx.set_camera(0, cam)
        ctx.render([0])
        ctx.sensor(0, p)
        ctx.isp(0, q)
        raw = ctx.readback(0, lib.VT_KIND_RAW, 640, 480)
        rgb = ctx.readback(0, lib.VT_KIND_RGB8, 640, 480)
        rgba = ctx.readback(0, lib.VT_KIND_RGBA8, 640, 480)
        yuv = ctx.readback(0, lib.VT_KIND_YUV420, 640, 480)
        stats = ctx.readback(0, lib.VT_KIND_STATS, 640, 480)
        ms = ctx.stats()["ms_isp"]
        ctx.isp(0, dict(q, chroma420=0))
        with pytest.raises(RuntimeError):
            ctx.readback(0, lib.VT_KIND_YUV420, 640, 480)
    finally:
        ctx.close()
    want = oracle.isp(dict(SENS, cfa=p["cfa"], black=p["black"], max_dn=p["max_dn"], read_noise=p["read_noise"],
                           conv_gain=p["conv_gain"], gain=p["gain"]), q, raw)
    print("ISP 640x480: %.3f ms; mean luma %.3f" % (ms, stats[0] / (640 * 480 * 255)))
    assert np.array_equal(rgba, want["rgba"])
    assert np.array_equal(rgb, want["rgba"][..., :3])
    assert np.array_equal(stats, want["stats"])
    for k in ("y", "cb", "cr"):
        assert np.array_equal(yuv[k], want[k]), k
    assert 0.02 < stats[0] / (640 * 480 * 255) < 0.98


@pytest.mark.gpu
@pytest.mark.parametrize("level,frac,sigma", [(16, 0.05, 2.0), (8, 0.2, 5.5), (5, 1.0, 0.7)])
def test_gpu_glare_dn_bit_exact(level, frac, sigma):
    from vtwin import lib
    from test_gpu_parity import _sensor_case
    rad, ri = _sensor_case(203, 157, 31, 3.0)
    rad[40:60, 100:130, :3] = 400.0          # a bright patch that glares
    p = _glare_sensor(glare_level=level, glare_fraction=frac, glare_sigma=sigma)
    p.update(noise_on=True, frame=2)
    print("glare F %d k %.2f sigma %.1f" % (level, frac, sigma), flush=True)
    ctx = lib.Context(0)
    try:
        got = ctx.sensor_test(p, rad, ri)
    finally:
        ctx.close()
    want = oracle.sensor_expose(p, rad, ri)
    off = oracle.sensor_expose(dict(p, glare_fraction=0.0), rad, ri)
    diff = np.count_nonzero(got != want)
    assert diff == 0, "%d DNs differ" % diff
    assert np.count_nonzero(want != off) > 0.05 * want.size


def _physics_sensor():
    """the reference's physics.json sensor / isp leaves (tests/golden/physics_sensor.json) as vt_sensor and vt_isp
    parameters, by the mapping of gpu/sensor.js:158-177"""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "physics_sensor.json")) as f:
        ph = json.load(f)
    v = lambda path: (lambda d: d["value"])(
        __import__("functools").reduce(lambda d, k: d[k], path.split("."), ph))
    esf = v("sensor.lens.esf_sigma_px")
    sens = dict(cfa={"RGGB": 0, "GRBG": 1, "GBRG": 2, "BGGR": 3}[v("sensor.bayer_order")], frame=4, seed=21,
                optics_k=math.pi * v("sensor.lens.transmission") / (4 * v("sensor.lens.f_number") ** 2),
                esf_sigma=math.sqrt(max(0.0, esf * esf - 1.0 / 12.0)), exposure=3e-3, gain=1.0,
                m=np.array(v("sensor.response_matrix.M"), np.float32), read_noise=v("sensor.read_noise"),
                full_well=v("sensor.full_well"), dark=v("sensor.dark_current"),
                conv_gain=v("sensor.conversion_gain_base"), max_dn=2.0 ** v("sensor.adc_bits") - 1,
                black=float(v("sensor.black_level")), noise_on=True, cam_index=0, pupil_exp=1.0,
                glare_fraction=v("sensor.lens.veiling_glare.fraction"),
                glare_sigma=v("sensor.lens.veiling_glare.sigma_frac_diag") * math.hypot(640, 480) / 16.0,
                glare_level=16)
    isp = dict(demosaic=1, nr=1, nr_sigma_px=v("isp.noise_reduction_sigma_px"), nr_range_scale=2.0, tone=3,
               s_curve_c=v("isp.s_curve_c"), black_offset=v("isp.black_offset_linear"), gamma=v("isp.gamma"),
               wb=v("isp.white_balance_gains"), ccm=v("isp.ccm"), sharpen_amount=v("isp.sharpening.amount"),
               sharpen_radius=v("isp.sharpening.radius_px"), chroma420=1)
    assert "edge" in v("isp.demosaic") and "edge" in v("isp.noise_reduction") and v("isp.tone_curve") == "srgb"
    return sens, isp


@pytest.mark.gpu
def test_gpu_physics_sensor_chain_on_a_rendered_camera():
    """The twin's own sensor and ISP settings (glare 0.1, black 168, full colour response matrix, dark current,
    edge-aware demosaic, edge-preserving NR, sRGB + S-curve, 4:2:0) on a rendered c1 frame: raw DN against the
    oracle's sensor on the GPU's radiance and RI map, then every 8-bit output against the oracle's ISP."""
    from vtwin import lib
    sc, cams = scenes.c1()
    cam = dict(cams[0], spp=2)
    sens, isp = _physics_sensor()
    print("physics sensor chain: render, sensor, isp", flush=True)
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(sc)
        ctx.set_camera(0, cam)
        ctx.render([0])
        lin = ctx.readback(0, lib.VT_KIND_LINEAR, 640, 480)
        ctx.sensor(0, sens)
        ctx.isp(0, isp)
        raw = ctx.readback(0, lib.VT_KIND_RAW, 640, 480)
        ri = ctx.readback(0, lib.VT_KIND_RI, 640, 480)
        rgba = ctx.readback(0, lib.VT_KIND_RGBA8, 640, 480)
        yuv = ctx.readback(0, lib.VT_KIND_YUV420, 640, 480)
        stats = ctx.readback(0, lib.VT_KIND_STATS, 640, 480)
    finally:
        ctx.close()
    want_dn = oracle.sensor_expose(sens, lin, ri)
    d = np.count_nonzero(raw != want_dn)
    assert d == 0, "%d DNs differ" % d
    want = oracle.isp(dict(SENS, cfa=sens["cfa"], black=sens["black"], max_dn=sens["max_dn"],
                           read_noise=sens["read_noise"], conv_gain=sens["conv_gain"], gain=sens["gain"]), isp, raw)
    assert np.array_equal(rgba, want["rgba"]) and np.array_equal(stats, want["stats"])
    for k in ("y", "cb", "cr"):
        assert np.array_equal(yuv[k], want[k]), k
    print("mean luma %.3f, raw mean %.1f DN" % (stats[0] / (640 * 480 * 255), raw.mean()))


@pytest.mark.gpu
def test_gpu_ri_map_matches_oracle_with_vignetting():
    """The device's relative-illumination map (vt_sensor's, read back as VT_KIND_RI) for the c1 KB4 camera against
    the oracle's relativeIlluminationMap with a vignetting table, and again after the table changes (the map is
    rebuilt for it)."""
    from vtwin import lib
    sc, cams = scenes.c1()
    cam = dict(cams[0], spp=1)
    lens = cam["lens"]
    Ls = oracle.lens(*[float(x) for x in lens[:4]], [float(x) for x in lens[4:8]], float(lens[8]))
    ctx = lib.Context(0)
    try:
        ctx.upload_scene(sc)
        ctx.set_camera(0, cam)
        ctx.render([0])
        for vt, vv, q in (([0.0, 20.0, 40.0, 60.0], [1.0, 0.95, 0.7, 0.4], 1.0), ((), (), 1.0),
                          ([10.0, 50.0], [0.9, 0.5], 0.5)):
            sens = dict(scenes.default_sensor(), pupil_exp=q, vignetting_theta_deg=vt, vignetting=vv)
            ctx.sensor(0, sens)
            ri = ctx.readback(0, lib.VT_KIND_RI, 640, 480)
            want = oracle.ri_map(Ls, 640, 480, q, vt, vv)
            err = np.max(np.abs(ri.astype(np.float64) - want) / np.maximum(want, 1e-6))
            print("RI map, vignetting %s, q %.1f: max relative difference %.2e" % (list(vv), q, err))
            assert err < 1e-6
    finally:
        ctx.close()

