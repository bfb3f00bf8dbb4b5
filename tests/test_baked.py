"""The lite renderer's radiosity on its baked node table (lite/shaders.js:234-283; lite/renderer.js:214-222,
486-517): gatherOcc's transmission bytes and the Jacobi E_ind = pi/K sum_k t_k B(hit_k), B = (E_dir + hh E_ind)/pi.

Oracle pins (CPU): the segment-segment distance of gatherOcc against a bounded numerical minimiser (generic,
parallel and degenerate pairs); the bytes on hand-made capsules (none, far, zero radius, one cutting a known ray
set, the smallest rho of two, rho clamped); the Jacobi's first step from zero (E_dir / pi exactly), the closed
box's fixed point e / (pi (1 - rho)), and on the twin's real table (twin_assets/lite/lite_nodes.bin) the iterates
against a Krylov (GMRES) solve of (I - diag(hh / K) W) B = E_dir / pi, which fixes pi / K and 1 / pi; a node whose
every ray a black capsule blocks gathers nothing.
GPU: the device kernels on the real table against the oracle (bytes equal but where FP32 and FP64 distances
straddle the radius; Jacobi steps on the device's bytes to 1e-5), warm starts, and the argument checks.
"""
import os

import numpy as np
import pytest

import oracle
from vtwin import lib, scenes, twin

ROOT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "twin_assets")
needs_lite = pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "lite", "lite_nodes.bin")),
                                reason="twin_assets/lite missing: python tests/golden/fetch_twin_assets.py")


def _byte(rho):
    """round(255 rho) as binary32 arithmetic rounds it (half to even)"""
    return int(np.rint(np.float32(255.0) * np.float32(rho)))


def _seg_seg_numeric(p0, p1, q0, q1):
    from scipy.optimize import minimize
    f = lambda x: np.linalg.norm((p0 + (p1 - p0) * x[0]) - (q0 + (q1 - q0) * x[1]))
    best = np.inf
    for s0 in (0.0, 0.5, 1.0):
        for t0 in (0.0, 0.5, 1.0):
            r = minimize(f, [s0, t0], bounds=[(0, 1), (0, 1)], method="L-BFGS-B", options=dict(ftol=1e-15, gtol=1e-12))
            best = min(best, r.fun)
    return best


def _two_planes(nx=6, ny=5, K=8, gap=0.1, seed=3):
    """nodes on z = 0 (facing +z) and z = gap (facing -z); every ray of a node hits a random node of the other
    plane; meshes 0 and 1."""
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.linspace(0, 0.1, nx), np.linspace(0, 0.08, ny))
    P = np.stack([xs.ravel(), ys.ravel()], axis=1)
    m = P.shape[0]
    n = 2 * m
    a = np.zeros((n, 4), np.float32)
    a[:m, 0:2], a[m:, 0:2], a[m:, 2], a[:, 3] = P, P, gap, 1e-4
    nn = np.zeros((n, 4), np.float32)
    nn[:m, 2], nn[m:, 2] = 1.0, -1.0
    nn[m:, 3] = np.array([1], np.uint32).view(np.float32)[0]
    g = np.zeros((n, K), np.uint16)
    g[:m] = rng.integers(m, n, (m, K))
    g[m:] = rng.integers(0, m, (m, K))
    return dict(nodes=n, K=K, gather=g, node_a=a, node_n=nn, hh=np.array([[0.5, 0.4, 0.3, 0], [0.6, 0.6, 0.6, 0]], np.float32))


def test_oracle_seg_seg_matches_a_numerical_minimiser():
    rng = np.random.default_rng(8)
    worst = 0.0
    cases = []
    for _ in range(60):
        cases.append(rng.normal(size=(4, 3)))
    for _ in range(10):                       # parallel and anti-parallel
        p0, d, q0 = rng.normal(size=(3, 3))
        cases.append(np.array([p0, p0 + d, q0, q0 + d * rng.uniform(-2, 2)]))
    for _ in range(10):                       # a point against a segment, both ways
        p0, q0, q1 = rng.normal(size=(3, 3))
        cases.append(np.array([p0, p0, q0, q1]))
        cases.append(np.array([q0, q1, p0, p0]))
    for c in cases:
        cap = np.zeros((1, 8), np.float32)
        cap[0, 0:3], cap[0, 4:7], cap[0, 3], cap[0, 7] = c[2], c[3], 0.0, 0.5
        c = c.astype(np.float32).astype(np.float64)
        # the oracle's distance through gatherOcc's bookkeeping: a one-node, one-ray table from c[0] (lifted by
        # 1 mm along the normal, here 0) to the node at c[1]
        tbl = dict(nodes=2, K=4, gather=np.array([[1, 1, 1, 1], [0xffff] * 4], np.uint16),
                   node_a=np.array([[*c[0], 0], [*c[1], 0]], np.float32), node_n=np.zeros((2, 4), np.float32),
                   hh=np.zeros((1, 4), np.float32))
        _, dist = oracle.baked_occ(tbl, cap, with_dist=True)
        want = _seg_seg_numeric(c[0], c[1], cap[0, 0:3].astype(np.float64), cap[0, 4:7].astype(np.float64))
        worst = max(worst, abs(dist[0, 0] - want))
    print("segment distance: max |oracle - minimiser| %.2e over %d pairs" % (worst, len(cases)))
    assert worst < 1e-7


def test_oracle_baked_occ_pins():
    t = _two_planes()
    n, K = t["nodes"], t["K"]
    assert np.all(oracle.baked_occ(t, np.zeros((0, 8))) == 255)
    far = np.array([[5, 5, 5, 0.01, 5, 6, 5, 0.2]], np.float32)
    assert np.all(oracle.baked_occ(t, far) == 255)
    # a capsule along y through the gap's middle, over x in [0.045, 0.055]: it cuts the rays whose segment
    # passes within r of the line; here r = 0 cuts nothing, r large cuts every ray
    line = np.array([[0.05, -1, 0.05, 0.0, 0.05, 1, 0.05, 0.3]], np.float32)
    assert np.all(oracle.baked_occ(t, line) == 255)
    big = line.copy()
    big[0, 3] = 10.0
    occ = oracle.baked_occ(t, big)
    assert np.all(occ == _byte(0.3))
    # r = 2 cm: a ray is cut when its midpoint's x is within 2 cm of 0.05 (every ray crosses z = 0.05 at its
    # midpoint, and the capsule's axis runs along y there), computed by hand from the node positions
    mid = big.copy()
    mid[0, 3] = 0.02
    occ = oracle.baked_occ(t, mid)
    a = t["node_a"].astype(np.float64)
    xm = 0.5 * (a[:, None, 0] + a[t["gather"], 0])
    zs = a[:, 2] + 1e-3 * t["node_n"][:, 2]
    # the segment from the lifted node to the hit node: at z = 0.05 its x is the interpolated value
    zh = a[t["gather"], 2]
    s = (0.05 - zs[:, None]) / (zh - zs[:, None])
    x = a[:, None, 0] + (a[t["gather"], 0] - a[:, None, 0]) * s
    # closest distance of the segment to the line x = 0.05, z = 0.05 (the line along y is infinite here)
    dz = zh - zs[:, None]
    dx = a[t["gather"], 0] - a[:, None, 0]
    dist = np.abs((x - 0.05)) * np.abs(dz) / np.hypot(dx, dz)
    want = np.where(dist < 0.02, _byte(0.3), 255)
    clear = np.abs(dist - 0.02) > 1e-9
    assert np.array_equal(occ[clear], want[clear]) and 0 < np.mean(occ == 255) < 1
    del xm
    # two capsules: the smaller byte wins; rho is clamped to [0, 1]
    two = np.concatenate([big, big])
    two[1, 7] = 0.1
    assert np.all(oracle.baked_occ(t, two) == _byte(0.1))
    clamp = big.copy()
    clamp[0, 7] = -0.5
    assert np.all(oracle.baked_occ(t, clamp) == 0)
    clamp[0, 7] = 3.0
    assert np.all(oracle.baked_occ(t, clamp) == 255)
    # a miss never carries a capsule's byte
    t2 = dict(t, gather=t["gather"].copy())
    t2["gather"][:, 0] = 0xffff
    assert np.all(oracle.baked_occ(t2, big)[:, 0] == 255)


def test_oracle_baked_jacobi_closed_forms():
    t = _two_planes()
    n, K = t["nodes"], t["K"]
    rng = np.random.default_rng(2)
    e = np.zeros((n, 4), np.float32)
    e[:, 0:3] = rng.uniform(0.1, 2.0, (n, 3))
    occ = np.full((n, K), 255, np.uint8)
    b1, ei = oracle.baked_jacobi(t, occ, e, np.zeros((n, 4), np.float32), 1)
    assert np.array_equal(ei, np.zeros_like(ei)) and np.allclose(b1[:, 0:3], e[:, 0:3] / np.pi, rtol=1e-7)
    # closed box: every ray hits, t = 1, one albedo rho, uniform E_dir = e0: B* = e0 / (pi (1 - rho))
    rho, e0 = 0.6, 1.7
    tb = dict(t, hh=np.array([[rho] * 3 + [0]] * 2, np.float32))
    eu = np.zeros((n, 4), np.float32)
    eu[:, 0:3] = e0
    b, ei = oracle.baked_jacobi(tb, occ, eu, np.zeros((n, 4), np.float32), 120)
    assert np.allclose(b[:, 0:3], e0 / (np.pi * (1 - rho)), rtol=1e-6)
    assert np.allclose(ei[:, 0:3], np.pi * e0 / (np.pi * (1 - rho)), rtol=1e-6)
    # a black capsule that blocks every ray of the plane z = 0 (any ray crosses z = gap / 2): those nodes gather 0
    blk = np.array([[-1, 0.04, 0.05, 0.2, 2, 0.04, 0.05, 0.0]], np.float32)
    ob = oracle.baked_occ(t, blk)
    assert np.all(ob == 0)
    _, ei = oracle.baked_jacobi(t, ob, e, b1, 1)
    assert np.all(ei == 0)


@needs_lite
def test_oracle_baked_jacobi_on_the_twin_table_meets_a_krylov_solve():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    L = twin.lite_nodes(ROOT)
    n, K = L["count"], L["K"]
    r = scenes.Rng(41)
    nm = int(L["mesh"].max()) + 1
    hh = r.uniform(0.2, 0.8, nm * 3).reshape(nm, 3)
    t = twin.baked_table(L, hh)
    e = np.zeros((n, 4), np.float32)
    e[:, 0:3] = r.uniform(0.0, 3.0, n * 3).reshape(n, 3)
    occ = np.full((n, K), 255, np.uint8)
    occ[:, ::7] = 128                                       # some partial transmission
    hits = L["hits"].astype(np.int64)
    ok = hits != 0xFFFF
    rows = np.repeat(np.arange(n), K).reshape(n, K)[ok]
    W = sp.csr_matrix(((occ[ok] / 255.0), (rows, hits[ok])), shape=(n, n))
    b, ei = oracle.baked_jacobi(t, occ, e, np.zeros((n, 4), np.float32), 160)
    hhn = t["hh"][L["mesh"], 0:3].astype(np.float64)
    worst = 0.0
    for c in range(3):
        A = sp.identity(n, format="csr") - sp.diags(hhn[:, c] / K) @ W
        rhs = e[:, c].astype(np.float64) / np.pi
        x, info = spl.gmres(A, rhs, rtol=1e-13, atol=0.0, restart=60, maxiter=200)
        assert info == 0 and np.linalg.norm(A @ x - rhs) <= 1e-11 * np.linalg.norm(rhs)
        worst = max(worst, np.abs(b[:, c] - x).max() / np.abs(x).max())
    print("baked Jacobi, 160 steps vs the GMRES solve: max relative difference %.2e (%d nodes, %.1f %% misses)" % (
        worst, n, 100 * (1 - ok.mean())))
    assert worst < 1e-5


# ---------------------------------------------------------------------------- GPU

@pytest.fixture(scope="module")
def ctx():
    c = lib.Context(0)
    yield c
    c.close()


def _caps(seed=5):
    """six capsules among the twin's nodes: a few cm long, 4-12 mm thick, mean albedos 0.1-0.6"""
    L = twin.lite_nodes(ROOT)
    c = L["pos"].astype(np.float64).mean(axis=0)
    rng = np.random.default_rng(seed)
    out = np.zeros((6, 8), np.float32)
    for i in range(6):
        a = c + rng.normal(size=3) * 0.02
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        out[i, 0:3], out[i, 4:7] = a, a + d * rng.uniform(0.03, 0.08)
        out[i, 3], out[i, 7] = rng.uniform(0.004, 0.012), rng.uniform(0.1, 0.6)
    return out


@pytest.mark.gpu
@needs_lite
def test_gpu_baked_table_matches_oracle(ctx):
    L = twin.lite_nodes(ROOT)
    n, K = L["count"], L["K"]
    r = scenes.Rng(43)
    nm = int(L["mesh"].max()) + 1
    t = twin.baked_table(L, r.uniform(0.2, 0.8, nm * 3).reshape(nm, 3))
    e = np.zeros((n, 4), np.float32)
    e[:, 0:3] = r.uniform(0.0, 3.0, n * 3).reshape(n, 3)
    print("baked table: %d nodes x %d rays" % (n, K), flush=True)
    ctx.baked_upload(t)
    assert np.all(ctx.baked_read(2) == 255) and np.all(ctx.baked_read(0) == 0)
    caps = _caps()
    ctx.baked_occlusion(caps)
    occ = ctx.baked_read(2)
    occ_o, dist = oracle.baked_occ(t, caps, with_dist=True)
    diff = occ != occ_o
    print("gatherOcc: %d of %d bytes differ (closest |dist - r| among them %.2e); %.2f %% of the rays cut" % (
        diff.sum(), diff.size, np.abs(dist[diff]).max() if diff.any() else 0.0, 100 * np.mean(occ_o < 255)))
    assert np.mean(occ_o < 255) > 0.005
    assert np.all(np.abs(dist[diff]) < 1e-5)
    ctx.baked_jacobi(e, 6)
    b6 = ctx.baked_read(0)
    ei = ctx.baked_read(1)
    bo, eio = oracle.baked_jacobi(t, occ, e, np.zeros((n, 4), np.float32), 6)
    rel = np.abs(b6 - bo).max() / np.abs(bo).max()
    print("baked Jacobi, 6 steps: max relative difference %.2e" % rel)
    assert rel < 1e-5 and np.abs(ei - eio).max() <= 1e-5 * np.abs(eio).max()
    # warm start: 6 more steps without a new E_dir continue from B; 12 in one call from zero agree bit for bit
    ctx.baked_jacobi(None, 6)
    b12 = ctx.baked_read(0)
    ctx.baked_upload(t)
    ctx.baked_occlusion(caps)
    ctx.baked_jacobi(e, 12)
    assert np.array_equal(ctx.baked_read(0), b12) and not np.array_equal(b12, b6)
    with pytest.raises(lib.VTError):
        ctx.baked_occlusion(np.zeros((17, 8), np.float32))
    bad = dict(t, gather=t["gather"].copy())
    bad["gather"][0, 0] = n
    with pytest.raises(lib.VTError):
        ctx.baked_upload(bad)
