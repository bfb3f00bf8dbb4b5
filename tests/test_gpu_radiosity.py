"""Node radiosity on the GPU against the oracle: the reference's sparse gather Jacobi step, the dense form
factors, and the dense solves (FP32, BF16 on MFMA with FP32 refinement, FP64)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from vtwin import lib, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = lib.Context(0)
    yield c
    c.close()


def _tables(N=1500, A=1100, K=64, nbrk=16, nmat=5, bins=32, seed=17):
    r = scenes.Rng(seed)
    nrm = np.stack([r.uniform(-1, 1, N), r.uniform(-1, 1, N), r.uniform(-1, 1, N)], 1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    normals = np.zeros((N, 4), np.float32)
    normals[:, 0:3] = nrm
    active = np.sort(r.int(0, N - 1, n=A * 2))
    active = np.unique(active)[:A].astype(np.uint32)
    A = active.size
    gt = r.int(0, N - 1, n=N * K).astype(np.uint32)
    gt[r.u(N * K) < 0.2] = 0xFFFFFFFF
    node_mat = r.int(0, nmat - 1, n=N).astype(np.uint32)
    alb = np.zeros((nmat, bins, 2, 4), np.float32)
    alb[:, :, 0, 0:3] = r.uniform(0.0, 0.3, nmat * bins * 3).reshape(nmat, bins, 3)
    alb[:, :, 1, 0:3] = r.uniform(0.1, 0.6, nmat * bins * 3).reshape(nmat, bins, 3)
    hh = np.zeros((nmat, 4), np.float32)
    hh[:, 0:3] = r.uniform(0.2, 0.8, nmat * 3).reshape(nmat, 3)
    nbr = r.int(0, A - 1, n=A * nbrk).astype(np.uint32).reshape(A, nbrk)
    nbr_w = r.uniform(0.1, 1.0, A * nbrk).astype(np.float32).reshape(A, nbrk)
    cut = r.int(1, nbrk, n=A)
    for k in range(A):
        nbr_w[k, cut[k]:] = 0
    st = np.zeros((N, 16), np.float32)
    st[:, 3] = r.uniform(0.5, 1.0, N)           # capsule visibility V
    st[:, 4:7] = r.uniform(0, 2, N * 3).reshape(N, 3)
    st[:, 12:15] = r.uniform(0, 3, N * 3).reshape(N, 3)
    t = dict(n_static=N, n_active=A, gather_k=K, nbr_k=nbrk, n_mat=nmat, albedo_bins=bins, active=active, gtable=gt,
             normals=normals, node_mat=node_mat, albedo=alb, hh=hh, nbr=nbr, nbr_w=nbr_w)
    return t, st


def _oracle_jacobi(t, st, iters):
    """The oracle's gather + update steps on the same tables: the new state and the last step's aux."""
    want = st.copy()
    N, A, K = t["n_static"], t["n_active"], t["gather_k"]
    verts = np.zeros((N, 8), np.float32)
    verts[:, 4:7] = t["normals"][:, 0:3]
    spec = np.ascontiguousarray(t["albedo"][:, :, 0, 0:3]).reshape(-1)
    diff = np.ascontiguousarray(t["albedo"][:, :, 1, 0:3]).reshape(-1)
    hh = np.ascontiguousarray(t["hh"][:, 0:3]).reshape(-1)
    aux_o = np.zeros((A, 4), np.float32)
    nbr, nbr_w = np.ascontiguousarray(t["nbr"]), np.ascontiguousarray(t["nbr_w"])
    slot = 1
    for _ in range(iters):
        oracle.lib().or_jacobi_gather(A, oracle.ptr(t["active"]), oracle.ptr(t["gtable"]), K, oracle.ptr(verts),
                                      oracle.ptr(t["node_mat"]), oracle.ptr(spec), oracle.ptr(diff), 32, oracle.ptr(hh),
                                      oracle.ptr(want), slot, oracle.ptr(aux_o))
        oracle.lib().or_jacobi_update(A, oracle.ptr(t["active"]), oracle.ptr(nbr), oracle.ptr(nbr_w), t["nbr_k"],
                                      oracle.ptr(aux_o), oracle.ptr(t["node_mat"]), oracle.ptr(hh), oracle.ptr(want), 3 - slot)
        slot = 3 - slot
    return want, aux_o


def test_sparse_jacobi_matches_oracle(ctx):
    t, st = _tables()
    got, aux = ctx.jacobi_sparse(t, st, 3)
    want, aux_o = _oracle_jacobi(t, st, 3)
    assert np.allclose(aux[:, 0:3], aux_o[:, 0:3], rtol=2e-5, atol=1e-6)
    assert np.allclose(got, want, rtol=2e-5, atol=1e-6)


def test_sparse_jacobi_on_the_baked_table(ctx):
    """Real sparse input: the lite renderer's baked gather table of the twin (12,976 nodes x 256 hits, 4.3 % misses)
    with its node normals and mesh ids, smoothing neighbours from its node positions; materials and the starting
    state seeded. Every node active; 4 steps against the oracle."""
    import os
    from vtwin import twin
    root = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "twin_assets")
    if not os.path.exists(os.path.join(root, "lite", "lite_nodes.bin")):
        pytest.skip("twin_assets/lite missing: python tests/golden/fetch_twin_assets.py")
    L = twin.lite_nodes(root)
    n, K = L["count"], L["K"]
    meshes, node_mat = np.unique(L["mesh"], return_inverse=True)
    nmat, bins = meshes.size, 32
    r = scenes.Rng(23)
    normals = np.zeros((n, 4), np.float32)
    normals[:, 0:3] = L["nrm"]
    alb = np.zeros((nmat, bins, 2, 4), np.float32)
    alb[:, :, 0, 0:3] = r.uniform(0.0, 0.2, nmat * bins * 3).reshape(nmat, bins, 3)
    alb[:, :, 1, 0:3] = r.uniform(0.1, 0.7, nmat * bins * 3).reshape(nmat, bins, 3)
    hh = np.zeros((nmat, 4), np.float32)
    hh[:, 0:3] = r.uniform(0.2, 0.8, nmat * 3).reshape(nmat, 3)
    nbr, nbr_w = lib.node_neighbours(L["pos"], L["nrm"], 16, 0.014, 0.00385)
    st = np.zeros((n, 16), np.float32)
    st[:, 3] = r.uniform(0.5, 1.0, n)
    st[:, 4:7] = r.uniform(0, 2, n * 3).reshape(n, 3)
    st[:, 12:15] = r.uniform(0, 3, n * 3).reshape(n, 3)
    t = dict(n_static=n, n_active=n, gather_k=K, nbr_k=16, n_mat=nmat, albedo_bins=bins,
             active=np.arange(n, dtype=np.uint32), gtable=L["gtable"], normals=normals,
             node_mat=node_mat.astype(np.uint32), albedo=alb, hh=hh, nbr=nbr, nbr_w=nbr_w)
    print("baked table: %d nodes x %d hits, %d materials" % (n, K, nmat), flush=True)
    got, aux = ctx.jacobi_sparse(t, st, 4)
    want, aux_o = _oracle_jacobi(t, st, 4)
    assert np.allclose(aux[:, 0:3], aux_o[:, 0:3], rtol=2e-5, atol=1e-6)
    assert np.allclose(got, want, rtol=2e-5, atol=1e-6)


def _dense_problem(n, r, seed=5):
    g = scenes.Rng(seed)
    # a form-factor-like matrix: distance falloff between random points, rows scaled to sum 0.9
    p = np.stack([g.uniform(-1, 1, n), g.uniform(-1, 1, n), g.uniform(-1, 1, n)], 1)
    d2 = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1) + 0.05
    F = 1.0 / d2
    np.fill_diagonal(F, 0)
    F = (F / F.sum(1, keepdims=True) * 0.9).astype(np.float32)
    e = np.zeros((n, r), np.float32)
    idx = g.int(0, n - 1, n=32)
    e[idx] = g.uniform(5, 20, 32 * r).reshape(32, r)
    return F, e


# the solution is returned in binary32, so the FP64 solve is checked at binary32 rounding (6e-8)
@pytest.mark.parametrize("method,tol,err,inner", [(lib.VT_DENSE_FP64, 1e-12, 1e-7, 0), (lib.VT_DENSE_FP32, 1e-6, 2e-5, 0),
                                                  (lib.VT_DENSE_BF16_REFINE, 1e-6, 1e-5, 24),
                                                  (lib.VT_DENSE_BF16_REFINE, 1e-6, 1e-5, 0)])
def test_dense_solves_match_direct_solve(ctx, method, tol, err, inner):
    """FP64 and FP32 (MFMA) Jacobi, BF16 MFMA sweeps with FP32-residual refinement at a fixed and at the
    adaptive inner count, against LAPACK on a 1024-node form-factor-like system with 48 right-hand sides."""
    n, r = 1024, 48
    F, e = _dense_problem(n, r)
    want = np.linalg.solve(np.eye(n) - F.astype(np.float64), e.astype(np.float64))
    ctx.dense_upload(F, e)
    x, res = ctx.dense_solve(n, r, method, tol=tol, max_sweeps=3000, inner_sweeps=inner)
    rel = np.linalg.norm(x - want) / np.linalg.norm(want)
    print("method %d: sweeps %d, refinements %d, FP64 residual %.2e, error vs direct %.2e" % (method, res["sweeps"], res["refinements"], res["residual"], rel))
    assert res["residual"] <= max(tol, 2e-7) * 1.5
    assert rel < err


def test_form_factors_match_cpu(ctx):
    nodes, sc = scenes.c4_nodes(16, 16)          # 512 nodes on the dome, visibility through the table box
    n = nodes.shape[0]
    ctx.upload_scene(sc)
    e = scenes.c4_emission(n, configs=1)
    ctx.dense_formfactors(nodes, 0.9, e)
    F = ctx.dense_F(n)
    p, a, nr = nodes[:, 0:3].astype(np.float64), nodes[:, 3].astype(np.float64), nodes[:, 4:7].astype(np.float64)
    d = p[None, :, :] - p[:, None, :]
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, 1.0)
    dn = d / np.sqrt(r2)[..., None]
    ci = (nr[:, None, :] * dn).sum(-1)
    cj = -(nr[None, :, :] * dn).sum(-1)
    G = np.where((ci > 0) & (cj > 0), ci * cj * a[None, :] / (math.pi * r2), 0.0)
    np.fill_diagonal(G, 0)
    # visibility by the oracle's any-hit traversal of the same scene
    ii, jj = np.nonzero(G > 0)
    o = p[ii] + nr[ii] * 1e-4
    dd = p[jj] - o
    ln = np.linalg.norm(dd, axis=1)
    rays = np.zeros((ii.size, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 4:7] = dd / ln[:, None]
    rays[:, 7] = ln - 1e-4
    hits = oracle.trace(sc["nodes"], sc["tris"], rays, any_hit=True)
    V = np.zeros_like(G)
    V[ii, jj] = (hits[:, 1] < 0)
    G = G * V
    want = G * (0.9 / np.maximum(1.0, G.sum(1)))[:, None]
    close = np.isclose(F, want, rtol=2e-4, atol=1e-9)
    print("form factors: %d of %d entries differ (visibility at grazing pairs)" % ((~close).sum(), F.size))
    assert np.mean(~close) < 1e-3
    assert np.allclose(F.sum(1), want.sum(1), rtol=5e-3)
