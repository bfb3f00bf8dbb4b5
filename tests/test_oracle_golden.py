"""The oracle against the reference's own outputs (tests/golden/ref_vectors.json, made by make_golden.mjs
running control/rng.js, gpu/bvh.js and gpu/shade_ref.js under Node)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")))


def _f32(h):
    return np.frombuffer(bytes.fromhex(h), dtype=np.float32).copy()


def _u32(h):
    return np.frombuffer(bytes.fromhex(h), dtype=np.uint32).copy()


@pytest.mark.parametrize("seed", [0, 1, 1001, 0xDEADBEEF, 4294967295])
def test_mulberry32_bit_exact(seed):
    want = np.array(GOLD["rng"]["mulberry32_%d" % seed], dtype=np.float64)
    got = np.zeros(64)
    oracle.lib().or_mulberry32_fill(seed, oracle.ptr(got), 64)
    assert np.array_equal(np.round(got * 4294967296.0), want)


@pytest.mark.parametrize("case", [c["name"] for c in GOLD["bvh"]])
def test_build_bvh_bit_exact(case):
    g = next(c for c in GOLD["bvh"] if c["name"] == case)
    pos, idx = _f32(g["pos"]), _u32(g["idx"])
    b = oracle.build_bvh(pos, idx)
    assert b["node_count"] == g["nodeCount"]
    assert b["max_depth"] == g["maxDepth"]
    assert np.array_equal(b["nodes"].view(np.uint32), _u32(g["nodes"]))
    assert np.array_equal(b["order"], _u32(g["order"])) or g["ntri"] == 0
    if g["ntri"]:
        T = oracle.pack_triangles(pos, idx, b["order"])
        want = _f32(g["tris"]).reshape(-1, 16)
        got = T.reshape(-1, 16)
        geo = [0, 1, 2, 4, 5, 6, 8, 9, 10]
        assert np.array_equal(got[:, geo], want[:, geo])
        sah = oracle.lib().or_sah_cost(oracle.ptr(b["nodes"]), b["node_count"],
                                       oracle.ptr(b["bmin"].astype(np.float32)), oracle.ptr(b["bmax"].astype(np.float32)), 1.0, 1.0)
        assert sah == pytest.approx(g["sah"], rel=1e-12)


def _check_hits(got, want):
    for k, (h, w) in enumerate(zip(got, want)):
        if w[0] is None:
            assert h[1] == -1, k
        else:
            assert int(h[1]) == w[1], k
            assert h[0] == w[0] and h[2] == w[2] and h[3] == w[3], k


@pytest.mark.parametrize("case", [c["name"] for c in GOLD["bvh"] if c["ntri"] > 0])
def test_traverse_matches_reference(case):
    g = next(c for c in GOLD["bvh"] if c["name"] == case)
    pos, idx = _f32(g["pos"]), _u32(g["idx"])
    b = oracle.build_bvh(pos, idx)
    T = oracle.pack_triangles(pos, idx, b["order"])
    r6 = _f32(g["rays"]).reshape(-1, 6)
    rays = np.zeros((r6.shape[0], 8), np.float32)
    rays[:, 0:3], rays[:, 4:7] = r6[:, 0:3], r6[:, 3:6]
    rays[:, 7] = np.inf
    _check_hits(oracle.trace(b["nodes"], T, rays), g["closest"])
    _check_hits(oracle.brute(T, g["ntri"], rays), g["brute"])
    rays[:, 7] = 4.0
    _check_hits(oracle.trace(b["nodes"], T, rays, any_hit=True), g["any"])


def test_shade_ref_functions():
    L = oracle.lib()
    pr = GOLD["shade"]["profile"]
    kn, p = np.array(pr["knots"]), np.array(pr["p"])
    for t, v in zip(pr["theta"], pr["value"]):
        assert L.or_profile(t, oracle.ptr(kn), oracle.ptr(p), len(kn)) == pytest.approx(v, rel=1e-15, abs=1e-300)
    for c in GOLD["shade"]["capsule"]:
        a = [np.array(c[k], dtype=np.float64) for k in ("p", "n", "a", "b")]
        got = L.or_capsule_occ(*[oracle.ptr(x) for x in a], c["r"])
        assert got == pytest.approx(c["value"], rel=1e-12, abs=1e-15)
    gl = GOLD["shade"]["gaussLegendre12"]
    x, w = np.zeros(12), np.zeros(12)
    L.or_gauss_legendre(12, oracle.ptr(x), oracle.ptr(w))
    assert np.allclose(x, gl["x"], rtol=0, atol=1e-15) and np.allclose(w, gl["w"], rtol=0, atol=1e-15)
    for c in GOLD["shade"]["disc"]:
        a = [np.array(c[k], dtype=np.float64) for k in ("x", "n", "c", "m")]
        got = L.or_disc_irradiance(*[oracle.ptr(v) for v in a], c["a"], c["L"], 24, 48)
        assert got == pytest.approx(c["value"], rel=1e-12)


def test_rng_normal_matches_v8_to_last_ulps():
    g = GOLD["rng"]["rng1003"]
    # re-run the same Rng(1003) call sequence: 8 uniform, 8 int, 9 normal, 8 pick
    s = np.zeros(200)
    oracle.lib().or_mulberry32_fill(1003, oracle.ptr(s), 200)
    it = iter(s)
    uni = [-2 + 7 * next(it) for _ in range(8)]
    assert uni == pytest.approx(g["uniform"], rel=0, abs=0)
    ints = [int(np.floor(next(it) * 16)) for _ in range(8)]
    assert ints == g["int"]
    normals, spare = [], None
    for _ in range(9):
        if spare is not None:
            normals.append(1 + 2 * spare)
            spare = None
            continue
        u = next(it)
        while u <= 1e-300:
            u = next(it)
        v = next(it)
        r = np.sqrt(-2 * np.log(u))
        spare = r * np.sin(2 * np.pi * v)
        normals.append(1 + 2 * r * np.cos(2 * np.pi * v))
    assert normals == pytest.approx(g["normal"], rel=1e-12)
    picks = [[1, 4, 16, 64][int(np.floor(next(it) * 4))] for _ in range(8)]
    assert picks == g["pick"]
