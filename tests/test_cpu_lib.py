"""Host-side checks of the C-ABI library that need no GPU: it loads, exports every entry point include/vt.h
declares, and its native BVH builder reproduces the reference's (the oracle is pinned to gpu/bvh.js)."""
import json
import os
import re

import numpy as np
import pytest

import oracle
from vtwin import lib, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_vectors.json")))


def test_library_exports_every_declared_entry_point():
    L = lib.load()
    hdr = open(os.path.join(ROOT, "include", "vt.h")).read()
    declared = sorted(set(re.findall(r"^(?:int|void|const char \*)\s*(vt_\w+)\(", hdr, re.M)))
    assert declared, "no declarations parsed"
    assert sorted(lib.EXPORTS) == declared
    for name in declared:
        assert hasattr(L, name), name


@pytest.mark.parametrize("case", [c["name"] for c in GOLD["bvh"]])
def test_native_bvh_equals_reference_golden(case):
    g = next(c for c in GOLD["bvh"] if c["name"] == case)
    pos = np.frombuffer(bytes.fromhex(g["pos"]), np.float32)
    idx = np.frombuffer(bytes.fromhex(g["idx"]), np.uint32)
    nodes, order = lib.build_bvh(pos, idx)
    assert np.array_equal(nodes.view(np.uint32), np.frombuffer(bytes.fromhex(g["nodes"]), np.uint32))
    if g["ntri"]:
        assert np.array_equal(order, np.frombuffer(bytes.fromhex(g["order"]), np.uint32))


def test_native_bvh_equals_oracle_on_the_c1_scene():
    sc, _ = scenes.c1()
    o = oracle.build_bvh(sc["pos"], sc["idx"])
    assert np.array_equal(o["nodes"].view(np.uint32), sc["nodes"].view(np.uint32))
    assert np.array_equal(o["order"], sc["order"])


def test_build_bvh_rejects_bad_input():
    with pytest.raises(lib.VTError):
        lib.build_bvh(np.zeros(9, np.float32), np.array([0, 1, 5], np.uint32))   # vertex id past the end


def test_mulberry32_stream_matches_reference_golden():
    for seed in (0, 1, 1001, 0xDEADBEEF, 4294967295):
        want = np.array(GOLD["rng"]["mulberry32_%d" % seed], dtype=np.float64)
        got = scenes.Rng(seed).u(64)
        assert np.array_equal(np.round(got * 4294967296.0), want)


def test_scenes_are_deterministic_and_sized():
    a, ca = scenes.c1()
    b, cb = scenes.c1()
    assert np.array_equal(a["tris"], b["tris"]) and np.array_equal(ca[0]["lens"], cb[0]["lens"])
    assert a["tris"].size // 16 == 19900 + 12 + 8
    assert a["emit"].size // 16 == 8
    assert ca[0]["width"] == 640 and ca[0]["spp"] == 16 and ca[0]["max_depth"] == 3
