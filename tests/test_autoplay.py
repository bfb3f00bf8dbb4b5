"""Seeded autoplay (vtwin/autoplay.py; control/scripts.js makeScript / tipSpeeds and the clamp half of
control/kinematics.js). The reference's scripts.js cannot run in the image's Node 12 (it uses ?? and ?.), so no
trajectory of its own is a golden vector; this planner draws its numbers in its own order. Pinned instead:
the minimum-jerk profile by its closed-form properties (ends, symmetry, zero end velocity, peak = 1.875 x mean),
coneClamp by construction (inside unchanged; outside onto the cone along the same azimuth), aimAt against the
forward kinematics, workspaceCentre on camera axes that meet in one point; and on the twin's scene the properties
the reference's planner promises: the same state for the same seed, five labelled phases, every hinge target
inside 85 % of its port's cone and the insertion limits and above the bench, mean move speeds 8-50 mm/s and no
sampled tip speed above 100 mm/s, the scope inside its cone.
"""
import math
import os

import numpy as np
import pytest

from vtwin import autoplay, twin

ASSETS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "twin_assets")
needs_assets = pytest.mark.skipif(not os.path.exists(os.path.join(ASSETS, "scene.json")),
                                  reason="twin_assets/ missing (python tests/golden/fetch_twin_assets.py)")


def test_min_jerk_profile():
    s = autoplay.min_jerk
    assert s(0.0) == 0.0 and s(1.0) == 1.0 and s(-1.0) == 0.0 and s(2.0) == 1.0
    assert abs(s(0.5) - 0.5) < 1e-15
    u = np.linspace(0, 1, 200001)
    v = np.array([s(x) for x in u])
    dv = np.gradient(v, u)
    assert abs(dv[0]) < 1e-6 and abs(dv[-1]) < 1e-6
    assert np.all(np.diff(v) >= 0)
    assert abs(dv.max() - autoplay.MJ_PEAK) < 1e-6                 # peak speed = 1.875 x the mean (1)
    for x in (0.13, 0.5, 0.77):                                     # symmetric: s(u) + s(1 - u) = 1
        assert abs(s(x) + s(1 - x) - 1) < 1e-14


def test_cone_clamp_and_aim_at():
    rng = np.random.default_rng(4)
    cone = 30 * autoplay.DEG
    for _ in range(200):
        q = dict(yaw=rng.uniform(-1.2, 1.2), pitch=rng.uniform(-1.2, 1.2), roll=0.3, insertion=0.1)
        c, clamped = autoplay.cone_clamp(q, cone)
        ang = autoplay.cone_angle(q)
        assert clamped == (ang > cone + 1e-12)
        if not clamped:
            assert c == q
            continue
        assert abs(autoplay.cone_angle(c) - cone) < 1e-12
        # the same azimuth about the port axis
        az = lambda r: math.atan2(-math.sin(r["pitch"]), math.cos(r["pitch"]) * math.sin(r["yaw"]))
        assert abs(math.remainder(az(c) - az(q), 2 * math.pi)) < 1e-9
        assert c["roll"] == q["roll"] and c["insertion"] == q["insertion"]
    # yaw alone past the cone lands exactly on it
    c, _ = autoplay.cone_clamp(dict(yaw=0.9, pitch=0.0, roll=0.0, insertion=0.1), cone)
    assert abs(c["yaw"] - cone) < 1e-12 and abs(c["pitch"]) < 1e-12
    # aimAt inverts the forward kinematics
    R = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    R *= np.sign(np.linalg.det(R))
    port = dict(pivot=np.array([0.1, -0.2, 0.3]), R=R)
    for _ in range(50):
        h = port["pivot"] + rng.normal(size=3) * 0.1
        q = autoplay.aim_at(port, h, 0.4, 0.2)
        assert np.allclose(autoplay.hinge_of(port, q), h, atol=1e-12)


def test_workspace_centre_meets_camera_axes():
    c = np.array([0.05, -0.02, 0.12])
    cams = []
    for d in ([1, 0.2, -1], [-0.5, 1, -0.8], [0.3, -1, -0.6]):
        z = np.array(d, float) / np.linalg.norm(d)
        o = c - 0.2 * z
        x = np.cross([0, 0, 1.0], z)
        x /= np.linalg.norm(x)
        y = np.cross(z, x)
        T = np.eye(4)
        T[:3, 0], T[:3, 1], T[:3, 2], T[:3, 3] = x, y, z, o
        cams.append(dict(T_D_C=T.tolist()))
    got = autoplay.workspace_centre(dict(cameras=cams, bench=dict(z=0.0)))
    assert np.allclose(got, c, atol=1e-12)
    low = autoplay.workspace_centre(dict(cameras=cams, bench=dict(z=0.2)))
    assert np.allclose(low[:2], c[:2]) and low[2] == pytest.approx(0.23)


@pytest.fixture(scope="module")
def scene():
    return twin.load(ASSETS)[0]


@needs_assets
def test_task_set_is_seeded_and_phased(scene):
    a = autoplay.make_script("task_set", scene, seed=7)
    b = autoplay.make_script("task_set", scene, seed=7)
    c = autoplay.make_script("task_set", scene, seed=8)
    print("task_set: %.1f s, phases %s" % (a.duration, [(round(t0, 1), round(t1, 1), n) for t0, t1, n in a.labels]))
    assert [n for *_, n in a.labels] == ["reach", "grasp", "cross", "still_hold", "scope_reaim"]
    assert 30 < a.duration < 200
    ts = np.linspace(0, a.duration, 97)
    for t in ts:
        sa, sb = a.state_at(t), b.state_at(t)
        assert sa == sb or all(np.allclose(list(x.values())[1:], list(y.values())[1:])
                               for x, y in zip(sa["instruments"], sb["instruments"]))
    assert any(abs(a.state_at(t)["instruments"][0]["yaw"] - c.state_at(t)["instruments"][0]["yaw"]) > 1e-3
               for t in ts)
    # the still hold is still
    t0, t1 = next((t0, t1) for t0, t1, n in a.labels if n == "still_hold")
    s0, s1 = a.state_at(t0 + 0.01), a.state_at(t1 - 0.01)
    assert s0["instruments"] == s1["instruments"] and s0["scope"] == s1["scope"]
    with pytest.raises(ValueError):
        autoplay.make_script("juggle", scene)


@needs_assets
def test_task_set_limits(scene):
    s = autoplay.make_script("task_set", scene, seed=3)
    for tid, tool in s.tools.items():
        sp, p = tool["spec"], tool["port"]
        moves = [g for g in tool["track"].segs if g[1] > g[0] and not np.array_equal(g[2]["hinge"], g[3]["hinge"])]
        assert len(moves) >= 10
        for t0, t1, a, b, kind in moves:
            q = autoplay.aim_at(p, b["hinge"])
            assert autoplay.cone_angle(q) <= p["cone"] * autoplay.TARGET_CONE + 1e-9, (tid, kind)
            assert sp.get("min_insertion", 0) + 0.003 - 1e-12 <= q["insertion"] <= sp["max_insertion"] * 0.95 + 1e-12
            assert b["hinge"][2] >= float((scene.get("bench") or {}).get("z", 0.0)) + 0.008 - 1e-12
            mean = np.linalg.norm(b["hinge"] - a["hinge"]) / (t1 - t0)
            assert mean <= 0.050 + 1e-12, (tid, kind, mean)
    sp = autoplay.tip_speeds(s, scene, dt=0.01)
    for tid, v in sp.items():
        print("%s: tip speed max %.1f mm/s, moving mean %.1f mm/s" % (tid, v.max(), v[v > 1].mean()))
        assert v.max() <= 100.0 * 1.01
    cone = s.scope_port["cone"]
    for t in np.linspace(0, s.duration, 301):
        assert autoplay.cone_angle(s.state_at(t)["scope"]) <= cone + 1e-9
