"""Seeded autoplay of the twin: the pose source that drives vt_set_instances in end-to-end runs (control/scripts.js
makeScript / tipSpeeds :563-586 and its task_set plan :259-354; the clamp half of control/kinematics.js:
coneClamp :165-174, aimAt / inverseDirection :118-131, workspaceCentre :201-227, scopeDefault :152-159, spec :90-96).

A script is planned once from (scene, seed) into one track per body -- each instrument's jaw hinge (a point in the
world, D), roll and jaw opening, and the scope's joints -- made of minimum-jerk moves s(u) = 10u^3 - 15u^4 + 6u^5
(Flash and Hogan 1985) and holds; state_at(t) is then a pure function of t, and twin.poses_from_state(scene,
state_at(t)) gives the poses. The task set has the reference's five phases: reaches across the workspace (the two
hands alternating), grasp cycles (open wide, close, hold, reopen), crossings (each tip to the other side, at
different heights, so the shafts pass each other in the views), a still hold of everything, and scope re-aims.

The plan is this module's own: its random draws (mulberry32, control/rng.js, as scenes.Rng) come in its own order,
so a seed gives other trajectories than the reference's planner, whose scripts.js the image's Node 12 cannot run.
What the reference fixes for every script is kept and tested: tip speeds (minimum-jerk mean speed 8-50 mm/s per
move, peaks at most 100 mm/s), every hinge target inside 85 % of its port's cone and within the insertion limits,
the hinge above the bench, the scope inside 85 % of its cone, and the same state for the same (scene, seed, t).
"""
import bisect
import math

import numpy as np

from . import scenes, twin

DEG = math.pi / 180.0
MJ_PEAK = 1.875                       # peak speed of a minimum-jerk move / its mean speed
SPEC_DEFAULTS = dict(shaft_radius=0.0025, housing_length=0.010, jaw_length=0.016, jaw_radius=0.0012,
                     jaw_max=60 * DEG, max_insertion=0.33)
V_MEAN = (0.008, 0.050)               # m/s: a move's mean tip speed is drawn in this range
TARGET_CONE = 0.85                    # targets stay inside this fraction of the port's cone half-angle


def min_jerk(u):
    u = min(max(u, 0.0), 1.0)
    return u * u * u * (10.0 + u * (-15.0 + 6.0 * u))


class Track:
    """One body's plan: segments (t0, t1, a, b, kind) blending a -> b by min_jerk; the value holds before the first
    segment and after the last."""

    def __init__(self, value):
        self.segs = []
        self.starts = []
        self.v = value
        self.t = 0.0

    def _add(self, t0, t1, a, b, kind):
        self.segs.append((t0, t1, a, b, kind))
        self.starts.append(t0)
        self.v, self.t = b, t1
        return t1

    def move(self, t0, dur, value, kind="move"):
        return self._add(t0, t0 + dur, self.v, value, kind)

    def hold(self, t0, dur, kind="hold"):
        return self._add(t0, t0 + max(dur, 0.0), self.v, self.v, kind)

    def hold_until(self, t, kind="hold"):
        if self.t < t:
            self.hold(self.t, t - self.t, kind)

    def at(self, t, blend):
        if not self.segs:
            return self.v
        if t <= self.segs[0][0]:
            return self.segs[0][2]
        g = self.segs[bisect.bisect_right(self.starts, t) - 1]
        if t >= g[1]:
            return g[3]
        return blend(g[2], g[3], min_jerk((t - g[0]) / (g[1] - g[0])))

    def kind_at(self, t):
        k = bisect.bisect_right(self.starts, t) - 1
        if k >= 0 and t < self.segs[k][1]:
            return self.segs[k][4]
        return "idle"


def _blend_tool(a, b, w):
    return dict(hinge=a["hinge"] + (b["hinge"] - a["hinge"]) * w, roll=a["roll"] + (b["roll"] - a["roll"]) * w,
                jaw=a["jaw"] + (b["jaw"] - a["jaw"]) * w)


def _blend_joints(a, b, w):
    return {k: a[k] + (b[k] - a[k]) * w for k in ("yaw", "pitch", "roll", "insertion")}


# ---------------------------------------------------------------------------- kinematics (the clamp half)

def port_of(scene, pid):
    """A port's pivot, frame R_D_P, cone half-angle (rad; 30 deg when the scene gives none) and role."""
    p = next(q for q in scene["ports"] if q["id"] == pid)
    return dict(id=pid, pivot=np.array(p["pivot"], float), R=np.array(p["R_D_P"], float),
                cone=float(p.get("cone_half_deg", 30.0)) * DEG, role=p.get("role"))


def tool_spec(ins):
    sp = dict(SPEC_DEFAULTS, **(ins.get("spec") or {}))
    if ins.get("min_insertion") is not None and sp.get("min_insertion") is None:
        sp["min_insertion"] = float(ins["min_insertion"])
    return sp


def inverse_direction(R_D_P, d):
    """A shaft direction in the world -> (yaw, pitch) in the port frame (R_D_P Ry(yaw) Rx(pitch) e_z = d)."""
    v = R_D_P.T @ (np.asarray(d, float) / np.linalg.norm(d))
    return math.atan2(v[0], v[2]), -math.asin(min(max(v[1], -1.0), 1.0))


def aim_at(port, hinge, roll=0.0, jaw=0.0):
    v = np.asarray(hinge, float) - port["pivot"]
    yaw, pitch = inverse_direction(port["R"], v)
    return dict(yaw=yaw, pitch=pitch, roll=roll, insertion=float(np.linalg.norm(v)), jaw=jaw)


def cone_angle(q):
    """The angle between the shaft (yaw, pitch) and the port axis."""
    return math.acos(min(max(math.cos(q["pitch"]) * math.cos(q["yaw"]), -1.0), 1.0))


def cone_clamp(q, cone):
    """The joints with the shaft pulled back onto the cone of half-angle `cone` about the port axis, along the same
    azimuth; (q', clamped)."""
    cp = math.cos(q["pitch"])
    d = np.array([cp * math.sin(q["yaw"]), -math.sin(q["pitch"]), cp * math.cos(q["yaw"])])
    if math.acos(min(max(d[2], -1.0), 1.0)) <= cone + 1e-12:
        return dict(q), False
    l = math.hypot(d[0], d[1]) or 1.0
    e = np.array([d[0] / l * math.sin(cone), d[1] / l * math.sin(cone), math.cos(cone)])
    return dict(q, yaw=math.atan2(e[0], e[2]), pitch=-math.asin(min(max(e[1], -1.0), 1.0))), True


def workspace_centre(scene):
    """scene.workspace.centre, else the point nearest (least squares) to every tracking camera's optical axis, kept
    30 mm above the bench."""
    ws = (scene.get("workspace") or {}).get("centre")
    if ws:
        return np.array(ws, float)
    A, b = np.zeros((3, 3)), np.zeros(3)
    for c in scene["cameras"]:
        T = np.array(c["T_D_C"], float)
        o, z = T[:3, 3], T[:3, 2] / np.linalg.norm(T[:3, 2])
        M = np.eye(3) - np.outer(z, z)
        A += M
        b += M @ o
    p = np.linalg.solve(A, b)
    p[2] = max(p[2], float((scene.get("bench") or {}).get("z", 0.0)) + 0.03)
    return p


def scope_default(scene):
    df = (scene.get("scope") or {}).get("default") or {}
    p = port_of(scene, (scene.get("scope") or {}).get("port", "C"))
    yaw, pitch = df.get("yaw"), df.get("pitch")
    if yaw is None or pitch is None:
        yaw, pitch = inverse_direction(p["R"], df["dir_D"]) if df.get("dir_D") else (0.0, 0.0)
    return dict(yaw=float(yaw), pitch=float(pitch), roll=float(df.get("roll", 0.0)),
                insertion=float(df.get("insertion", 0.016)))


def hinge_of(port, q):
    R = twin.shaft_rotation(port["R"], q)
    return port["pivot"] + R[:, 2] * q["insertion"]


# ---------------------------------------------------------------------------- the task set

class TaskSet:
    """The task_set plan. reaches, grasps, crossings, still (s) and reaims size its phases (the reference's full set
    by default)."""

    def __init__(self, scene, seed=0, reaches=6, grasps=3, crossings=2, still=10.0, reaims=3):
        self.scene = scene
        self.seed = int(seed)
        self.rng = scenes.Rng(int(seed) & 0xFFFFFFFF)
        base = twin.default_state(scene)
        self.tools = {}
        for ins in scene["instruments"]:
            q = next(x for x in base["instruments"] if x["id"] == ins["id"])
            p = port_of(scene, ins["port"])
            self.tools[ins["id"]] = dict(ins=ins, port=p, spec=tool_spec(ins),
                                         track=Track(dict(hinge=hinge_of(p, q), roll=q["roll"], jaw=q["jaw"])))
        self.scope_port = port_of(scene, (scene.get("scope") or {}).get("port", "C"))
        self.scope0 = dict(base["scope"]) if base.get("scope") else scope_default(scene)
        self.scope = Track(dict(self.scope0))
        self.centre = workspace_centre(scene)
        self.bench_z = float((scene.get("bench") or {}).get("z", 0.0))
        self.labels = []
        self._plan(reaches, grasps, crossings, still, reaims)

    # ------------------------------------------------------------ targets
    def reachable(self, tool, hinge, jaw):
        """Inside TARGET_CONE of the port's cone, within the insertion limits (3 mm past the floor, 95 % of the
        maximum) and 8 mm above the bench."""
        q = aim_at(tool["port"], hinge, 0.0, jaw)
        sp = tool["spec"]
        return (cone_angle(q) <= tool["port"]["cone"] * TARGET_CONE and
                sp.get("min_insertion", 0.0) + 0.003 <= q["insertion"] <= sp["max_insertion"] * 0.95 and
                hinge[2] >= self.bench_z + 0.008)

    def target(self, tool, jaw=0.9, radius=0.05, side=None):
        """A random reachable hinge target near the workspace centre (within 1.8 radius): a direction drawn inside
        the cone (uniform over its disc of angles), a depth along it; side '+' / '-' keeps it on that side (world
        y) of the working ports' midline. None found in 400 draws: the tool stays."""
        ys = [p["pivot"][1] for p in self.scene["ports"] if p.get("role") != "camera"]
        y_mid = float(np.mean(ys)) if ys else 0.0
        p, sp = tool["port"], tool["spec"]
        lo = max(0.05, sp.get("min_insertion", 0.0) + 0.003)
        hi = max(lo + 0.01, min(lo + 0.16, sp["max_insertion"] * 0.9))
        for _ in range(400):
            r = math.sqrt(self.rng.u()) * p["cone"] * TARGET_CONE
            az = self.rng.uniform(0.0, 2.0 * math.pi)
            d = p["R"] @ np.array([math.sin(r) * math.cos(az), math.sin(r) * math.sin(az), math.cos(r)])
            h = p["pivot"] + d * self.rng.uniform(lo, hi)
            if np.linalg.norm(h - self.centre) > 1.8 * radius:
                continue
            if (side == "+" and h[1] < y_mid) or (side == "-" and h[1] > y_mid):
                continue
            if self.reachable(tool, h, jaw):
                return h
        return tool["track"].v["hinge"].copy()

    # ------------------------------------------------------------ moves
    def reach(self, tid, t0, hinge, roll=None, jaw=None, v=V_MEAN, kind="reach"):
        """A min-jerk move of the hinge at a mean speed drawn in v (m/s); at least 0.35 s."""
        tr = self.tools[tid]["track"]
        cur = tr.v
        goal = dict(hinge=np.asarray(hinge, float), roll=cur["roll"] if roll is None else roll,
                    jaw=cur["jaw"] if jaw is None else jaw)
        dist = float(np.linalg.norm(goal["hinge"] - cur["hinge"]))
        tr.hold_until(t0)
        return tr.move(max(t0, tr.t), max(0.35, dist / self.rng.uniform(*v)), goal, kind)

    def set_jaw(self, tid, t0, jaw, kind="jaw"):
        tr = self.tools[tid]["track"]
        tr.hold_until(t0)
        cur = tr.v
        return tr.move(max(t0, tr.t), max(0.25, abs(jaw - cur["jaw"]) / 1.5), dict(cur, jaw=jaw), kind)

    def sync(self, t=None):
        """Every tool held until t (default: the latest end); returns that time."""
        t = max([tl["track"].t for tl in self.tools.values()] + ([t] if t is not None else []))
        for tl in self.tools.values():
            tl["track"].hold_until(t)
        return t

    def _plan(self, reaches, grasps, crossings, still, reaims):
        ids = list(self.tools)
        t = 0.5
        # 1. reaches: every tool to a new target, the second hand starting a little later
        t0 = t
        for _ in range(reaches):
            for j, tid in enumerate(ids):
                tool = self.tools[tid]
                lag = self.rng.uniform(0.1, 0.6) if j else 0.0
                self.reach(tid, max(tool["track"].t, t + lag), self.target(tool),
                           roll=self.rng.uniform(-1.2, 1.2), jaw=self.rng.uniform(0.2, 0.8))
            t = self.sync()
        self.labels.append((t0, t, "reach"))
        # 2. grasp cycles at a nearby target: open wide, dwell, close, hold closed; then part-open
        t0 = t
        for tid in ids:
            tool = self.tools[tid]
            tt = self.reach(tid, t, self.target(tool, jaw=tool["spec"]["jaw_max"], radius=0.035), jaw=0.3)
            for _ in range(grasps):
                tt = self.set_jaw(tid, tt, tool["spec"]["jaw_max"] * self.rng.uniform(0.75, 1.0), "grasp_open")
                tt = tool["track"].hold(tt, self.rng.uniform(0.2, 0.5))
                tt = self.set_jaw(tid, tt, 0.0, "grasp_close")
                tt = tool["track"].hold(tt, self.rng.uniform(0.3, 0.8), "grasp_hold")
            self.set_jaw(tid, tt, 0.4)
        t = self.sync()
        self.labels.append((t0, t, "grasp"))
        # 3. crossings: each tip to the side away from its own port, 8 mm above / below the other, then back
        t0 = t
        if len(ids) >= 2:
            a, b = ids[0], ids[1]
            y_mid = 0.5 * (self.tools[a]["port"]["pivot"][1] + self.tools[b]["port"]["pivot"][1])
            away = {tid: ("-" if self.tools[tid]["port"]["pivot"][1] > y_mid else "+") for tid in (a, b)}
            for _ in range(crossings):
                ha = self.target(self.tools[a], jaw=0.6, radius=0.04, side=away[a]) + np.array([0.0, 0.0, 0.008])
                hb = self.target(self.tools[b], jaw=0.6, radius=0.04, side=away[b]) - np.array([0.0, 0.0, 0.008])
                ends = [t]
                if self.reachable(self.tools[a], ha, 0.6):
                    ends.append(self.reach(a, t, ha, v=(0.015, 0.035), kind="cross"))
                if self.reachable(self.tools[b], hb, 0.6):
                    ends.append(self.reach(b, t + 0.2, hb, v=(0.015, 0.035), kind="cross"))
                t = self.sync(max(ends) + 0.6)
                back = {tid: "+" if away[tid] == "-" else "-" for tid in (a, b)}
                ea = self.reach(a, t, self.target(self.tools[a], radius=0.04, side=back[a]), kind="uncross")
                eb = self.reach(b, t + 0.15, self.target(self.tools[b], radius=0.04, side=back[b]), kind="uncross")
                t = self.sync(max(ea, eb))
        self.labels.append((t0, t, "cross"))
        # 4. the still hold of everything (the stationary-instrument statistic)
        t0 = t
        for tl in self.tools.values():
            tl["track"].hold(tl["track"].t, t + still - tl["track"].t, "still")
        self.scope.hold_until(t + still, "still")
        t += still
        self.labels.append((t0, t, "still_hold"))
        # 5. scope re-aims inside 85 % of its cone at ~20 deg/s, the instruments working meanwhile
        t0 = t
        s0, cone = self.scope0, self.scope_port["cone"] * TARGET_CONE
        for _ in range(reaims):
            q, _ = cone_clamp(dict(yaw=s0["yaw"] + self.rng.uniform(-0.25, 0.25),
                                   pitch=s0["pitch"] + self.rng.uniform(-0.25, 0.25),
                                   roll=s0["roll"] + self.rng.uniform(-0.6, 0.6),
                                   insertion=s0["insertion"] + self.rng.uniform(-0.004, 0.012)), cone)
            cur = self.scope.v
            ang = math.hypot(q["yaw"] - cur["yaw"], q["pitch"] - cur["pitch"], 0.5 * (q["roll"] - cur["roll"]))
            self.scope.hold_until(t)
            t = self.scope.move(t, max(0.8, ang / (20 * DEG)), q, "scope_reaim")
            t = self.scope.hold(t, 1.0)
            for tid in ids:
                tool = self.tools[tid]
                tool["track"].hold_until(t - 1.5)
                self.reach(tid, tool["track"].t, self.target(tool, radius=0.04))
        self.scope.hold_until(t)
        t = self.scope.move(t, 1.2, dict(s0), "scope_reaim")
        t = self.sync(t)
        self.labels.append((t0, t, "scope_reaim"))
        self.duration = t + 0.5
        self.sync(self.duration)
        self.scope.hold_until(self.duration)

    # ------------------------------------------------------------ evaluation
    def state_at(self, t):
        """The joint state at time t (twin.default_state's layout), a pure function of t."""
        ins = []
        for tid, tool in self.tools.items():
            v = tool["track"].at(t, _blend_tool)
            q = aim_at(tool["port"], v["hinge"], v["roll"], v["jaw"])
            ins.append(dict(id=tid, **q))
        scope = self.scope.at(t, _blend_joints)
        return dict(t=float(t), instruments=ins, scope=dict(scope),
                    boards=[dict(name=b["name"], T_D_B=None, visible=False) for b in self.scene.get("boards") or []])

    def phase_at(self, t):
        for t0, t1, name in self.labels:
            if t0 <= t < t1:
                return name
        return "idle"


SCRIPTS = ("task_set",)


def make_script(name, scene, seed=0, **opts):
    """makeScript(name, scene, seed, opts): a planned script with duration, state_at(t), phase_at(t), labels."""
    if name == "task_set":
        return TaskSet(scene, seed, **opts)
    raise ValueError("unknown script %s; known: %s" % (name, ", ".join(SCRIPTS)))


def tip_speeds(script, scene, dt=0.01):
    """Hinge speed (mm/s) of every instrument between samples dt apart over the script (tipSpeeds)."""
    ports = {ins["id"]: port_of(scene, ins["port"]) for ins in scene["instruments"]}
    out = {tid: [] for tid in ports}
    prev = None
    for t in np.arange(0.0, script.duration + 1e-9, dt):
        st = script.state_at(float(t))
        cur = {q["id"]: hinge_of(ports[q["id"]], q) for q in st["instruments"]}
        if prev is not None:
            for tid in cur:
                out[tid].append(float(np.linalg.norm(cur[tid] - prev[tid])) / dt * 1e3)
        prev = cur
    return {tid: np.array(v) for tid, v in out.items()}
