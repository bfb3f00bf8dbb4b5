"""The real twin scene through the camera path's boundary: scene.json + meshes/*.bin + physics.json -> the upload
arrays of vt_upload_scene, the body poses of a joint state, and the three tracking cameras.

Follows the reference's scene preparation (gpu/scene_gpu.js:20-472: decodeMesh, weld + subdivide of the static
meshes by vt_weld_subdivide, benchMesh, materialsFromPhysics, emitterFromPhysics, bodyOfPart, the per-body triangle
sets, fitCapsule, the emitter records) and its kinematics (control/kinematics.js:97-244: shaftRotation, forward,
scopeForward, defaultState, posesFromState, cannulaT). Differences, by the library's design: the static triangles
keep their preparation order and the library builds every BVH at upload (the same binned-SAH trees); the charuco
board's print texture rides with its mesh (texture coordinates per vertex, the mask's red channel); FDM layer lines
are not carried. The emitter (the scope's light outlet) rides on the scope body as the
reference's does (its records and axis in the scope's frame, posed per sample); a scene without a scope body places
it in the world at the scope's pose as static geometry.
"""
import json
import math
import os

import numpy as np

from . import lib, scenes

MESH_FLAG = dict(STATIC=1, EMITTER=2, LAYER=4, BOARD=8, INSTRUMENT=16, SCOPE=32, BENCH=64, TEXTURED=128)
STATIC_BODY = scenes.STATIC_BODY
DEG = math.pi / 180


def unwrap(x):
    return x["value"] if isinstance(x, dict) and "value" in x else x


# ---------------------------------------------------------------------------- loading (gpu/scene_gpu.js:22-63)

def decode_mesh(meta, buf):
    """One mesh's .bin by its layout block: P, N float32 (n, 3), I uint32 (t, 3), UV float32 (n, 2) or None."""
    lay = meta["layout"]
    nv, nt = meta["vertex_count"], meta["triangle_count"]
    vec = lambda c: nv if (c == nv * 3 and nv * 3 != nv) else c
    idxc = lambda c: nt * 3 if c == nt else c
    po, pc = lay["positions"]
    P = np.frombuffer(buf, np.float32, vec(pc) * 3, po).reshape(-1, 3)
    I = np.frombuffer(buf, np.uint32, idxc(lay["indices"][1]), lay["indices"][0]).reshape(-1, 3)
    if lay.get("normals"):
        no, nc = lay["normals"]
        N = np.frombuffer(buf, np.float32, vec(nc) * 3, no).reshape(-1, 3)
    else:
        N = compute_normals(P, I)
    UV = None
    if lay.get("uvs"):
        uo, uc = lay["uvs"]
        UV = np.frombuffer(buf, np.float32, (nv if (uc == nv * 2 and nv * 2 != nv) else uc) * 2, uo).reshape(-1, 2)
    return P, N, I, UV


def compute_normals(P, I):
    """Area-weighted vertex normals for a mesh shipped without them (gpu/scene_gpu.js:38-52)."""
    P = P.astype(np.float64)
    n = np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]])
    N = np.zeros_like(P)
    for k in range(3):
        np.add.at(N, I[:, k], n)
    L = np.linalg.norm(N, axis=1, keepdims=True)
    return (N / np.where(L > 0, L, 1)).astype(np.float32)


def load(root):
    """scene.json, physics.json and every mesh of the scene under `root`; each board's printed texture decoded into
    board["mask"] (the PNG's red channel, uint8, row 0 the image's top; None when the file is absent), as
    _loadBoardTexture reads it (gpu/renderer.js:395-412)."""
    with open(os.path.join(root, "scene.json")) as f:
        scene = json.load(f)
    with open(os.path.join(root, "physics.json")) as f:
        physics = json.load(f)
    meshes = {}
    for m in scene["meshes"]:
        with open(os.path.join(root, m["file"]), "rb") as f:
            meshes[m["name"]] = decode_mesh(m, f.read())
    for b in scene.get("boards") or []:
        path = os.path.join(root, b["texture"]) if b.get("texture") else None
        b["mask"] = None
        if path and os.path.exists(path):
            from PIL import Image
            with Image.open(path) as im:
                b["mask"] = np.ascontiguousarray(np.asarray(im.convert("RGB"))[:, :, 0])
    return scene, meshes, physics


# ---------------------------------------------------------------------------- bench (gpu/scene_gpu.js:146-167)

def bench_mesh(z, extent, fine, h):
    """A grid of cells <= h over `fine` (clipped to the extent) and four coarse quads out to the extent."""
    P, I = [], []
    (ex0, ey0), (ex1, ey1) = extent
    fx0, fy0 = max(ex0, fine[0][0]), max(ey0, fine[0][1])
    fx1, fy1 = min(ex1, fine[1][0]), min(ey1, fine[1][1])
    nx = max(1, math.ceil((fx1 - fx0) / (h * 0.999)))
    ny = max(1, math.ceil((fy1 - fy0) / (h * 0.999)))

    def v(x, y):
        P.append((x, y, z))
        return len(P) - 1

    grid = [[v(fx0 + (fx1 - fx0) * i / nx, fy0 + (fy1 - fy0) * j / ny) for i in range(nx + 1)] for j in range(ny + 1)]
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = grid[j][i], grid[j][i + 1], grid[j + 1][i], grid[j + 1][i + 1]
            I.extend((a, b, d, a, d, c) if (i + j) & 1 else (a, b, c, b, d, c))

    def quad(x0, y0, x1, y1):
        if not (x1 > x0 and y1 > y0):
            return
        a, b, c, d = v(x0, y0), v(x1, y0), v(x0, y1), v(x1, y1)
        I.extend((a, b, d, a, d, c))

    quad(ex0, ey0, ex1, fy0)
    quad(ex0, fy1, ex1, ey1)
    quad(ex0, fy0, fx0, fy1)
    quad(fx1, fy0, ex1, fy1)
    P = np.array(P, np.float64)
    N = np.tile(np.array([0.0, 0.0, 1.0]), (len(P), 1))
    return P, N, np.array(I, np.uint32).reshape(-1, 3)


# ---------------------------------------------------------------------------- materials (gpu/scene_gpu.js:169-215)

def prepare_material(m):
    """gpu/bsdf.js:68-88 prepareMaterial on a physics material (values unwrapped) -> the reference's fields."""
    model = unwrap(m.get("model"))
    if model == "coated_conductor":
        r = max(float(unwrap(m.get("roughness")) if m.get("roughness") is not None else 0.3), 0.02)
        rc = max(float(unwrap(m.get("coat_roughness")) if m.get("coat_roughness") is not None else 0.05), 0.02)
        tau = unwrap(m.get("coat_transmittance"))
        tau = [1.0, 1.0, 1.0] if tau is None else [float(x) for x in np.atleast_1d(tau)]
        if len(tau) == 1:
            tau = tau * 3
        return dict(kind=scenes.KIND_COATED, c=[float(x) for x in unwrap(m.get("f0")) or (0.56, 0.57, 0.58)],
                    roughness=r, ior=float(unwrap(m.get("coat_ior")) or 1.5), tau=tau, coat_roughness=rc, fbar=0.0)
    metallic = unwrap(m.get("metallic"))
    metallic = (1 if model == "conductor" else 0) if metallic is None else int(float(metallic))
    textured = m.get("rho_d_paper") is not None and m.get("rho_d_ink") is not None
    rough = unwrap(m.get("roughness_paper" if textured else "roughness"))
    r = max(float(0.5 if rough is None else rough), 0.02)
    ior = unwrap(m.get("ior"))
    ior = 1.5 if ior is None else float(ior)
    if metallic:
        f0 = unwrap(m.get("f0"))
        return dict(kind=scenes.KIND_CONDUCTOR, c=[float(x) for x in (f0 if f0 is not None else (0.56, 0.57, 0.58))],
                    roughness=r, ior=ior, fbar=0.0)
    rho = unwrap(m.get("rho_d_paper" if textured else "rho_d"))
    fa = unwrap(m.get("fresnel_avg"))
    out = dict(kind=scenes.KIND_DIELECTRIC, c=[float(x) for x in (rho if rho is not None else (0.5, 0.5, 0.5))],
               roughness=r, ior=ior, fbar=None if fa is None else float(fa))
    if textured:
        # charuco_print: the ink's rho_d and roughness (gpu/scene_gpu.js:192; roughness_ink 0.4 by default)
        ri = unwrap(m.get("roughness_ink"))
        out["texture"] = dict(ink=[float(x) for x in unwrap(m.get("rho_d_ink"))],
                              rough_ink=0.4 if ri is None else float(ri))
    return out


def material_record(pm):
    """16-float record of a prepared material; a dielectric's Fbar is physics' fresnel_avg when given (the record
    keeps the computed one beside it, gpu/bsdf.js:84-86)."""
    rec = scenes.material_record(pm["kind"], pm["c"], roughness=pm["roughness"], ior=pm["ior"],
                                 tau=pm.get("tau", (1.0, 1.0, 1.0)), coat_roughness=pm.get("coat_roughness", 0.05))
    if pm["kind"] == scenes.KIND_DIELECTRIC and pm.get("fbar") is not None:
        rec[6] = pm["fbar"]
    if pm.get("texture"):
        # the ink in words 8-11 (gpu/renderer.js:293: ink rgb, roughInk^2)
        rec[8:11] = pm["texture"]["ink"]
        rec[11] = pm["texture"]["rough_ink"] ** 2
    return rec


def materials_from_physics(physics, names):
    """(records, index by name, prepared fields) for the materials named, in first-use order; a name physics does
    not hold gets the neutral grey dielectric (gpu/scene_gpu.js:170-203)."""
    src = (physics or {}).get("materials") or {}
    recs, index, prepared = [], {}, []
    for name in names:
        if name in index:
            continue
        pm = prepare_material(src.get(name) or {})
        pm["name"], pm["missing"] = name, name not in src
        pm["emitter"] = unwrap((src.get(name) or {}).get("model")) == "emitter"
        index[name] = len(recs)
        recs.append(material_record(pm))
        prepared.append(pm)
    return recs, index, prepared


def emitter_from_physics(physics):
    """L0 (rgb), the radiance profile (theta rad, p) and the emission axis angle from the shaft (gpu/scene_gpu.js:
    205-214)."""
    em = ((physics or {}).get("materials") or {}).get("emitter") or {}
    L0 = unwrap(em.get("L0"))
    L0 = [1.0, 1.0, 1.0] if L0 is None else [float(x) for x in L0]
    prof = unwrap(em.get("profile"))
    if isinstance(prof, str):
        node = physics
        for k in prof.split("."):
            node = node.get(k) if isinstance(node, dict) else None
        prof = node
    prof = prof or ((physics or {}).get("light") or {}).get("profile") or {}
    th = unwrap(prof.get("theta_deg"))
    p = unwrap(prof.get("p"))
    th = [0.0, 90.0] if th is None else th
    p = [1.0, 1.0] if p is None else p
    ax = ((physics or {}).get("scope") or {}).get("emission_axis") or {}
    ax = unwrap(ax.get("angle_from_shaft_deg"))
    return dict(L0=L0, theta=[float(t) * DEG for t in th], p=[float(x) for x in p],
                axis_deg=float("nan") if ax is None else float(ax))


# ---------------------------------------------------------------------------- bodies (gpu/scene_gpu.js:218-235)

def body_of_part(scene):
    """mesh name -> {key, kind, port, owner, emitter}: instrument parts, the scope (and its light outlet), boards,
    each camera's own lens front (a fixed body), any other dynamic mesh with a body key."""
    out = {}
    for ins in scene.get("instruments") or []:
        for part in ins.get("parts") or []:
            out[part["mesh"]] = dict(key="%s/%s" % (ins["id"], part["body"]), kind="instrument", port=ins.get("port"))
    sc = scene.get("scope") or {}
    for m in sc.get("mesh_parts") or []:
        out[m] = dict(key="scope", kind="scope", port=sc.get("port"))
    if sc.get("emitter_mesh"):
        out[sc["emitter_mesh"]] = dict(key="scope", kind="scope", port=sc.get("port"), emitter=True)
    for b in scene.get("boards") or []:
        out[b["mesh"]] = dict(key="board/" + b["name"], kind="board")
    for m in scene.get("meshes") or []:
        if m.get("owner_camera") and m["name"] not in out:
            out[m["name"]] = dict(key="fixed/" + m["name"], kind="fixed", owner=m["owner_camera"])
    ports = {p["id"] for p in scene.get("ports") or []}
    for m in scene.get("meshes") or []:
        if m["name"] in out or not m.get("dynamic") or not m.get("body"):
            continue
        head = m["body"].split("/")[0]
        kind = "scope" if m["body"] == "scope" else "board" if head == "board" else "fixture"
        out[m["name"]] = dict(key=m["body"], kind=kind, port=m.get("port", head if head in ports else None))
    return out


def fit_capsule(P):
    """Segment a-b and radius r around a vertex set: axis the principal direction by 64 power iterations from
    (0.3, 0.5, 0.8), r the largest distance from it (gpu/scene_gpu.js:509-532)."""
    P = np.asarray(P, np.float64).reshape(-1, 3)
    if not len(P):
        return dict(a=[0.0, 0.0, 0.0], b=[0.0, 0.0, 0.0], r=0.0)
    c = P.mean(axis=0)
    d = P - c
    C = d.T @ d
    ax = np.array([0.3, 0.5, 0.8])
    for _ in range(64):
        y = C @ ax
        L = np.linalg.norm(y)
        if not L > 0:
            break
        ax = y / L
    s = d @ ax
    r = float(np.max(np.linalg.norm(d - s[:, None] * ax, axis=1)))
    return dict(a=list(c + s.min() * ax), b=list(c + s.max() * ax), r=r)


# ---------------------------------------------------------------------------- kinematics (control/kinematics.js)

def rot_x(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def rot_y(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def rot_z(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def port(scene, pid):
    for p in scene["ports"]:
        if p["id"] == pid:
            return dict(pivot=np.array(p["pivot"], float), R=np.array(p["R_D_P"], float),
                        cannula_body=p.get("cannula_body"))
    raise KeyError("scene has no port %s" % pid)


def shaft_rotation(R_D_P, q):
    """R_D_S = R_D_P Ry(yaw) Rx(pitch) Rz(roll) (kinematics.js:97-99)"""
    return R_D_P @ rot_y(q["yaw"]) @ rot_x(q["pitch"]) @ rot_z(q["roll"])


def make_T(R, t):
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = t
    return T


def default_state(scene):
    """scene.default_state (kinematics.js:171-178); the scope falls back to scene.scope.default."""
    d = scene["default_state"]
    scope = dict(d["scope"]) if d.get("scope") else None
    if scope is None:
        df = (scene.get("scope") or {}).get("default") or {}
        scope = dict(yaw=df.get("yaw", 0.0), pitch=df.get("pitch", 0.0), roll=df.get("roll", 0.0),
                     insertion=df.get("insertion", 0.016))
    return dict(t=d.get("t", 0.0), instruments=[dict(q) for q in d["instruments"]], scope=scope,
                boards=[dict(name=b["name"], T_D_B=None, visible=False) for b in scene.get("boards") or []])


def poses_from_state(scene, state):
    """{body key: 4x4 row-major T_D_body or None} (kinematics.js:212-233): instrument shaft and jaws about the
    hinge, the scope at its lens point, each port's cannula (the tool's yaw and pitch, no roll, at the pivot),
    boards from the state (hidden without a pose)."""
    bodies = {}
    for ins in scene["instruments"]:
        q = next((s for s in state["instruments"] if s["id"] == ins["id"]), None)
        if q is None:
            continue
        p = port(scene, ins["port"])
        R = shaft_rotation(p["R"], q)
        hinge = p["pivot"] + R[:, 2] * q["insertion"]
        h = np.array(ins.get("hinge_in_S") or [0.0, 0.0, 0.0], float)
        bodies[ins["id"] + "/shaft"] = make_T(R, hinge - R @ h)
        bodies[ins["id"] + "/jaw_a"] = make_T(R @ rot_x(-q["jaw"] / 2), hinge)
        bodies[ins["id"] + "/jaw_b"] = make_T(R @ rot_x(q["jaw"] / 2), hinge)
        if p["cannula_body"]:
            bodies[p["cannula_body"]] = make_T(p["R"] @ rot_y(q["yaw"]) @ rot_x(q["pitch"]), p["pivot"])
    if scene.get("scope") and state.get("scope"):
        p = port(scene, scene["scope"].get("port", "C"))
        R = shaft_rotation(p["R"], state["scope"])
        bodies["scope"] = make_T(R, p["pivot"] + R[:, 2] * state["scope"]["insertion"])
        if p["cannula_body"]:
            q = state["scope"]
            bodies[p["cannula_body"]] = make_T(p["R"] @ rot_y(q["yaw"]) @ rot_x(q["pitch"]), p["pivot"])
    for b in scene.get("boards") or []:
        sb = next((x for x in state.get("boards") or [] if x["name"] == b["name"]), None)
        ok = sb is not None and sb.get("visible") is not False and sb.get("T_D_B") is not None
        bodies["board/" + b["name"]] = np.array(sb["T_D_B"], float) if ok else None
    return bodies


# ---------------------------------------------------------------------------- the whole scene

def prepare(scene, meshes, physics, node_spacing=0.003, poses=None):
    """Upload arrays of the twin scene (pack_scene's keys, with tri_body / body_count) plus: bodies (per body key,
    kind, owner, port, capsule and occlusion capsule in the body frame, the area per material for rho_eff),
    shadow_skip, emit_body, profiles (the outlet's radiance profile about its axis, in the scope's frame), stats.
    poses: {key: 4x4} placing the light outlet when there is no scope body (default: the scene's default state)."""
    if poses is None:
        poses = poses_from_state(scene, default_state(scene))
    bodies_of = body_of_part(scene)
    metas = list(scene["meshes"])
    bench_mat = (scene.get("bench") or {}).get("material", "bench_white")
    recs, mindex, prepared = materials_from_physics(physics, [m["material"] for m in metas] + [bench_mat])
    em = emitter_from_physics(physics)
    mesh_table = []

    def add_mesh(name, material, flags):
        mesh_table.append(dict(name=name, mat=mindex[material], flags=flags))
        return len(mesh_table) - 1

    # static: weld + subdivide each static mesh, then the bench
    parts = []        # (P float64, N, I, mesh id, body)
    part_uv = {}      # part index -> texture coordinates (textured meshes)
    lo, hi = np.array([np.inf, np.inf]), np.array([-np.inf, -np.inf])
    raw_static = 0
    for m in metas:
        if m.get("dynamic") or m["name"] in bodies_of:
            continue
        P, N, I, _ = meshes[m["name"]]
        raw_static += I.shape[0]
        flags = MESH_FLAG["STATIC"]
        ll = ((physics.get("materials") or {}).get(m["material"]) or {}).get("layer_lines")
        if m.get("build_dir") and ll and unwrap(ll.get("height_m")) and float(unwrap(ll.get("depth_m")) or 0) > 0:
            flags |= MESH_FLAG["LAYER"]
        mid = add_mesh(m["name"], m["material"], flags)
        P2, N2, I2 = lib.weld_subdivide(P, N, I, node_spacing)
        parts.append((P2, N2, I2, mid, None))
        lo = np.minimum(lo, P2[:, :2].min(axis=0))
        hi = np.maximum(hi, P2[:, :2].max(axis=0))
    if scene.get("bench"):
        b = scene["bench"]
        ext = b.get("extent") or [[-0.5, -0.5], [0.5, 0.5]]
        fine = [list(lo), list(hi)] if np.isfinite(lo[0]) else ext
        fb = [[fine[0][0] - 0.006, fine[0][1] - 0.006], [fine[1][0] + 0.006, fine[1][1] + 0.006]]
        P2, N2, I2 = bench_mesh(b.get("z", 0.0), ext, fb, node_spacing)
        mid = add_mesh("bench", bench_mat, MESH_FLAG["STATIC"] | MESH_FLAG["BENCH"])
        parts.append((P2, N2, I2, mid, None))

    # bodies (bodyOfPart): the light outlet belongs to the scope's body, as in the reference's per-body sets
    emit_parts = []
    body_list, body_index = [], {}
    for m in metas:
        b = bodies_of.get(m["name"])
        if not b:
            continue
        if b.get("emitter"):
            emit_parts.append(m)
            continue
        if b["key"] not in body_index:
            body_index[b["key"]] = len(body_list)
            body_list.append(dict(b, meshes=[]))
        body_list[body_index[b["key"]]]["meshes"].append(m)
    if len(body_list) > 24:
        raise ValueError("too many bodies (%d > 24)" % len(body_list))
    light_idx = 0
    emit_body = body_index.get("scope") if emit_parts else None
    if emit_parts and "emitter" not in mindex:
        mindex["emitter"] = len(recs)
        recs.append(scenes.material_record(scenes.KIND_DIELECTRIC, (0.02, 0.02, 0.02), roughness=0.05))
        prepared.append(dict(name="emitter", emitter=True))
    # the outlet in the emitter's frame (the scope's, or the world at the scope's pose without a scope body)
    T = np.eye(4)
    if emit_body is None and poses.get("scope") is not None:
        T = np.asarray(poses["scope"], np.float64)
    emit_local = None
    emit_rows = []
    for m in emit_parts:
        P, N, I, _ = meshes[m["name"]]
        Pe = P.astype(np.float64) @ T[:3, :3].T + T[:3, 3]
        Ne = N.astype(np.float64) @ T[:3, :3].T
        mid = add_mesh(m["name"], "emitter", MESH_FLAG["EMITTER"] | MESH_FLAG["SCOPE"])
        emit_rows.append((Pe, Ne, np.asarray(I, np.uint32), mid, emit_body))
        emit_local = (Pe, Ne, np.asarray(I, np.uint32))
    if emit_body is None:
        parts.extend(emit_rows)

    # bodies: their meshes concatenated in the body frame, in scene order (the scope's outlet after its parts)
    for bi, body in enumerate(body_list):
        for m in body["meshes"]:
            P, N, I, _ = meshes[m["name"]]
            # a mesh on a printed material with texture coordinates samples the print mask
            # (gpu/scene_gpu.js:324-325: not emitter, textured material, UVs present)
            uv = meshes[m["name"]][3]
            texd = uv is not None and bool(prepared[mindex[m["material"]]].get("texture"))
            flags = ((MESH_FLAG["BOARD"] if body["kind"] == "board" else 0) |
                     (MESH_FLAG["INSTRUMENT"] if body["kind"] == "instrument" else 0) |
                     (MESH_FLAG["SCOPE"] if body["kind"] == "scope" else 0) |
                     (MESH_FLAG["TEXTURED"] if texd else 0))
            mid = add_mesh(m["name"], m["material"], flags)
            parts.append((P.astype(np.float64), N.astype(np.float64), np.asarray(I, np.uint32), mid, bi))
            part_uv[len(parts) - 1] = uv
        if bi == emit_body:
            parts.extend(emit_rows)

    # capsules and the area per material of each body (its rho_eff needs the materials' albedo tables)
    overview = {m["name"] for m in metas if m.get("overview_only")}
    for bi, body in enumerate(body_list):
        Ps = [meshes[m["name"]][0] for m in body["meshes"]]
        body["capsule"] = fit_capsule(np.concatenate(Ps))
        keep = [meshes[m["name"]] for m in body["meshes"] if m["name"] not in overview]
        if len(keep) < len(body["meshes"]):
            if keep:
                body["occ_capsule"] = fit_capsule(np.concatenate([P[np.unique(I.reshape(-1))] for P, N, I, _ in keep]))
            else:
                body["occ_capsule"] = dict(a=[0.0] * 3, b=[0.0] * 3, r=0.0)
        else:
            body["occ_capsule"] = body["capsule"]
        area = {}
        for m in body["meshes"]:
            if m["name"] in overview:
                continue
            P, N, I, _ = meshes[m["name"]]
            P = P.astype(np.float64)
            a = 0.5 * np.linalg.norm(np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]]), axis=1).sum()
            k = mindex[m["material"]]
            area[k] = area.get(k, 0.0) + float(a)
        body["area_by_material"] = area

    # one vertex / triangle list: static parts first, bodies after (the outlet with the scope)
    Pl, Nl, Il, Ml, Bl, UVl = [], [], [], [], [], []
    base = 0
    for k, (P, N, I, mid, bi) in enumerate(parts):
        uv = part_uv.get(k)
        UVl.append(np.zeros((len(P), 2), np.float32) if uv is None else np.asarray(uv, np.float32).reshape(-1, 2))
        Pl.append(np.asarray(P, np.float64))
        Nl.append(np.asarray(N, np.float64))
        Il.append(np.asarray(I, np.uint32) + base)
        Ml.append(np.full(I.shape[0], mid, np.uint32))
        Bl.append(np.full(I.shape[0], STATIC_BODY if bi is None else bi, np.uint32))
        base += P.shape[0]
    P = np.concatenate(Pl).astype(np.float32)
    N = np.concatenate(Nl).astype(np.float32)
    I = np.concatenate(Il)
    tri_mesh = np.concatenate(Ml)
    tri_body = np.concatenate(Bl)
    meshes_arr = np.zeros((len(mesh_table), 4), np.uint32)
    for k, mt in enumerate(mesh_table):
        meshes_arr[k] = (mt["mat"], mt["flags"], light_idx, 0)
    out = scenes.pack_scene(P, N, I, tri_mesh, meshes_arr, np.array(recs, np.float32),
                            np.array([[em["L0"][0], em["L0"][1], em["L0"][2], 0.0]], np.float32),
                            tri_body=tri_body, body_count=len(body_list))
    if any(mt["flags"] & MESH_FLAG["TEXTURED"] for mt in mesh_table):
        # texture coordinates per vertex and the first board's print mask (one paper texel without it)
        out["uvs"] = np.concatenate(UVl).reshape(-1)
        boards = scene.get("boards") or []
        out["board_mask"] = boards[0].get("mask") if boards else None
    # the outlet's axis: the scope's +z turned toward +y by the emission angle (physics scope.emission_axis), else
    # the outlet's mean normal; in the emitter's frame
    profiles = {}
    shadow_skip = 0
    if emit_local is not None:
        Pw, Nw, Ie = emit_local
        a, b, c = Pw[Ie[:, 0]], Pw[Ie[:, 1]], Pw[Ie[:, 2]]
        n = np.cross(b - a, c - a)
        area = 0.5 * np.linalg.norm(n, axis=1)
        vn = Nw[Ie[:, 0]] + Nw[Ie[:, 1]] + Nw[Ie[:, 2]]
        n[np.sum(n * vn, axis=1) < 0] *= -1
        mean_n = (n / np.linalg.norm(n, axis=1, keepdims=True) * area[:, None]).sum(axis=0)
        mean_n /= np.linalg.norm(mean_n)
        if math.isfinite(em["axis_deg"]):
            ang = em["axis_deg"] * DEG
            axis = T[:3, :3] @ np.array([0.0, math.sin(ang), math.cos(ang)])
        else:
            axis = mean_n
        profiles[0] = (tuple(axis), em["theta"], em["p"])
        # shadow rays skip the scope's own BVH when no part of it lies in front of the outlet's plane
        if "scope" in body_index:
            cen = (a * area[:, None]).sum(axis=0) / area.sum()
            ahead = -np.inf
            for m in body_list[body_index["scope"]]["meshes"]:
                Pb = meshes[m["name"]][0].astype(np.float64) @ T[:3, :3].T + T[:3, 3]
                ahead = max(ahead, float(np.max((Pb - cen) @ mean_n)))
            if ahead <= 1e-6:
                shadow_skip = 1 << body_index["scope"]
    if emit_body is not None:
        out["emit_body"] = emit_body
    # the ports the bodies go through: what cavityCapsule cuts the occlusion capsules with
    out["ports"] = {p["id"]: dict(pivot=[float(x) for x in p["pivot"]], axis_out=[float(x) for x in p["axis_out"]],
                                  cannula_toe_insertion=p.get("cannula_toe_insertion"),
                                  cannula_outer_radius=p.get("cannula_outer_radius"))
                    for p in scene.get("ports") or [] if p.get("pivot") and p.get("axis_out")}
    out.update(bodies=body_list, shadow_skip=shadow_skip, profiles=profiles, mesh_table=mesh_table,
               materials_prepared=prepared, material_index=mindex,
               stats=dict(raw_static_tris=raw_static, static_tris=int(np.sum(tri_body == STATIC_BODY)),
                          static_verts=int(sum(p[0].shape[0] for p in parts if p[4] is None)),
                          static_nodes=int(sum(p[0].shape[0] for p in parts
                                               if mesh_table[p[3]]["flags"] & MESH_FLAG["STATIC"])),
                          bodies=len(body_list), emitter_area=float(out["emit_area"])))
    return out


def quat_from_mat(R):
    """Unit quaternion (x, y, z, w) of a rotation matrix (Shepperd's branch on the largest diagonal term)."""
    R = np.asarray(R, float)
    tr = R[0, 0] + R[1, 1] + R[2, 2]
    if tr > 0:
        s = math.sqrt(tr + 1.0) * 2
        q = ((R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s, 0.25 * s)
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = math.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        q = (0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s, (R[2, 1] - R[1, 2]) / s)
    elif R[1, 1] > R[2, 2]:
        s = math.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        q = ((R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s, (R[0, 2] - R[2, 0]) / s)
    else:
        s = math.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        q = ((R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s, (R[1, 0] - R[0, 1]) / s)
    q = np.array(q)
    return q / np.linalg.norm(q)


def rho_eff(prep, hh):
    """Each body's area-weighted luminance of its materials' hemispherical albedo (gpu/scene_gpu.js:455-466);
    hh: (materials, 3) from the albedo tables."""
    hh = np.asarray(hh, float).reshape(-1, 3)
    for body in prep["bodies"]:
        a = sum(body["area_by_material"].values())
        r = sum(A * (0.2126 * hh[k, 0] + 0.7152 * hh[k, 1] + 0.0722 * hh[k, 2])
                for k, A in body["area_by_material"].items())
        body["rho_eff"] = r / a if a else 0.0


def _finite(x):
    try:
        return math.isfinite(float(x))
    except (TypeError, ValueError):
        return False


def cavity_of(prep, body):
    """cavityCapsule's cut (gpu/scene_gpu.js:486-503) of a body through a port, as vt_instance's cavity fields: the
    port's pivot and inward axis (-axis_out); a cannula (fixture) keeps s >= 0 with the sheath's radius
    cannula_outer_radius, an instrument part or the scope s >= cannula_toe_insertion. None without a port (and for
    lens fronts and boards, which have no capsule at all)."""
    port = (prep.get("ports") or {}).get(body.get("port"))
    if port is None or body["kind"] in ("fixed", "board"):
        return None
    fixture = body["kind"] == "fixture"
    toe = port.get("cannula_toe_insertion")
    s0 = 0.0 if fixture or not _finite(toe) else float(toe)
    rr = port.get("cannula_outer_radius")
    r = float(rr) if fixture and _finite(rr) else 0.0
    return dict(origin=list(port["pivot"]), axis=[-x for x in port["axis_out"]], s0=s0, r=r)


def instances(prep, poses0, poses1=None):
    """vt_set_instances records (lib.Context.set_instances) of the prepared bodies at shutter open and close from
    posesFromState matrices; a body without a pose is hidden, a camera's lens front stays where the scene has it;
    occlusion capsules cut to the cavity (cavity_of)."""
    poses1 = poses0 if poses1 is None else poses1
    out = []
    for body in prep["bodies"]:
        T0, T1 = poses0.get(body["key"]), poses1.get(body["key"])
        if body["kind"] == "fixed":
            T0 = T1 = np.eye(4)
        if T0 is None:
            out.append(None)
            continue
        T0 = np.asarray(T0, float)
        T1 = T0 if T1 is None else np.asarray(T1, float)
        oc = body["occ_capsule"]
        out.append(dict(q0=quat_from_mat(T0[:3, :3]), p0=T0[:3, 3], q1=quat_from_mat(T1[:3, :3]), p1=T1[:3, 3],
                        occ_a=oc["a"], occ_b=oc["b"], occ_r=0.0 if body["kind"] in ("fixed", "board") else oc["r"],
                        rho_eff=body.get("rho_eff", 0.0), cavity=cavity_of(prep, body)))
    return out


def skip_mask(prep, camera_name):
    """Bodies a camera's own rays skip: its lens front (gpu/renderer.js:881-886)."""
    m = 0
    for i, body in enumerate(prep["bodies"]):
        if body.get("owner") == camera_name:
            m |= 1 << i
    return m


def cameras(scene, spp=16, depth=3, seed=0, light_samples=1, prep=None):
    """The scene's tracking cameras (KB4 lens, T_D_C, size) as vt_camera records (R_D_C columns, centre); with
    prep, each skips its own lens front."""
    out = []
    for k, c in enumerate(scene["cameras"]):
        K, D = c["lens"]["K"], c["lens"]["D"]
        fx = K[0][0]
        lens = [fx, K[1][1], K[0][2], K[1][2], D[0], D[1], D[2], D[3], K[0][1] / fx]
        T = np.array(c["T_D_C"], float)
        R = T[:3, :3].T.reshape(-1)          # columns of R_D_C, flattened
        w, h = c["size"]
        cam = scenes.camera(R.astype(np.float32), T[:3, 3], lens, w, h, spp, k, depth, light_samples=light_samples,
                            seed=seed)
        cam["name"] = c["name"]
        if prep is not None:
            cam["skip_bodies"] = skip_mask(prep, c["name"])
        out.append(cam)
    return out


def world_triangles(prep, poses):
    """Every triangle in the world at one pose set, hidden bodies left out: (P float32 (n, 3), I (t, 3), mesh per
    triangle). The static twin of a posed frame, for checks that trace without instances."""
    tris = np.asarray(prep["tris"], np.float32).reshape(-1, 16)
    tu = tris.view(np.uint32)
    idx = tu[:, [3, 7, 11]]
    P = np.asarray(prep["verts"], np.float32).reshape(-1, 8)[:, :3].astype(np.float64).copy()
    tb = np.asarray(prep["tri_body"], np.uint32)
    keep = np.ones(len(tris), bool)
    for bi, body in enumerate(prep["bodies"]):
        T = np.eye(4) if body["kind"] == "fixed" else poses.get(body["key"])
        sel = tb == bi
        if T is None:
            keep[sel] = False
            continue
        T = np.asarray(T, float)
        v = np.unique(idx[sel].reshape(-1))
        P[v] = P[v] @ T[:3, :3].T + T[:3, 3]
    return P.astype(np.float32), idx[keep], tu[keep, 12]



def lite_nodes(root):
    """The lite renderer's baked node table (assets/lite.json `files.nodes`, lite/renderer.js:214-221): per node
    8 floats (position, area, normal, mesh id as bits) and K uint16 gather hits (0xffff: the ray escaped), here as
    uint32 ids with 0xffffffff for a miss, the gather table layout of the node radiosity step."""
    with open(os.path.join(root, "lite", "lite.json")) as f:
        meta = json.load(f)
    info = meta["files"]["nodes"]
    raw = np.fromfile(os.path.join(root, "lite", info["file"]), np.uint8)
    lay = info["layout"]

    def view(name, dtype):
        off, count, _ = lay[name]
        return np.frombuffer(raw[off:off + count * np.dtype(dtype).itemsize].tobytes(), dtype)

    n, K = info["count"], info["K"]
    rec = view("nodes", np.float32).reshape(n, 8)
    hits = view("gather", np.uint16).reshape(n, K)
    g = hits.astype(np.uint32)
    g[hits == 0xFFFF] = 0xFFFFFFFF
    return dict(count=n, K=K, pos=rec[:, 0:3].copy(), area=rec[:, 3].copy(), nrm=rec[:, 4:7].copy(),
                mesh=rec[:, 7].copy().view(np.uint32), gtable=g.reshape(-1), hits=hits.copy(), stats=meta["stats"])


def baked_table(L, hh):
    """The baked table of lite_nodes as vt_baked_upload's arrays (nodeA / nodeN, lite/renderer.js:214-217), with
    hh (meshes x 3 or 4): the meshes' hemispherical albedos."""
    n = L["count"]
    a = np.zeros((n, 4), np.float32)
    a[:, 0:3], a[:, 3] = L["pos"], L["area"]
    nn = np.zeros((n, 4), np.float32)
    nn[:, 0:3] = L["nrm"]
    nn[:, 3] = L["mesh"].view(np.float32)
    h = np.zeros((np.asarray(hh).shape[0], 4), np.float32)
    h[:, 0:3] = np.asarray(hh, np.float32)[:, 0:3]
    return dict(nodes=n, K=L["K"], gather=L["hits"], node_a=a, node_n=nn, hh=h)
