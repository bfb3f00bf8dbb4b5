"""CPU oracle of the camera-rendering path — TEST INFRASTRUCTURE ONLY.

ctypes bindings of liboracle.so (built from the C sources beside this file by `make -C oracle`). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package, and only as the checker.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        f64, f32, u32, i32, vp = C.c_double, C.c_float, C.c_uint32, C.c_int, C.c_void_p
        sig = {
            "or_mulberry32_fill": (None, [u32, vp, C.c_size_t]),
            "or_pcg": (u32, [u32]),
            "or_hash3": (u32, [u32, u32, u32]),
            "or_u01": (f32, [u32]),
            "or_kr": (f32, [u32, f32, f32]),
            "or_distort": (f64, [vp, f64]),
            "or_distort_deriv": (f64, [vp, f64]),
            "or_undistort_theta": (f64, [vp, f64, i32]),
            "or_project": (None, [vp, vp, vp]),
            "or_unproject": (None, [vp, vp, vp]),
            "or_max_monotone_theta": (f64, [vp, f64]),
            "or_max_image_theta": (f64, [vp, i32, i32]),
            "or_unproject_f32": (None, [vp, vp, vp]),
            "or_build_bvh": (i32, [vp, vp, u32, i32, i32, vp, u32, vp, vp]),
            "or_pack_triangles": (None, [vp, vp, vp, u32, vp]),
            "or_sah_cost": (f64, [vp, u32, vp, vp, f64, f64]),
            "or_trace_batch": (None, [vp, vp, vp, u32, i32, vp, vp, i32]),
            "or_brute_batch": (None, [vp, u32, vp, u32, vp, i32]),
            "or_fresnel_dielectric": (f64, [f64, f64]),
            "or_fresnel_average": (f64, [f64, i32]),
            "or_ggx_d": (f64, [f64, f64]),
            "or_schlick": (f64, [f64, f64]),
            "or_ggx_lambda": (f64, [f64, f64]),
            "or_ggx_g1": (f64, [f64, f64]),
            "or_ggx_g2": (f64, [f64, f64, f64]),
            "or_eval_bsdf": (None, [vp, vp, vp, vp, vp]),
            "or_pdf_bsdf": (f64, [vp, vp, vp, vp]),
            "or_sample_bsdf": (f64, [vp, vp, vp, vp, vp]),
            "or_spec_prob": (f64, [vp, f64]),
            "or_mat_unpack": (None, [vp, vp]),
            "or_albedo_table": (None, [vp, i32, i32, vp, vp, vp, vp]),
            "or_radinv2": (f64, [u32]),
            "or_profile": (f64, [f64, vp, vp, i32]),
            "or_capsule_occ": (f64, [vp, vp, vp, vp, f64]),
            "or_gauss_legendre": (None, [i32, vp, vp]),
            "or_disc_irradiance": (f64, [vp, vp, vp, vp, f64, f64, i32, i32]),
            "or_expf": (f32, [f32]),
            "or_board_mask": (f32, [vp, u32, u32, f32, f32]),
            "or_logf": (f32, [f32]),
            "or_cosf": (f32, [f32]),
            "or_erfa": (f32, [f32]),
            "or_sensor_expose": (None, [vp, vp, vp, vp]),
            "or_glare_level": (None, [vp, vp, vp, vp]),
            "or_isp": (None, [vp, vp, vp, vp, vp, vp]),
            "or_isp_green": (None, [vp, vp, vp]),
            "or_isp_tone": (f32, [vp, f32]),
            "or_ri_map": (None, [vp, i32, i32, f64, i32, vp, vp, vp]),
            "or_raw_noise_dn": (f64, [f64, f64, f64, f64]),
            "or_render_tile": (None, [vp, vp, u32, u32, u32, u32, u32, u32, vp, vp, vp, i32]),
            "or_camera_ray": (None, [vp, u32, u32, u32, vp, vp]),
            "or_visible": (None, [vp, vp, vp, u32, f64, f64, vp]),
            "or_gather_build": (None, [vp, vp, vp, u32, u32, vp, i32]),
            "or_gather_remap": (None, [vp, vp, vp, u32, u32, vp, vp, i32]),
            "or_node_direct": (None, [vp, vp, vp, vp, vp, u32, u32, vp, vp, i32]),
            "or_capsule_vis": (f64, [vp, vp, vp]),
            "or_baked_occ": (None, [u32, u32, vp, vp, vp, vp, u32, vp, vp]),
            "or_baked_jacobi": (None, [u32, u32, vp, vp, vp, vp, vp, vp, vp, vp]),
            "or_jacobi_dense": (i32, [vp, vp, u32, u32, f64, i32, vp, vp]),
            "or_jacobi_gather": (None, [u32, vp, vp, u32, vp, vp, vp, vp, i32, vp, vp, i32, vp]),
            "or_jacobi_update": (None, [u32, vp, vp, vp, u32, vp, vp, vp, vp, i32]),
            "or_gather_dir": (None, [u32, u32, u32, vp, vp]),
            "or_backward_tile": (None, [vp, vp, u32, u32, u32, u32, u32, u32, vp, vp, vp, i32]),
            "or_camera_grad_tile": (None, [vp, vp, u32, u32, u32, u32, u32, u32, vp, C.c_double, vp, vp, i32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Lens(C.Structure):
    _fields_ = [("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
                ("alpha", C.c_double), ("k", C.c_double * 4)]


def lens(fx, fy, cx, cy, k, alpha=0.0):
    return Lens(fx, fy, cx, cy, alpha, (C.c_double * 4)(*k))


class Mat(C.Structure):
    _fields_ = [("kind", C.c_int), ("c", C.c_double * 3), ("alpha", C.c_double), ("roughness", C.c_double),
                ("ior", C.c_double), ("fbar", C.c_double), ("fbar_computed", C.c_double),
                ("alpha_c", C.c_double), ("tau", C.c_double * 3)]


def mat_from_record(rec):
    m = Mat()
    r = np.ascontiguousarray(rec, dtype=np.float32)
    lib().or_mat_unpack(ptr(r), C.byref(m))
    return m


class Sensor(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("cfa", C.c_uint32), ("frame", C.c_uint32),
                ("optics_k", C.c_float), ("esf_sigma", C.c_float), ("exposure", C.c_float), ("gain", C.c_float),
                ("m", C.c_float * 9), ("read_noise", C.c_float), ("full_well", C.c_float), ("dark", C.c_float),
                ("conv_gain", C.c_float), ("max_dn", C.c_float), ("black", C.c_float), ("noise_on", C.c_float),
                ("seed", C.c_uint32), ("cam_index", C.c_uint32), ("glare_k", C.c_float),
                ("glare_sigma", C.c_float), ("glare_f", C.c_uint32)]


class Isp(C.Structure):
    _fields_ = [("w", C.c_uint32), ("h", C.c_uint32), ("cfa", C.c_uint32), ("demosaic", C.c_uint32),
                ("nr", C.c_uint32), ("tone", C.c_uint32), ("chroma420", C.c_uint32), ("max_dn", C.c_float),
                ("black", C.c_float), ("read_noise", C.c_float), ("conv_gain", C.c_float), ("gain", C.c_float),
                ("wb", C.c_float * 3), ("ccm", C.c_float * 9), ("nr_sigma", C.c_float), ("nr_scale", C.c_float),
                ("gamma", C.c_float), ("s_c", C.c_float), ("black_off", C.c_float), ("sharpen_amount", C.c_float),
                ("sharpen_rad", C.c_int)]


class Scene(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("node_count", C.c_uint32), ("tris", C.c_void_p), ("tri_count", C.c_uint32),
                ("verts", C.c_void_p), ("vert_count", C.c_uint32), ("meshes", C.c_void_p), ("mesh_count", C.c_uint32),
                ("materials", C.c_void_p), ("mat_count", C.c_uint32), ("lights", C.c_void_p),
                ("light_count", C.c_uint32), ("emit", C.c_void_p), ("emit_count", C.c_uint32),
                ("emit_area", C.c_float), ("bodies", C.c_void_p), ("body_count", C.c_uint32),
                ("shadow_skip", C.c_uint32), ("light_prof", C.c_void_p), ("knots", C.c_void_p),
                ("node_state", C.c_void_p), ("alb", C.c_void_p), ("hh", C.c_void_p), ("slot", C.c_uint32),
                ("e_avg", C.c_float * 3), ("emit_c", C.c_float * 4), ("sample_materials", C.c_void_p),
                ("emit_posed", C.c_uint32), ("emit_body", C.c_uint32), ("uvs", C.c_void_p), ("tex", C.c_void_p),
                ("tex_w", C.c_uint32), ("tex_h", C.c_uint32)]


class Body(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("tris", C.c_void_p), ("tri_count", C.c_uint32), ("q0", C.c_float * 4),
                ("q1", C.c_float * 4), ("p0", C.c_float * 3), ("p1", C.c_float * 3), ("hidden", C.c_uint32),
                ("occ_a", C.c_float * 3), ("occ_b", C.c_float * 3), ("occ_r", C.c_float), ("rho_eff", C.c_float),
                ("bc_a", C.c_float * 4), ("bc_b", C.c_float * 4), ("cavity", C.c_uint32),
                ("cavity_origin", C.c_float * 3), ("cavity_axis", C.c_float * 3), ("cavity_s0", C.c_float),
                ("cavity_r", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("R", C.c_float * 9), ("pos", C.c_float * 3), ("mode", C.c_uint32), ("lens", C.c_float * 9),
                ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("seed", C.c_uint32),
                ("cam_index", C.c_uint32), ("light_samples", C.c_uint32), ("max_depth", C.c_uint32),
                ("flags", C.c_uint32), ("shutter", C.c_float * 2), ("skip_bodies", C.c_uint32),
                ("refine", C.c_uint32), ("refine_light_samples", C.c_uint32)]


# ---------------------------------------------------------------------------- helpers

def build_bvh(pos, idx, leaf_size=4, bins=16):
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
    idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    n = idx.size // 3
    cap = max(4, n + 2)
    nodes = np.zeros(cap * 16, dtype=np.float32)
    order = np.zeros(max(n, 1), dtype=np.uint32)

    class Info(C.Structure):
        _fields_ = [("node_count", C.c_uint32), ("max_depth", C.c_int), ("bmin", C.c_float * 3), ("bmax", C.c_float * 3)]
    info = Info()
    rc = lib().or_build_bvh(ptr(pos), ptr(idx), n, leaf_size, bins, ptr(nodes), cap, ptr(order), C.byref(info))
    assert rc == 0
    nc = info.node_count
    return dict(nodes=nodes[: nc * 16].copy(), node_count=nc, order=order[:n].copy(), max_depth=info.max_depth,
                bmin=np.array(info.bmin[:]), bmax=np.array(info.bmax[:]))


def pack_triangles(pos, idx, order):
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
    idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    order = np.ascontiguousarray(order, dtype=np.uint32)
    out = np.zeros(order.size * 16, dtype=np.float32)
    lib().or_pack_triangles(ptr(pos), ptr(idx), ptr(order), order.size, ptr(out))
    return out


def trace(nodes, tris, rays, any_hit=False, nthreads=8, visits=False):
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    n = rays.shape[0]
    hits = np.zeros((n, 4), dtype=np.float64)
    vis = np.zeros((n, 2), dtype=np.uint64) if visits else None
    lib().or_trace_batch(ptr(nodes), ptr(tris), ptr(rays), n, int(any_hit), ptr(hits),
                         ptr(vis) if visits else None, nthreads)
    return (hits, vis) if visits else hits


def brute(tris, count, rays, nthreads=8):
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    hits = np.zeros((rays.shape[0], 4), dtype=np.float64)
    lib().or_brute_batch(ptr(tris), count, ptr(rays), rays.shape[0], ptr(hits), nthreads)
    return hits


def _group_bvh(sc, members):
    """BVH2 (this oracle's builder) over the listed rows of sc["tris"], and those rows in its slot order."""
    tris = np.asarray(sc["tris"], np.float32).reshape(-1, 16)
    idx = tris[members].view(np.uint32)[:, [3, 7, 11]]
    pos = np.asarray(sc["verts"], np.float32).reshape(-1, 8)[:, :3]
    b = build_bvh(pos, idx)
    return b["nodes"], np.ascontiguousarray(tris[members[b["order"]]].reshape(-1))


def scene_struct(sc):
    """sc: dict of the upload arrays (nodes, tris, verts, meshes, materials, lights, emit, emit_area). With posed
    bodies (tri_body, body_count, poses: one dict(q0, p0[, q1, p1]) or None per body) every group's BVH is built
    here: the static scene over the rows whose tri_body is 0xffffffff, then each body over its own rows."""
    keep = {k: np.ascontiguousarray(sc[k]) for k in ("tris", "verts", "meshes", "materials", "lights", "emit")}
    nb = int(sc.get("body_count", 0))
    bodies = None
    if nb:
        tb = np.asarray(sc["tri_body"], np.uint32)
        keep["nodes"], keep["tris"] = _group_bvh(sc, np.nonzero(tb == 0xFFFFFFFF)[0])
        bodies = (Body * nb)()
        poses = sc.get("poses") or [None] * nb
        for b in range(nb):
            n, t = _group_bvh(sc, np.nonzero(tb == b)[0])
            keep["bn%d" % b], keep["bt%d" % b] = n, t
            B = bodies[b]
            B.nodes, B.tris, B.tri_count = ptr(n), ptr(t), t.size // 16
            P = poses[b]
            if P is None:
                B.q0[:] = B.q1[:] = [0.0, 0.0, 0.0, 1.0]
                B.hidden = 1
            else:
                B.q0[:] = [float(x) for x in P["q0"]]
                B.q1[:] = [float(x) for x in P.get("q1", P["q0"])]
                B.p0[:] = [float(x) for x in P["p0"]]
                B.p1[:] = [float(x) for x in P.get("p1", P["p0"])]
                B.occ_a[:] = [float(x) for x in P.get("occ_a", (0, 0, 0))]
                B.occ_b[:] = [float(x) for x in P.get("occ_b", (0, 0, 0))]
                B.occ_r = float(P.get("occ_r", 0.0))
                B.rho_eff = float(P.get("rho_eff", 0.0))
                cav = P.get("cavity")
                if cav is not None:
                    B.cavity = 1
                    B.cavity_origin[:] = [float(x) for x in cav["origin"]]
                    B.cavity_axis[:] = [float(x) for x in cav["axis"]]
                    B.cavity_s0 = float(cav["s0"])
                    B.cavity_r = float(cav.get("r", 0.0))
                ca, cb = body_capsule(t.reshape(-1, 16), P)
                B.bc_a[:] = [float(x) for x in ca]
                B.bc_b[:] = [float(x) for x in cb] + [0.0]
        keep["bodies"] = bodies
    else:
        keep["nodes"] = np.ascontiguousarray(sc["nodes"])
    s = Scene(ptr(keep["nodes"]), keep["nodes"].size // 16, ptr(keep["tris"]), keep["tris"].size // 16,
              ptr(keep["verts"]), keep["verts"].size // 8, ptr(keep["meshes"]), keep["meshes"].size // 4,
              ptr(keep["materials"]), keep["materials"].size // 16, ptr(keep["lights"]), keep["lights"].size // 4,
              ptr(keep["emit"]), keep["emit"].size // 16, float(sc["emit_area"]),
              C.cast(bodies, C.c_void_p) if nb else None, nb, int(sc.get("shadow_skip", 0)) if nb else 0)
    if nb and sc.get("emit_body") is not None:
        # the light rides on this body: emit records and light axes in its frame
        s.emit_posed, s.emit_body = 1, int(sc["emit_body"])
    if sc.get("profiles"):
        # {light: (axis, theta[], p[])}: the upload layout of vt_set_light_profile
        nl = keep["lights"].size // 4
        lp = np.zeros((nl, 4), np.float32)
        lp[:, 2] = 1.0
        kn = np.zeros((nl, 64, 2), np.float32)
        for light, (axis, th, p) in sc["profiles"].items():
            ax = np.asarray(axis, np.float64)
            lp[light, :3] = ax / np.linalg.norm(ax)
            lp[light].view(np.uint32)[3] = len(th)
            kn[light, : len(th), 0] = th
            kn[light, : len(th), 1] = p
        keep["light_prof"], keep["knots"] = lp, kn
        s.light_prof, s.knots = ptr(lp), ptr(kn)
    if sc.get("node_state") is not None:
        for k in ("node_state", "alb", "hh"):
            keep[k] = np.ascontiguousarray(sc[k], np.float32)
        s.node_state, s.alb, s.hh = ptr(keep["node_state"]), ptr(keep["alb"]), ptr(keep["hh"])
        s.slot = int(sc.get("slot", 1))
        s.e_avg[:] = [float(x) for x in sc.get("e_avg", (0, 0, 0))]
        s.emit_c[:] = [float(x) for x in sc.get("emit_c", emitter_centroid(sc))]
    if sc.get("sample_materials") is not None:
        keep["sample_materials"] = np.ascontiguousarray(sc["sample_materials"], np.float32)
        assert keep["sample_materials"].size == keep["materials"].size
        s.sample_materials = ptr(keep["sample_materials"])
    if sc.get("uvs") is not None:
        # the printed board: texture coordinates per vertex and the mask (255 = paper; one paper texel without it)
        keep["uvs"] = np.ascontiguousarray(sc["uvs"], np.float32)
        s.uvs = ptr(keep["uvs"])
        if sc.get("board_mask") is not None:
            keep["tex"] = np.ascontiguousarray(sc["board_mask"], np.uint8)
            s.tex, s.tex_h, s.tex_w = ptr(keep["tex"]), keep["tex"].shape[0], keep["tex"].shape[1]
    s._keep = keep
    return s


def board_mask(mask, u, v):
    """or_board_mask at each (u, v): the bilinear clamp-to-edge sample of an (h, w) uint8 mask, in [0, 1]."""
    m = np.ascontiguousarray(mask, np.uint8)
    u = np.atleast_1d(np.asarray(u, np.float32))
    v = np.atleast_1d(np.asarray(v, np.float32))
    out = np.empty(u.shape, np.float32)
    f = lib().or_board_mask
    for k in range(u.size):
        out.flat[k] = f(ptr(m), m.shape[1], m.shape[0], float(u.flat[k]), float(v.flat[k]))
    return out


def _quat_mat(q):
    x, y, z, w = np.asarray(q, np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _slerp(a, b, t):
    a = np.asarray(a, np.float64) / np.linalg.norm(a)
    b = np.asarray(b, np.float64) / np.linalg.norm(b)
    c = a @ b
    if c < 0:
        b, c = -b, -c
    if c > 0.9995:
        q = a + (b - a) * t
        return q / np.linalg.norm(q)
    th = math.acos(c)
    return (a * math.sin((1 - t) * th) + b * math.sin(t * th)) / math.sin(th)


def body_capsule(tris, P):
    """The bounding capsule a posed body is culled with (the library's: its box's longest axis posed at
    mid-exposure, radius padded by the largest motion of its points over the exposure): (a.xyz + r, b.xyz)."""
    tr = np.asarray(tris, np.float32).reshape(-1, 16)
    v0 = tr[:, 0:3].astype(np.float64)
    pts = np.concatenate([v0, v0 + tr[:, 4:7], v0 + tr[:, 8:11]]).astype(np.float32).astype(np.float64)
    lo, hi = pts.min(0), pts.max(0)
    q0 = np.asarray(P["q0"], np.float64)
    q1 = np.asarray(P.get("q1", P["q0"]), np.float64)
    p0 = np.asarray(P["p0"], np.float64)
    p1 = np.asarray(P.get("p1", P["p0"]), np.float64)
    q0n, q1n = q0 / np.linalg.norm(q0), q1 / np.linalg.norm(q1)
    moving = bool(np.any(q0n.astype(np.float32) != q1n.astype(np.float32)) or np.any(p0 != p1))
    ctr, half = 0.5 * (lo + hi), 0.5 * (hi - lo)
    rad = np.linalg.norm(ctr) + np.linalg.norm(half)
    phi = 2 * math.acos(min(1.0, abs(float(q0n @ q1n)))) if moving else 0.0
    ax = (0 if half[0] >= half[2] else 2) if half[0] >= half[1] else (1 if half[1] >= half[2] else 2)
    ea, eb = ctr.copy(), ctr.copy()
    ea[ax] -= half[ax]
    eb[ax] += half[ax]
    rr = math.sqrt(sum(half[k] ** 2 for k in range(3) if k != ax))
    t = 0.5 if moving else 0.0
    R = _quat_mat(_slerp(q0n, q1n, t))
    pm = p0 + (p1 - p0) * t
    wa, wb = R @ ea + pm, R @ eb + pm
    motion = 1.01 * (2 * math.sin(phi / 4) * rad + 0.5 * np.linalg.norm(p1 - p0)) if moving else 0.0
    scale = max(np.abs(lo).max(), np.abs(hi).max(), 1e-3)
    return np.concatenate([wa, [rr * 1.0001 + motion + 4e-6 * scale + 1e-6]]), wb


def emitter_centroid(sc):
    """centroid and radius of the emitter triangles (gpu/scene_gpu.js:418-422)"""
    e = np.asarray(sc["emit"], np.float32).reshape(-1, 16).astype(np.float64)
    if e.shape[0] == 0:
        return (0.0, 0.0, 0.0, 0.0)
    cen = ((e[:, 0:3] + (e[:, 4:7] + e[:, 8:11]) / 3) * e[:, 7:8]).sum(0) / max(e[:, 7].sum(), 1e-300)
    pts = np.concatenate([e[:, 0:3], e[:, 0:3] + e[:, 4:7], e[:, 0:3] + e[:, 8:11]])
    return tuple(cen) + (float(np.linalg.norm(pts - cen, axis=1).max()),)


def camera_struct(cam):
    c = Camera()
    c.R[:] = [float(x) for x in cam["R"]]
    c.pos[:] = [float(x) for x in cam["pos"]]
    c.mode = int(cam.get("mode", 0))
    c.lens[:] = [float(x) for x in cam["lens"]]
    for k in ("width", "height", "spp", "seed", "cam_index", "light_samples", "max_depth", "flags"):
        setattr(c, k, int(cam.get(k, 0)))
    c.shutter[:] = [float(x) for x in cam.get("shutter", (0.0, 1.0))]
    c.skip_bodies = int(cam.get("skip_bodies", 0))
    c.refine = int(cam.get("refine", 0))
    c.refine_light_samples = int(cam.get("refine_light_samples", 0))
    return c


def render_tile(sc, cam, x0, y0, w, h, spp_begin=0, spp_end=None, nthreads=8):
    S = scene_struct(sc)
    Cm = camera_struct(cam)
    spp_end = cam["spp"] if spp_end is None else spp_end
    out = np.zeros((h, w, 4), dtype=np.float32)
    rays = C.c_uint64(0)
    vis = np.zeros(5, dtype=np.uint64)
    lib().or_render_tile(C.byref(S), C.byref(Cm), x0, y0, w, h, spp_begin, spp_end, ptr(out), C.byref(rays), ptr(vis),
                         nthreads)
    return out, int(rays.value), vis


def visible(sc, cam, points, tol=2e-4, time=0.5):
    """FP64 visible query (gpu/renderer.js:1238-1258) of world points from the camera centre."""
    S = scene_struct(sc)
    Cm = camera_struct(cam)
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(max(1, pts.shape[0]), np.uint8)
    lib().or_visible(C.byref(S), C.byref(Cm), ptr(pts), pts.shape[0], float(tol), float(time), ptr(out))
    return out[: pts.shape[0]].astype(bool)


def camera_ray(cam, x, y, i):
    Cm = camera_struct(cam)
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib().or_camera_ray(C.byref(Cm), x, y, i, ptr(o), ptr(d))
    return o, d


def sensor_expose(params, radiance, ri):
    radiance = np.ascontiguousarray(radiance, dtype=np.float32)
    ri = np.ascontiguousarray(ri, dtype=np.float32)
    h, w = ri.shape
    dn = np.zeros((h, w), dtype=np.uint16)
    s = sensor_struct(params, w, h)
    lib().or_sensor_expose(C.byref(s), ptr(radiance), ptr(ri), ptr(dn))
    return dn


def sensor_struct(p, w, h):
    s = Sensor()
    s.width, s.height, s.cfa, s.frame = w, h, int(p["cfa"]), int(p["frame"])
    s.optics_k, s.esf_sigma, s.exposure, s.gain = p["optics_k"], p["esf_sigma"], p["exposure"], p["gain"]
    s.m[:] = [float(x) for x in np.asarray(p["m"]).reshape(-1)]
    s.read_noise, s.full_well, s.dark, s.conv_gain = p["read_noise"], p["full_well"], p["dark"], p["conv_gain"]
    s.max_dn, s.black, s.noise_on = p["max_dn"], p["black"], 1.0 if p["noise_on"] else 0.0
    s.seed, s.cam_index = int(p["seed"]), int(p["cam_index"])
    s.glare_k, s.glare_sigma = p.get("glare_fraction", 0.0), p.get("glare_sigma", 1.0)
    s.glare_f = int(p.get("glare_level", 16))
    return s


def glare_level(params, radiance, ri):
    """the blurred glare level (gh, gw, 3)"""
    radiance = np.ascontiguousarray(radiance, dtype=np.float32)
    ri = np.ascontiguousarray(ri, dtype=np.float32)
    h, w = ri.shape
    s = sensor_struct(params, w, h)
    F = s.glare_f
    out = np.zeros(((h + F - 1) // F, (w + F - 1) // F, 3), np.float32)
    lib().or_glare_level(C.byref(s), ptr(radiance), ptr(ri), ptr(out))
    return out


ISP_DEFAULTS = dict(demosaic=0, nr=0, tone=0, chroma420=1, wb=(1.0, 1.0, 1.0),
                    ccm=(1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0), gamma=2.2, s_curve_c=1.0, black_offset=0.0,
                    nr_sigma_px=0.4, nr_range_scale=2.0, sharpen_amount=0.0, sharpen_radius=1.0)


def isp_struct(sensor, isp, w, h):
    """or_isp from the raw stage's parameters (cfa, max_dn, black, read_noise, conv_gain, gain) and the ISP's
    (keys of ISP_DEFAULTS)"""
    q = dict(ISP_DEFAULTS, **isp)
    s = Isp()
    s.w, s.h, s.cfa = w, h, int(sensor["cfa"])
    s.demosaic, s.tone, s.chroma420 = int(q["demosaic"]), int(q["tone"]), 1 if q["chroma420"] else 0
    s.nr = 1 if (q["nr"] and q["nr_sigma_px"] > 0) else 0
    s.max_dn, s.black, s.read_noise = sensor["max_dn"], sensor["black"], sensor["read_noise"]
    s.conv_gain, s.gain = sensor["conv_gain"], sensor["gain"]
    s.wb[:] = [float(x) for x in q["wb"]]
    s.ccm[:] = [float(x) for x in np.asarray(q["ccm"]).reshape(-1)]
    s.nr_sigma, s.nr_scale, s.gamma = q["nr_sigma_px"], q["nr_range_scale"], q["gamma"]
    s.s_c, s.black_off, s.sharpen_amount = q["s_curve_c"], q["black_offset"], q["sharpen_amount"]
    r = np.round(np.float32(q["sharpen_radius"]))
    s.sharpen_rad = int(min(4, max(1, r)))
    return s


def isp(sensor, isp_params, dn):
    """the ISP on raw DN (h, w): dict(rgba (h, w, 4) u8, y, cb, cr planes (4:2:0 only), stats u32[4], lin (h, w, 4))"""
    dn = np.ascontiguousarray(dn, dtype=np.uint16)
    h, w = dn.shape
    s = isp_struct(sensor, isp_params, w, h)
    cw, ch = (w + 1) // 2, (h + 1) // 2
    rgba = np.zeros((h, w, 4), np.uint8)
    yuv = np.zeros(w * h + 2 * cw * ch, np.uint8)
    stats = np.zeros(4, np.uint32)
    lin = np.zeros((h, w, 4), np.float32)
    lib().or_isp(C.byref(s), ptr(dn), ptr(rgba), ptr(yuv), ptr(stats), ptr(lin))
    out = dict(rgba=rgba, stats=stats, lin=lin, yuv=yuv)
    if s.chroma420:
        out.update(y=yuv[:w * h].reshape(h, w), cb=yuv[w * h:w * h + cw * ch].reshape(ch, cw),
                   cr=yuv[w * h + cw * ch:].reshape(ch, cw))
    return out


def isp_green(sensor, isp_params, dn):
    dn = np.ascontiguousarray(dn, dtype=np.uint16)
    h, w = dn.shape
    s = isp_struct(sensor, isp_params, w, h)
    g = np.zeros((h, w), np.float32)
    lib().or_isp_green(C.byref(s), ptr(dn), ptr(g))
    return g


def isp_tone(isp_params, x):
    s = isp_struct(dict(cfa=0, max_dn=4095.0, black=0.0, read_noise=0.0, conv_gain=1.0, gain=1.0), isp_params, 1, 1)
    return np.array([lib().or_isp_tone(C.byref(s), float(v)) for v in np.asarray(x, np.float32).reshape(-1)],
                    np.float32)


def ri_map(L, w, h, pupil_exp=1.0, vignetting_theta_deg=(), vignetting=()):
    """gpu/sensor.js:103-124 relativeIlluminationMap: the KB4 term times the vignetting table (degrees)"""
    out = np.zeros((h, w), dtype=np.float32)
    vt = np.radians(np.asarray(vignetting_theta_deg, np.float64))
    vv = np.asarray(vignetting, np.float64)
    assert vt.size == vv.size
    lib().or_ri_map(C.byref(L), w, h, float(pupil_exp), int(vt.size), ptr(vt) if vt.size else None,
                    ptr(vv) if vv.size else None, ptr(out))
    return out


def backward_tile(sc, cam, x0, y0, w, h, adj, spp_begin=0, spp_end=None, nthreads=8):
    """FP64 path-replay gradients of sum adj . I over a tile: (materials x 3, lights x 3)."""
    S = scene_struct(sc)
    Cm = camera_struct(cam)
    spp_end = cam["spp"] if spp_end is None else spp_end
    adj = np.ascontiguousarray(adj, dtype=np.float32)
    gm = np.zeros(S.mat_count * 3)
    gl = np.zeros(max(S.light_count, 1) * 3)
    lib().or_backward_tile(C.byref(S), C.byref(Cm), x0, y0, w, h, spp_begin, spp_end, ptr(adj), ptr(gm), ptr(gl), nthreads)
    return gm.reshape(-1, 3), gl[: 3 * S.light_count].reshape(-1, 3)


def camera_grad_tile(sc, cam, x0, y0, w, h, adj, rel_step=1e-6, spp_begin=0, spp_end=None, nthreads=8):
    """FP64 central-difference gradient of sum adj . I over a tile with respect to the camera's 14 parameters
    (fx fy cx cy k1..k4, rotation about the world axes, centre), interior term: per sample, with the random
    numbers and the triangles held fixed. Returns (grad[14], pairs left out because the path changed)."""
    S = scene_struct(sc)
    Cm = camera_struct(cam)
    spp_end = cam["spp"] if spp_end is None else spp_end
    adj = np.ascontiguousarray(adj, dtype=np.float32)
    g = np.zeros(14)
    changed = np.zeros(1, np.uint64)
    lib().or_camera_grad_tile(C.byref(S), C.byref(Cm), x0, y0, w, h, spp_begin, spp_end, ptr(adj), rel_step, ptr(g),
                              ptr(changed), nthreads)
    return g, int(changed[0])


# ---------------------------------------------------------------------------- node radiosity cache

def node_inputs(sc):
    """Per vertex: npos (pos + static flag), nnorm (unit normal), node material and area, as the library's
    vt_radiosity_build derives them: static nodes are the vertices of static-flagged triangles of the static scene,
    the material that of the vertex's own mesh (vertex word 3), a third of every such triangle's area."""
    verts = np.asarray(sc["verts"], np.float32).reshape(-1, 8)
    tris = np.asarray(sc["tris"], np.float32).reshape(-1, 16)
    tu = tris.view(np.uint32)
    meshes = np.asarray(sc["meshes"], np.uint32).reshape(-1, 4)
    V = verts.shape[0]
    sel = (meshes[tu[:, 12], 1] & 1) != 0
    if int(sc.get("body_count", 0)):
        sel &= np.asarray(sc["tri_body"], np.uint32) == 0xFFFFFFFF
    ids = tu[sel][:, [3, 7, 11]]
    stat = np.zeros(V, bool)
    stat[ids.reshape(-1)] = True
    P = verts[:, 0:3].astype(np.float64)
    a = P[ids[:, 0]]
    ar = 0.5 * np.linalg.norm(np.cross(P[ids[:, 1]] - a, P[ids[:, 2]] - a), axis=1) / 3
    area = np.zeros(V)
    for k in range(3):
        np.add.at(area, ids[:, k], ar)
    nmat = np.where(stat, meshes[verts.view(np.uint32)[:, 3] % max(meshes.shape[0], 1), 0], 0).astype(np.uint32)
    n = verts[:, 4:7].astype(np.float64)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    n = np.where(ln > 0, n / np.where(ln > 0, ln, 1), np.array([0.0, 0.0, 1.0]))
    npos = np.zeros((V, 4), np.float32)
    npos[:, 0:3] = verts[:, 0:3]
    npos[:, 3] = stat
    nnorm = np.zeros((V, 4), np.float32)
    nnorm[:, 0:3] = n
    return dict(npos=npos, nnorm=nnorm, nmat=nmat, static=stat, area=area)


def albedo_tables(sc, samples=1024, bins=32):
    """FP64 albedo tables of every material (albedoTable, gpu/bsdf.js:210-254) in the device layout: per
    material bins x [spec.rgb, -, diff.rgb, -], and the hemispherical albedo (4 floats per material)."""
    mats = np.asarray(sc["materials"], np.float32).reshape(-1, 16)
    alb = np.zeros((mats.shape[0], bins, 8), np.float32)
    hh = np.zeros((mats.shape[0], 4), np.float32)
    for m in range(mats.shape[0]):
        M = mat_from_record(mats[m])
        spec = np.zeros(bins * 3)
        diff = np.zeros(bins * 3)
        h = np.zeros(3)
        hs = np.zeros(3)
        lib().or_albedo_table(C.byref(M), bins, samples, ptr(spec), ptr(diff), ptr(h), ptr(hs))
        alb[m, :, 0:3] = spec.reshape(bins, 3)
        alb[m, :, 4:7] = diff.reshape(bins, 3)
        hh[m, 0:3] = h
    return alb, hh


def gather_table(sc, ni, K, nthreads=8):
    """gatherBuild + gatherCount (kernels.wgsl.js:283-333) over the static scene: (table V x K, misses)."""
    S = scene_struct(sc)
    V = ni["npos"].shape[0]
    g = np.zeros(V * K, np.uint32)
    lib().or_gather_build(C.byref(S), ptr(ni["npos"]), ptr(ni["nnorm"]), V, K, ptr(g), nthreads)
    miss = np.where(ni["static"], (g.reshape(V, K) == 0xFFFFFFFF).sum(1), K).astype(np.float32)
    return g, miss


def gather_remap(sc, ni, K, gtable, miss, nthreads=8):
    S = scene_struct(sc)
    g = np.ascontiguousarray(gtable, np.uint32).copy()
    lib().or_gather_remap(C.byref(S), ptr(ni["npos"]), ptr(ni["nnorm"]), ni["npos"].shape[0], K,
                          ptr(np.ascontiguousarray(miss, np.float32)), ptr(g), nthreads)
    return g


def node_direct(sc, ni, active, samples, alb, state, nthreads=8):
    """nodeDirect (kernels.wgsl.js:348-406) into state (V x 16, in place)."""
    S = scene_struct(sc)
    act = np.ascontiguousarray(active, np.uint32)
    lib().or_node_direct(C.byref(S), ptr(ni["npos"]), ptr(ni["nnorm"]), ptr(ni["nmat"]), ptr(act), act.size,
                         samples, ptr(np.ascontiguousarray(alb, np.float32)), ptr(state), nthreads)


def baked_occ(table, caps, with_dist=False):
    """gatherOcc (lite/shaders.js:234-258) on a baked table dict(nodes, K, gather uint16 (nodes, K), node_a, node_n
    (nodes, 4) float32 with the mesh id bits in node_n[:, 3], hh (meshes, 4)); caps (n, 8) [a, r, b, rho].
    Returns the transmission bytes (nodes, K) and, with_dist, each ray's closest capsule distance minus radius."""
    n, K = int(table["nodes"]), int(table["K"])
    g = np.ascontiguousarray(table["gather"], np.uint16)
    a = np.ascontiguousarray(table["node_a"], np.float32)
    nn = np.ascontiguousarray(table["node_n"], np.float32)
    cp = np.ascontiguousarray(np.asarray(caps, np.float32).reshape(-1, 8))
    occ = np.zeros((n, K), np.uint8)
    dist = np.zeros((n, K), np.float64) if with_dist else None
    lib().or_baked_occ(n, K, ptr(g), ptr(a), ptr(nn), ptr(cp), cp.shape[0], ptr(occ), ptr(dist) if with_dist else None)
    return (occ, dist) if with_dist else occ


def baked_jacobi(table, occ, e_dir, b, iters):
    """`iters` baked Jacobi steps (lite/shaders.js:260-283) from B = b; returns (B, E_ind of the last step)."""
    n, K = int(table["nodes"]), int(table["K"])
    g = np.ascontiguousarray(table["gather"], np.uint16)
    nn = np.ascontiguousarray(table["node_n"], np.float32)
    hh = np.ascontiguousarray(table["hh"], np.float32)
    o = np.ascontiguousarray(occ, np.uint8)
    ed = np.ascontiguousarray(e_dir, np.float32)
    cur = np.ascontiguousarray(b, np.float32).copy()
    nxt = np.zeros_like(cur)
    ei = np.zeros((n, 4), np.float32)
    for _ in range(iters):
        lib().or_baked_jacobi(n, K, ptr(g), ptr(o), ptr(nn), ptr(hh), ptr(ed), ptr(cur), ptr(nxt), ptr(ei))
        cur, nxt = nxt, cur
    return cur, ei


def jacobi_steps(sc, ni, active, gtable, K, nbr, nbr_w, alb, hh, state, slot, iters):
    """`iters` steps of jacobiGather + jacobiUpdate (kernels.wgsl.js:416-461), ping-pong source slots; returns the
    slot holding the last sources."""
    verts = np.ascontiguousarray(np.asarray(sc["verts"], np.float32).reshape(-1, 8))
    act = np.ascontiguousarray(active, np.uint32)
    alb = np.asarray(alb, np.float32)
    spec = np.ascontiguousarray(alb[:, :, 0:3])
    diff = np.ascontiguousarray(alb[:, :, 4:7])
    h3 = np.ascontiguousarray(np.asarray(hh, np.float32)[:, 0:3])
    nb = np.ascontiguousarray(nbr, np.uint32)
    nw = np.ascontiguousarray(nbr_w, np.float32)
    nbr_k = nb.shape[1] if nb.ndim == 2 else 0
    aux = np.zeros((max(act.size, 1), 4), np.float32)
    g = np.ascontiguousarray(gtable, np.uint32)
    for _ in range(iters):
        lib().or_jacobi_gather(act.size, ptr(act), ptr(g), K, ptr(verts), ptr(ni["nmat"]), ptr(spec), ptr(diff),
                               alb.shape[1], ptr(h3), ptr(state), slot, ptr(aux))
        lib().or_jacobi_update(act.size, ptr(act), ptr(nb), ptr(nw), nbr_k, ptr(aux), ptr(ni["nmat"]), ptr(h3),
                               ptr(state), 3 - slot)
        slot = 3 - slot
    return slot


def e_avg(ni, active, state, hh):
    """nodeSummary's eAvg (gpu/renderer.js:824-852): area-weighted E_free V + B_dir / A_hh over the active nodes."""
    act = np.asarray(active, np.int64)
    a = ni["area"][act][:, None]
    st = np.asarray(state, np.float64)
    h = np.maximum(np.asarray(hh, np.float64)[ni["nmat"][act], 0:3], 1e-3)
    A = a.sum()
    E = (a * st[act, 0:3] * st[act, 3:4]).sum(0)
    Ed = (a * st[act, 12:15] / h).sum(0)
    return (E + Ed) / A if A > 0 else np.zeros(3)
