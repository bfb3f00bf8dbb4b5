"""ctypes binding of libvt.so (include/vt.h). Raises when the HIP library is missing: there is no CPU fallback."""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBPATH = os.path.join(PKG, "libvt.so")

VT_OK, VT_ERR_ARG, VT_ERR_SIZE, VT_ERR_HIP, VT_ERR_LOST, VT_ERR_STATE, VT_ERR_NOMEM = 0, -1, -2, -3, -4, -5, -6
VT_RENDER_BATCHED, VT_RENDER_PER_CAMERA, VT_RENDER_WAVEFRONT, VT_RENDER_AUTO = 0, 1, 2, 3
VT_KIND_LINEAR, VT_KIND_RAW, VT_KIND_RI = 0, 1, 2
VT_KIND_RGB8, VT_KIND_RGBA8, VT_KIND_YUV420, VT_KIND_STATS = 3, 4, 5, 6
VT_DENSE_FP32, VT_DENSE_BF16_REFINE, VT_DENSE_FP64 = 0, 1, 2

EXPORTS = [
    "vt_create", "vt_destroy", "vt_last_error", "vt_build_bvh", "vt_upload_scene", "vt_set_camera", "vt_render",
    "vt_sensor", "vt_readback", "vt_sync", "vt_frame_stats_get", "vt_trace_test", "vt_lens_test",
    "vt_camera_ray_test", "vt_bsdf_test", "vt_sensor_test", "vt_jacobi_sparse", "vt_dense_upload",
    "vt_dense_formfactors", "vt_dense_solve", "vt_dense_download_F", "vt_backward", "vt_update_materials",
    "vt_update_lights", "vt_autotune", "vt_copy_bandwidth", "vt_read_bandwidth", "vt_set_instances", "vt_visible",
    "vt_set_light_profile", "vt_set_board_texture", "vt_radiosity_build", "vt_interior_bvh", "vt_radiosity_update", "vt_radiosity_update_moved", "vt_baked_upload", "vt_baked_occlusion", "vt_baked_jacobi", "vt_baked_read", "vt_radiosity_state", "vt_node_neighbours",
    "vt_radiosity_tables", "vt_capsule_test", "vt_isp", "vt_isp_test", "vt_weld_subdivide", "vt_sensor_raw_device",
]

# camera option bits (vt_camera.flags)
VT_OPT_NO_INDIRECT, VT_OPT_DIRECT_ONLY, VT_OPT_NO_CAPSULE, VT_OPT_NORMALS, VT_OPT_NO_MIS = 4, 8, 32, 64, 128
VT_OPT_REFERENCE, VT_OPT_CULL_SHADOWS, VT_OPT_REFINE_SHADOWS = 1 << 16, 1 << 17, 1024
VT_OPT_HIT_CACHE, VT_OPT_SHADOW_CACHE = 2048, 8192

vp, u32, i32, f32, f64, sz = C.c_void_p, C.c_uint32, C.c_int, C.c_float, C.c_double, C.c_size_t
CAMERA_PARAMS = 14      # VT_CAMERA_PARAMS: fx fy cx cy k1 k2 k3 k4, rotation rx ry rz, centre tx ty tz


class CreateInfo(C.Structure):
    _fields_ = [("device", C.c_int), ("flags", u32)]


class Scene(C.Structure):
    _fields_ = [("nodes", vp), ("node_count", u32), ("tris", vp), ("tri_count", u32), ("verts", vp),
                ("vert_count", u32), ("meshes", vp), ("mesh_count", u32), ("materials", vp),
                ("material_count", u32), ("lights", vp), ("light_count", u32), ("emit", vp), ("emit_count", u32),
                ("emit_area", f32), ("tri_body", vp), ("body_count", u32), ("shadow_skip", u32),
                ("emit_posed", u32), ("emit_body", u32), ("uvs", vp)]


MAX_BODIES = 24
STATIC_BODY = 0xFFFFFFFF


class Instance(C.Structure):
    _fields_ = [("q0", f32 * 4), ("q1", f32 * 4), ("p0", f32 * 3), ("p1", f32 * 3), ("hidden", u32),
                ("occ_a", f32 * 3), ("occ_b", f32 * 3), ("occ_r", f32), ("rho_eff", f32),
                ("cavity", u32), ("cavity_origin", f32 * 3), ("cavity_axis", f32 * 3), ("cavity_s0", f32),
                ("cavity_r", f32)]


class BakedTable(C.Structure):
    _fields_ = [("nodes", u32), ("K", u32), ("gather", vp), ("node_a", vp), ("node_n", vp), ("hh", vp), ("meshes", u32)]


class RadiosityOpts(C.Structure):
    _fields_ = [("gather_k", u32), ("node_samples", u32), ("nbr_k", u32), ("smooth_radius", f32),
                ("smooth_sigma", f32)]


class RadiosityInfo(C.Structure):
    _fields_ = [("nodes", u32), ("static_nodes", u32), ("active", u32), ("ms_build", f64)]


class Camera(C.Structure):
    _fields_ = [("R", f32 * 9), ("pos", f32 * 3), ("mode", u32), ("lens", f32 * 9), ("width", u32), ("height", u32),
                ("spp", u32), ("seed", u32), ("cam_index", u32), ("light_samples", u32), ("max_depth", u32),
                ("flags", u32), ("scale", f32), ("shutter", f32 * 2), ("skip_bodies", u32), ("refine", u32),
                ("refine_light_samples", u32)]


class RenderOpts(C.Structure):
    _fields_ = [("mode", u32), ("spp_begin", u32), ("spp_end", u32), ("accumulate", u32), ("chunk", u32),
                ("block", u32), ("stack_lds", u32), ("queues", u32), ("graph", u32), ("batch", u32),
                ("sort_bits", u32), ("shadow_sort", u32)]


class TuneResult(C.Structure):
    _fields_ = [("block", u32), ("stack_lds", u32), ("chunk", u32), ("queues", u32), ("candidates", u32),
                ("probe_mrays_best", f64), ("probe_mrays_default", f64), ("probe_fraction", f64), ("mode", u32),
                ("wf_batch", u32), ("wf_stack_lds", u32), ("wf_refill", u32), ("probe_mrays_wavefront", f64),
                ("wf_sort_bits", u32)]


class SensorParams(C.Structure):
    _fields_ = [("cfa", u32), ("frame", u32), ("seed", u32), ("optics_k", f32), ("esf_sigma", f32),
                ("exposure", f32), ("gain", f32), ("m", f32 * 9), ("read_noise", f32), ("full_well", f32),
                ("dark", f32), ("conv_gain", f32), ("max_dn", f32), ("black", f32), ("noise_on", u32),
                ("pupil_exp", f32), ("glare_fraction", f32), ("glare_sigma", f32), ("glare_level", u32),
                ("vignetting_count", u32), ("vignetting_theta_deg", f32 * 8), ("vignetting", f32 * 8)]


class IspParams(C.Structure):
    _fields_ = [("demosaic", u32), ("nr", u32), ("tone", u32), ("chroma420", u32), ("wb", f32 * 3), ("ccm", f32 * 9),
                ("gamma", f32), ("s_curve_c", f32), ("black_offset", f32), ("nr_sigma_px", f32),
                ("nr_range_scale", f32), ("sharpen_amount", f32), ("sharpen_radius", f32)]


ISP_DEFAULTS = dict(demosaic=0, nr=0, tone=0, chroma420=1, wb=(1.0, 1.0, 1.0),
                    ccm=(1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0), gamma=2.2, s_curve_c=1.0, black_offset=0.0,
                    nr_sigma_px=0.4, nr_range_scale=2.0, sharpen_amount=0.0, sharpen_radius=1.0)


def isp_struct(p):
    q = dict(ISP_DEFAULTS, **p)
    s = IspParams()
    s.demosaic, s.nr, s.tone, s.chroma420 = int(q["demosaic"]), int(bool(q["nr"])), int(q["tone"]), int(bool(q["chroma420"]))
    s.wb[:] = [float(x) for x in q["wb"]]
    s.ccm[:] = [float(x) for x in np.asarray(q["ccm"]).reshape(-1)]
    s.gamma, s.s_curve_c, s.black_offset = q["gamma"], q["s_curve_c"], q["black_offset"]
    s.nr_sigma_px, s.nr_range_scale = q["nr_sigma_px"], q["nr_range_scale"]
    s.sharpen_amount, s.sharpen_radius = q["sharpen_amount"], q["sharpen_radius"]
    return s


class FrameStats(C.Structure):
    _fields_ = [("rays", f64), ("samples", f64), ("ms_render", f64), ("ms_sensor", f64), ("ms_kernel_avg", f64),
                ("launches", u32), ("stack_overflows", u32), ("ms_backward", f64), ("rays_backward", f64),
                ("graph_used", u32), ("graph_builds", u32), ("rays_extension", f64), ("rays_shadow", f64),
                ("ms_ext_avg", f64), ("ext_launches", u32), ("ms_isp", f64)]


class Grads(C.Structure):
    _fields_ = [("materials", vp), ("material_count", u32), ("lights", vp), ("light_count", u32), ("cameras", vp),
                ("camera_count", u32)]


class GatherTables(C.Structure):
    _fields_ = [("n_static", u32), ("n_active", u32), ("gather_k", u32), ("nbr_k", u32), ("n_mat", u32),
                ("albedo_bins", u32), ("active", vp), ("gtable", vp), ("normals", vp), ("node_mat", vp),
                ("albedo", vp), ("hh", vp), ("nbr", vp), ("nbr_w", vp)]


class DenseOpts(C.Structure):
    _fields_ = [("n", u32), ("r", u32), ("method", u32), ("tol", f64), ("max_sweeps", u32), ("inner_sweeps", u32)]


class DenseResult(C.Structure):
    _fields_ = [("residual", f64), ("sweeps", u32), ("refinements", u32), ("ms", f64), ("ms_per_sweep", f64)]


_LIB = None


def load():
    """Load libvt.so; raises if it has not been built (HIP library required, no fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise RuntimeError("libvt.so is not built (run __graft_entry__.build()); the HIP path has no fallback")
    L = C.CDLL(LIBPATH)
    sig = {
        "vt_create": (i32, [vp, vp]),
        "vt_destroy": (None, [vp]),
        "vt_last_error": (C.c_char_p, [vp]),
        "vt_build_bvh": (i32, [vp, u32, vp, u32, i32, i32, vp, u32, vp, vp]),
        "vt_upload_scene": (i32, [vp, vp]),
        "vt_set_board_texture": (i32, [vp, vp, u32, u32]),
        "vt_set_camera": (i32, [vp, u32, vp]),
        "vt_render": (i32, [vp, vp, u32, vp]),
        "vt_sensor": (i32, [vp, u32, vp]),
        "vt_readback": (i32, [vp, u32, i32, vp, sz]),
        "vt_sync": (i32, [vp]),
        "vt_frame_stats_get": (i32, [vp, vp]),
        "vt_trace_test": (i32, [vp, vp, u32, i32, vp]),
        "vt_lens_test": (i32, [vp, vp, vp, u32, vp]),
        "vt_camera_ray_test": (i32, [vp, u32, vp, u32, vp]),
        "vt_bsdf_test": (i32, [vp, vp, u32, vp]),
        "vt_sensor_test": (i32, [vp, vp, u32, u32, vp, vp, vp]),
        "vt_sensor_raw_device": (i32, [vp, vp, u32, u32, vp, vp, vp]),
        "vt_isp": (i32, [vp, u32, vp]),
        "vt_weld_subdivide": (i32, [vp, vp, u32, vp, u32, f64, u32, vp, vp, u32, vp, u32, vp, vp]),
        "vt_isp_test": (i32, [vp, vp, vp, u32, u32, vp, vp, vp, vp]),
        "vt_jacobi_sparse": (i32, [vp, vp, vp, u32, vp]),
        "vt_dense_upload": (i32, [vp, u32, u32, vp, vp]),
        "vt_dense_formfactors": (i32, [vp, u32, vp, f32, u32, vp]),
        "vt_dense_solve": (i32, [vp, vp, vp, vp]),
        "vt_dense_download_F": (i32, [vp, vp, sz]),
        "vt_backward": (i32, [vp, vp, u32, vp, vp, vp]),
        "vt_update_materials": (i32, [vp, vp, u32]),
        "vt_update_lights": (i32, [vp, vp, u32]),
        "vt_autotune": (i32, [vp, vp, u32, vp, u32, vp]),
        "vt_copy_bandwidth": (i32, [vp, C.c_uint64, u32, vp]),
        "vt_read_bandwidth": (i32, [vp, C.c_uint64, u32, vp]),
        "vt_set_instances": (i32, [vp, vp, u32]),
        "vt_visible": (i32, [vp, u32, vp, u32, f32, f32, vp]),
        "vt_set_light_profile": (i32, [vp, u32, vp, vp, vp, u32]),
        "vt_radiosity_build": (i32, [vp, vp, vp]),
        "vt_interior_bvh": (i32, [vp, u32, vp]),
        "vt_radiosity_update": (i32, [vp, u32, u32]),
        "vt_radiosity_update_moved": (i32, [vp, u32, u32, C.POINTER(u32)]),
        "vt_baked_upload": (i32, [vp, vp]),
        "vt_baked_occlusion": (i32, [vp, vp, u32]),
        "vt_baked_jacobi": (i32, [vp, vp, u32]),
        "vt_baked_read": (i32, [vp, u32, vp, sz]),
        "vt_radiosity_state": (i32, [vp, vp, sz, vp]),
        "vt_node_neighbours": (i32, [vp, vp, u32, u32, f32, f32, vp, vp]),
        "vt_radiosity_tables": (i32, [vp, u32, vp, sz]),
        "vt_capsule_test": (i32, [vp, vp, u32, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class VTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("vt error %d: %s" % (code, msg))
        self.code = code


def build_bvh(pos, idx, leaf_size=4, bins=16):
    """vt_build_bvh: the native binned-SAH BVH2 in the reference layout."""
    L = load()
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
    idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    n = idx.size // 3
    cap = n + 2
    nodes = np.zeros(cap * 16, np.float32)
    order = np.zeros(max(n, 1), np.uint32)
    cnt = C.c_uint32(0)
    rc = L.vt_build_bvh(ptr(pos), pos.size // 3, ptr(idx), n, leaf_size, bins, ptr(nodes), cap, C.byref(cnt), ptr(order))
    if rc:
        raise VTError(rc, "vt_build_bvh")
    return nodes[: cnt.value * 16].copy(), order[:n].copy()


def weld_subdivide(pos, nrm, idx, max_edge=0.003, max_passes=24):
    """vt_weld_subdivide: prepareScene's weld + conforming subdivision of one static mesh. Returns (P, N) as
    float64 (n, 3) and I uint32 (t, 3)."""
    L = load()
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
    nrm = np.ascontiguousarray(nrm, dtype=np.float32).reshape(-1)
    idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    nv, nt = C.c_uint32(0), C.c_uint32(0)
    args = (ptr(pos), ptr(nrm), pos.size // 3, ptr(idx), idx.size // 3, float(max_edge), int(max_passes))
    rc = L.vt_weld_subdivide(*args, None, None, 0, None, 0, C.byref(nv), C.byref(nt))
    if rc not in (0, VT_ERR_SIZE):
        raise VTError(rc, "vt_weld_subdivide")
    P = np.zeros((nv.value, 3), np.float64)
    N = np.zeros((nv.value, 3), np.float64)
    I = np.zeros((nt.value, 3), np.uint32)
    rc = L.vt_weld_subdivide(*args, ptr(P), ptr(N), nv.value, ptr(I), nt.value, C.byref(nv), C.byref(nt))
    if rc:
        raise VTError(rc, "vt_weld_subdivide")
    return P, N, I


class Context:
    """One vt_ctx (one device, one stream)."""

    def __init__(self, device=0):
        self.L = load()
        h = C.c_void_p()
        info = CreateInfo(device, 0)
        rc = self.L.vt_create(C.byref(info), C.byref(h))
        if rc:
            raise VTError(rc, "vt_create failed (no HIP device?)")
        self.h = h
        self._keep = {}

    def close(self):
        if self.h:
            self.L.vt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc:
            raise VTError(rc, self.L.vt_last_error(self.h).decode())

    # ---------------------------------------------------------------- scene / cameras
    def upload_scene(self, sc, with_nodes=True):
        """Upload a scene dict; with sc["tri_body"] (body index per triangle, STATIC_BODY for the static scene)
        and sc["body_count"] the bodies get BVHs of their own and are placed by set_instances; sc["emit_body"]
        (optional) names the body the light rides on, whose frame then holds the emit records."""
        a = {k: np.ascontiguousarray(sc[k]) for k in ("tris", "verts", "meshes", "materials", "lights", "emit")}
        nb = int(sc.get("body_count", 0))
        tb = np.ascontiguousarray(sc["tri_body"], dtype=np.uint32) if nb else None
        nodes = np.ascontiguousarray(sc["nodes"]) if with_nodes and not nb else None
        uvs = np.ascontiguousarray(sc["uvs"], dtype=np.float32) if sc.get("uvs") is not None else None
        s = Scene(ptr(nodes), 0 if nodes is None else nodes.size // 16, ptr(a["tris"]), a["tris"].size // 16,
                  ptr(a["verts"]), a["verts"].size // 8, ptr(a["meshes"]), a["meshes"].size // 4,
                  ptr(a["materials"]), a["materials"].size // 16, ptr(a["lights"]), a["lights"].size // 4,
                  ptr(a["emit"]), a["emit"].size // 16, float(sc["emit_area"]), ptr(tb), nb,
                  int(sc.get("shadow_skip", 0)), 1 if sc.get("emit_body") is not None else 0,
                  int(sc.get("emit_body") or 0), ptr(uvs))
        self.check(self.L.vt_upload_scene(self.h, C.byref(s)))
        if sc.get("board_mask") is not None:
            self.set_board_texture(sc["board_mask"])

    def set_board_texture(self, mask):
        """vt_set_board_texture: the print mask, an (h, w) uint8 array (255 = paper), row 0 the image's top."""
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self.check(self.L.vt_set_board_texture(self.h, ptr(m), m.shape[1], m.shape[0]))

    def set_instances(self, poses):
        """vt_set_instances: poses[i] = None (hidden) or dict(q0, p0[, q1, p1]) with quaternions (x, y, z, w), and
        optionally occ_a, occ_b, occ_r, rho_eff and cavity = dict(origin, axis, s0[, r])."""
        arr = (Instance * max(1, len(poses)))()
        for i, P in enumerate(poses):
            if P is None:
                arr[i].q0[:] = [0, 0, 0, 1]
                arr[i].q1[:] = [0, 0, 0, 1]
                arr[i].hidden = 1
                continue
            arr[i].q0[:] = [float(x) for x in P["q0"]]
            arr[i].q1[:] = [float(x) for x in P.get("q1", P["q0"])]
            arr[i].p0[:] = [float(x) for x in P["p0"]]
            arr[i].p1[:] = [float(x) for x in P.get("p1", P["p0"])]
            arr[i].occ_a[:] = [float(x) for x in P.get("occ_a", (0, 0, 0))]
            arr[i].occ_b[:] = [float(x) for x in P.get("occ_b", (0, 0, 0))]
            arr[i].occ_r = float(P.get("occ_r", 0.0))
            arr[i].rho_eff = float(P.get("rho_eff", 0.0))
            cav = P.get("cavity")
            if cav is not None:
                arr[i].cavity = 1
                arr[i].cavity_origin[:] = [float(x) for x in cav["origin"]]
                arr[i].cavity_axis[:] = [float(x) for x in cav["axis"]]
                arr[i].cavity_s0 = float(cav["s0"])
                arr[i].cavity_r = float(cav.get("r", 0.0))
        self.check(self.L.vt_set_instances(self.h, arr, len(poses)))

    def set_camera(self, slot, cam):
        c = Camera()
        c.R[:] = [float(x) for x in cam["R"]]
        c.pos[:] = [float(x) for x in cam["pos"]]
        c.mode = int(cam.get("mode", 0))
        c.lens[:] = [float(x) for x in cam["lens"]]
        for k in ("width", "height", "spp", "seed", "cam_index", "light_samples", "max_depth", "flags"):
            setattr(c, k, int(cam.get(k, 0)))
        c.scale = float(cam.get("scale", 1.0))
        c.shutter[:] = [float(x) for x in cam.get("shutter", (0.0, 1.0))]
        c.skip_bodies = int(cam.get("skip_bodies", 0))
        c.refine = int(cam.get("refine", 0))
        c.refine_light_samples = int(cam.get("refine_light_samples", 0))
        self.check(self.L.vt_set_camera(self.h, slot, C.byref(c)))

    def set_light_profile(self, light, axis=None, theta=(), p=()):
        """vt_set_light_profile: L_e = radiance * p(theta from axis), linear between knots; no knots = Lambertian."""
        th = np.ascontiguousarray(theta, dtype=np.float32)
        pp = np.ascontiguousarray(p, dtype=np.float32)
        ax = np.ascontiguousarray(axis if axis is not None else (0, 0, 1), dtype=np.float32)
        assert th.size == pp.size
        self.check(self.L.vt_set_light_profile(self.h, light, ptr(ax), ptr(th), ptr(pp), th.size))

    def upload_profiles(self, sc):
        """The scene dict's profiles ({light: (axis, theta[], p[])}), if any, after upload_scene."""
        for light, (axis, th, p) in (sc.get("profiles") or {}).items():
            self.set_light_profile(light, axis, th, p)

    # ---------------------------------------------------------------- node radiosity cache
    def radiosity_build(self, gather_k=64, node_samples=4, nbr_k=16, smooth_radius=0.006, smooth_sigma=0.003):
        o = RadiosityOpts(gather_k, node_samples, nbr_k, smooth_radius, smooth_sigma)
        info = RadiosityInfo()
        self.check(self.L.vt_radiosity_build(self.h, C.byref(o), C.byref(info)))
        return {k: getattr(info, k) for k, _ in RadiosityInfo._fields_}

    def interior_bvh(self, enable=True):
        """vt_interior_bvh: the interior tree on (built once per node cache) or off; returns the static triangles
        the walks see."""
        n = C.c_uint32(0)
        self.check(self.L.vt_interior_bvh(self.h, 1 if enable else 0, C.byref(n)))
        return n.value

    def radiosity_update(self, direct=True, iters=0):
        self.check(self.L.vt_radiosity_update(self.h, 1 if direct else 0, iters))

    # ---------------------------------------------------------------- the baked node table (lite renderer)
    def baked_upload(self, table):
        """vt_baked_upload: dict(nodes, K, gather uint16 (nodes, K), node_a, node_n (nodes, 4) float32, hh
        (meshes, 4) float32)."""
        self._baked = {k: np.ascontiguousarray(table[k], dt) for k, dt in
                       (("gather", np.uint16), ("node_a", np.float32), ("node_n", np.float32), ("hh", np.float32))}
        b = self._baked
        t = BakedTable(int(table["nodes"]), int(table["K"]), ptr(b["gather"]), ptr(b["node_a"]), ptr(b["node_n"]),
                       ptr(b["hh"]), b["hh"].size // 4)
        self.check(self.L.vt_baked_upload(self.h, C.byref(t)))
        self._baked_shape = (int(table["nodes"]), int(table["K"]))

    def baked_occlusion(self, caps):
        cp = np.ascontiguousarray(np.asarray(caps, np.float32).reshape(-1, 8))
        self.check(self.L.vt_baked_occlusion(self.h, ptr(cp), cp.shape[0]))

    def baked_jacobi(self, e_dir=None, iters=1):
        ed = None if e_dir is None else np.ascontiguousarray(e_dir, np.float32)
        self.check(self.L.vt_baked_jacobi(self.h, ptr(ed), iters))

    def baked_read(self, which):
        """0: B (nodes, 4) float32; 1: E_ind (nodes, 4); 2: transmission bytes (nodes, K) uint8."""
        n, K = self._baked_shape
        out = np.zeros((n, K), np.uint8) if which == 2 else np.zeros((n, 4), np.float32)
        self.check(self.L.vt_baked_read(self.h, which, ptr(out), out.nbytes))
        return out

    DIRECT_KINDS = ("none", "incremental", "full")

    def radiosity_update_moved(self, moved, iters=0):
        """The incremental update (_ensureRadiosity with a moved-body mask); returns the direct pass it ran:
        'none', 'incremental' or 'full'."""
        ran = C.c_uint32(0)
        self.check(self.L.vt_radiosity_update_moved(self.h, int(moved) & 0xffffffff, iters, C.byref(ran)))
        return self.DIRECT_KINDS[ran.value]

    def capsule_test(self, cases):
        cases = np.ascontiguousarray(cases, np.float32).reshape(-1, 16)
        out = np.zeros(max(1, cases.shape[0]), np.float32)
        self.check(self.L.vt_capsule_test(self.h, ptr(cases), cases.shape[0], ptr(out)))
        return out[: cases.shape[0]]

    def radiosity_tables(self, which, shape, dtype):
        out = np.zeros(shape, dtype)
        self.check(self.L.vt_radiosity_tables(self.h, which, ptr(out), out.nbytes))
        return out

    def radiosity_state(self, n_nodes=None):
        """(node state n x 16 or None, eAvg[3]); n_nodes = the scene's vertex count to download the state."""
        e = (C.c_float * 3)()
        st = np.zeros((n_nodes, 16), np.float32) if n_nodes else None
        self.check(self.L.vt_radiosity_state(self.h, ptr(st), 0 if st is None else st.nbytes, e))
        return st, np.array(e[:])

    def visible(self, slot, points, tol=2e-4, time=0.5):
        """vt_visible: bool per world point, seen unoccluded from the camera centre (to tol short of the point)."""
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.zeros(max(1, pts.shape[0]), np.uint8)
        self.check(self.L.vt_visible(self.h, slot, ptr(pts), pts.shape[0], float(tol), float(time), ptr(out)))
        return out[: pts.shape[0]].astype(bool)

    def render(self, slots, mode=VT_RENDER_BATCHED, spp_begin=0, spp_end=0, accumulate=False, chunk=0, block=0,
               stack_lds=0, queues=0, graph=False, batch=0, sort_bits=0, shadow_sort=0):
        arr = np.ascontiguousarray(slots, dtype=np.uint32)
        o = RenderOpts(mode, spp_begin, spp_end, 1 if accumulate else 0, chunk, block, stack_lds, queues,
                       1 if graph else 0, batch, sort_bits, shadow_sort)
        self.check(self.L.vt_render(self.h, ptr(arr), arr.size, C.byref(o)))

    def copy_bandwidth(self, nbytes, iters=50):
        g = C.c_double(0)
        self.check(self.L.vt_copy_bandwidth(self.h, nbytes, iters, C.byref(g)))
        return g.value

    def read_bandwidth(self, nbytes, iters=20):
        g = C.c_double(0)
        self.check(self.L.vt_read_bandwidth(self.h, nbytes, iters, C.byref(g)))
        return g.value

    def autotune(self, slots, spp_begin=0, spp_end=0, probe_stride=0):
        arr = np.ascontiguousarray(slots, dtype=np.uint32)
        o = RenderOpts(VT_RENDER_BATCHED, spp_begin, spp_end, 0, 0, 0, 0, 0, 0, 0, 0)
        r = TuneResult()
        self.check(self.L.vt_autotune(self.h, ptr(arr), arr.size, C.byref(o), probe_stride, C.byref(r)))
        return {k: getattr(r, k) for k, _ in TuneResult._fields_}

    def sensor(self, slot, params):
        p = sensor_struct(params)
        self.check(self.L.vt_sensor(self.h, slot, C.byref(p)))

    def isp(self, slot, params):
        q = isp_struct(params)
        self.check(self.L.vt_isp(self.h, slot, C.byref(q)))

    def readback(self, slot, kind, w, h):
        cw, ch = (w + 1) // 2, (h + 1) // 2
        if kind == VT_KIND_LINEAR:
            out = np.zeros((h, w, 4), np.float32)
        elif kind == VT_KIND_RAW:
            out = np.zeros((h, w), np.uint16)
        elif kind == VT_KIND_RGB8:
            out = np.zeros((h, w, 3), np.uint8)
        elif kind == VT_KIND_RGBA8:
            out = np.zeros((h, w, 4), np.uint8)
        elif kind == VT_KIND_YUV420:
            out = np.zeros(w * h + 2 * cw * ch, np.uint8)
        elif kind == VT_KIND_STATS:
            out = np.zeros(4, np.uint32)
        else:
            out = np.zeros((h, w), np.float32)
        self.check(self.L.vt_readback(self.h, slot, kind, ptr(out), out.nbytes))
        if kind == VT_KIND_YUV420:
            return dict(y=out[:w * h].reshape(h, w), cb=out[w * h:w * h + cw * ch].reshape(ch, cw),
                        cr=out[w * h + cw * ch:].reshape(ch, cw))
        return out

    def sync(self):
        self.check(self.L.vt_sync(self.h))

    def stats(self):
        s = FrameStats()
        self.check(self.L.vt_frame_stats_get(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in FrameStats._fields_}

    def update_materials(self, materials):
        m = np.ascontiguousarray(materials, dtype=np.float32).reshape(-1)
        self.check(self.L.vt_update_materials(self.h, ptr(m), m.size // 16))

    def update_lights(self, lights):
        a = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1)
        self.check(self.L.vt_update_lights(self.h, ptr(a), a.size // 4))

    def backward(self, slots, adj_images, n_mat, n_light, spp_begin=0, spp_end=0, chunk=0, cameras=False):
        """Gradients of sum adj . I w.r.t. material colours (n_mat x 3) and light radiances (n_light x 3); with
        cameras=True also w.r.t. each listed camera's 14 parameters (fx fy cx cy k1..k4, rotation, centre)."""
        arr = np.ascontiguousarray(slots, dtype=np.uint32)
        imgs = [np.ascontiguousarray(a, dtype=np.float32) for a in adj_images]
        ptrs = (C.c_void_p * len(imgs))(*[im.ctypes.data for im in imgs])
        gm = np.zeros((n_mat, 3), np.float32)
        gl = np.zeros((max(n_light, 1), 3), np.float32)
        gc = np.zeros((arr.size, CAMERA_PARAMS), np.float32) if cameras else None
        g = Grads(ptr(gm), n_mat, ptr(gl), n_light, ptr(gc), arr.size if cameras else 0)
        o = RenderOpts(0, spp_begin, spp_end, 0, chunk, 0, 0, 0, 0, 0)
        self.check(self.L.vt_backward(self.h, ptr(arr), arr.size, ptrs, C.byref(o), C.byref(g)))
        if cameras:
            return gm, gl[:n_light], gc
        return gm, gl[:n_light]

    # ---------------------------------------------------------------- test entry points
    def trace(self, rays, any_hit=False):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.zeros((rays.shape[0], 4), np.float32)
        self.check(self.L.vt_trace_test(self.h, ptr(rays), rays.shape[0], int(any_hit), ptr(out)))
        return out

    def lens_test(self, lens9, uv):
        lens9 = np.ascontiguousarray(lens9, dtype=np.float32)
        uv = np.ascontiguousarray(uv, dtype=np.float32).reshape(-1, 2)
        out = np.zeros((uv.shape[0], 3), np.float32)
        self.check(self.L.vt_lens_test(self.h, ptr(lens9), ptr(uv), uv.shape[0], ptr(out)))
        return out

    def camera_rays(self, slot, xyi):
        xyi = np.ascontiguousarray(xyi, dtype=np.uint32).reshape(-1, 3)
        out = np.zeros((xyi.shape[0], 6), np.float32)
        self.check(self.L.vt_camera_ray_test(self.h, slot, ptr(xyi), xyi.shape[0], ptr(out)))
        return out

    def bsdf_test(self, cases):
        cases = np.ascontiguousarray(cases, dtype=np.float32).reshape(-1, 16)
        out = np.zeros((cases.shape[0], 8), np.float32)
        self.check(self.L.vt_bsdf_test(self.h, ptr(cases), cases.shape[0], ptr(out)))
        return out

    def sensor_raw_device(self, params, width, height, radiance_ptr, ri_ptr, dn_ptr):
        """vt_sensor_raw_device: the raw stage on device buffers given by address (radiance w*h float4, ri w*h
        float32, dn w*h uint16), enqueued on the context's stream; allocates nothing."""
        p = sensor_struct(params)
        self.check(self.L.vt_sensor_raw_device(self.h, C.byref(p), width, height, C.c_void_p(radiance_ptr),
                                               C.c_void_p(ri_ptr), C.c_void_p(dn_ptr)))

    def sensor_test(self, params, radiance, ri):
        radiance = np.ascontiguousarray(radiance, dtype=np.float32)
        ri = np.ascontiguousarray(ri, dtype=np.float32)
        h, w = ri.shape
        dn = np.zeros((h, w), np.uint16)
        p = sensor_struct(params)
        self.check(self.L.vt_sensor_test(self.h, C.byref(p), w, h, ptr(radiance), ptr(ri), ptr(dn)))
        return dn

    def isp_test(self, sensor, params, dn):
        """the ISP on caller-given raw DN (h, w): dict(rgba, y/cb/cr (4:2:0), stats)"""
        dn = np.ascontiguousarray(dn, dtype=np.uint16)
        h, w = dn.shape
        cw, ch = (w + 1) // 2, (h + 1) // 2
        sp = sensor_struct(dict(dict(frame=0, seed=0, optics_k=1.0, esf_sigma=0.0, exposure=1.0, m=np.eye(3),
                                     full_well=1e9, dark=0.0, noise_on=0), **sensor))
        q = isp_struct(params)
        rgba = np.zeros((h, w, 4), np.uint8)
        yuv = np.zeros(w * h + 2 * cw * ch, np.uint8)
        stats = np.zeros(4, np.uint32)
        self.check(self.L.vt_isp_test(self.h, C.byref(sp), C.byref(q), w, h, ptr(dn), ptr(rgba), ptr(yuv),
                                      ptr(stats)))
        out = dict(rgba=rgba, stats=stats)
        if q.chroma420:
            out.update(y=yuv[:w * h].reshape(h, w), cb=yuv[w * h:w * h + cw * ch].reshape(ch, cw),
                       cr=yuv[w * h + cw * ch:].reshape(ch, cw))
        return out

    # ---------------------------------------------------------------- radiosity
    def jacobi_sparse(self, t, node_state, iters):
        keep = {k: np.ascontiguousarray(v) for k, v in t.items() if isinstance(v, np.ndarray)}
        g = GatherTables(t["n_static"], t["n_active"], t["gather_k"], t["nbr_k"], t["n_mat"], t["albedo_bins"],
                         ptr(keep["active"]), ptr(keep["gtable"]), ptr(keep["normals"]), ptr(keep["node_mat"]),
                         ptr(keep["albedo"]), ptr(keep["hh"]), ptr(keep.get("nbr")), ptr(keep.get("nbr_w")))
        st = np.ascontiguousarray(node_state, dtype=np.float32).copy()
        aux = np.zeros((t["n_active"], 4), np.float32)
        self.check(self.L.vt_jacobi_sparse(self.h, C.byref(g), ptr(st), iters, ptr(aux)))
        return st, aux

    def dense_upload(self, F, e):
        F = np.ascontiguousarray(F, dtype=np.float32)
        e = np.ascontiguousarray(e, dtype=np.float32)
        n, r = e.shape
        self.check(self.L.vt_dense_upload(self.h, n, r, ptr(F), ptr(e)))

    def dense_formfactors(self, nodes, row_max, e):
        nodes = np.ascontiguousarray(nodes, dtype=np.float32).reshape(-1, 8)
        e = np.ascontiguousarray(e, dtype=np.float32)
        self.check(self.L.vt_dense_formfactors(self.h, nodes.shape[0], ptr(nodes), float(row_max), e.shape[1], ptr(e)))

    def dense_solve(self, n, r, method, tol=1e-6, max_sweeps=2000, inner_sweeps=0, want_x=True):
        o = DenseOpts(n, r, method, tol, max_sweeps, inner_sweeps)
        res = DenseResult()
        x = np.zeros((n, r), np.float32) if want_x else None
        self.check(self.L.vt_dense_solve(self.h, C.byref(o), ptr(x), C.byref(res)))
        return x, {k: getattr(res, k) for k, _ in DenseResult._fields_}

    def dense_F(self, n):
        F = np.zeros((n, n), np.float32)
        self.check(self.L.vt_dense_download_F(self.h, ptr(F), F.nbytes))
        return F


def sensor_struct(p):
    s = SensorParams()
    s.cfa, s.frame, s.seed = int(p["cfa"]), int(p["frame"]), int(p["seed"])
    s.optics_k, s.esf_sigma, s.exposure, s.gain = p["optics_k"], p["esf_sigma"], p["exposure"], p["gain"]
    s.m[:] = [float(x) for x in np.asarray(p["m"]).reshape(-1)]
    s.read_noise, s.full_well, s.dark, s.conv_gain = p["read_noise"], p["full_well"], p["dark"], p["conv_gain"]
    s.max_dn, s.black = p["max_dn"], p["black"]
    s.noise_on = 1 if p["noise_on"] else 0
    s.pupil_exp = float(p.get("pupil_exp", 1.0))
    s.glare_fraction = float(p.get("glare_fraction", 0.0))
    s.glare_sigma = float(p.get("glare_sigma", 1.0))
    s.glare_level = int(p.get("glare_level", 16))
    vt, vv = list(p.get("vignetting_theta_deg", ())), list(p.get("vignetting", ()))
    assert len(vt) == len(vv) <= 8
    s.vignetting_count = len(vt)
    s.vignetting_theta_deg[:len(vt)] = [float(x) for x in vt]
    s.vignetting[:len(vv)] = [float(x) for x in vv]
    return s


def node_neighbours(pos, nrm, k=16, radius=0.006, sigma=0.003):
    """vt_node_neighbours (host only): the smoothing neighbour rows of the node radiosity build."""
    L = load()
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    idx = np.zeros((max(n, 1), k), np.uint32)
    w = np.zeros((max(n, 1), k), np.float32)
    rc = L.vt_node_neighbours(ptr(pos), ptr(nrm), n, k, float(radius), float(sigma), ptr(idx), ptr(w))
    if rc:
        raise VTError(rc, "vt_node_neighbours")
    return idx[:n], w[:n]
