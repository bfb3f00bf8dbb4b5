"""Seeded synthetic workloads of the camera path, built in the reference's upload format.

Every number comes from mulberry32 streams (control/rng.js:4-29), drawn in a fixed order, so a seed fixes the
scene. Seeds: scene 1001, materials 1002, lights 1003, cameras 1004; camera c's pixel seed is 1005 + c.

  c1()  the CPU-runnable case: hemisphere r = 3 m (19,900 triangles) + 1 x 2 x 0.8 m table box + 4 quad lights
        (radiance U[5, 20]); 16 Lambertian materials, albedo U[0.2, 0.8]; one KB4 camera 640x480, fx = fy = 300,
        k1..k4 ~ U[-0.05, 0.05]; 16 spp, depth 3.
  c2()  full scale: the dome (959,000 triangles) + 100 cylinders + 100 boxes (2,013,432 triangles), 16 quad
        lights, 256 materials (Lambert, GGX dielectric, GGX conductor; roughness U[0.05, 0.8]); three KB4 cameras
        1920x1080, 64 spp, depth 5, NEE; RGGB sensor with Poisson shot and 2 DN Gaussian read noise, 12 bit.
  c3()  the c2 scene with 24 cameras of uneven size (320x240 .. 2048x1536) and spp from {1, 4, 16, 64} on a
        2.5 m ring.
  c4()  node radiosity: 32,768 dome triangles as nodes, 3 channels x 16 emitter configurations.
  c5()  the c2 scene, three cameras 1280x960 at 32 spp.
"""
import math

import numpy as np

MULB = np.uint32(0x6D2B79F5)

KIND_DIELECTRIC, KIND_CONDUCTOR, KIND_COATED, KIND_LAMBERT = 0, 1, 2, 3
MESH_STATIC, MESH_EMITTER, MESH_TEXTURED = 1, 2, 128


class Rng:
    """control/rng.js Rng over a vectorised mulberry32: draw k of a stream has state seed + k * 0x6d2b79f5."""

    def __init__(self, seed):
        self.seed = np.uint32(seed & 0xFFFFFFFF)
        self.k = 0

    def u(self, n=None):
        m = 1 if n is None else int(n)
        with np.errstate(over="ignore"):
            a = (self.seed + (np.arange(self.k + 1, self.k + m + 1, dtype=np.uint64) * 0x6D2B79F5)).astype(np.uint64)
            a = (a & 0xFFFFFFFF).astype(np.uint32)
            t = a
            t = (t ^ (t >> np.uint32(15))) * (t | np.uint32(1))
            t = t ^ (t + (t ^ (t >> np.uint32(7))) * (t | np.uint32(61)))
            out = (t ^ (t >> np.uint32(14))).astype(np.float64) / 4294967296.0
        self.k += m
        return out[0] if n is None else out

    def uniform(self, lo=0.0, hi=1.0, n=None):
        return lo + (hi - lo) * self.u(n)

    def int(self, lo, hi, n=None):
        return (lo + np.floor(self.u(n) * (hi - lo + 1))).astype(np.int64) if n is not None else lo + int(math.floor(self.u() * (hi - lo + 1)))

    def pick(self, arr):
        return arr[int(math.floor(self.u() * len(arr)))]


# ---------------------------------------------------------------------------- geometry builders

class Builder:
    def __init__(self):
        self.P, self.N, self.I, self.M = [], [], [], []
        self.nv = 0

    def add(self, P, N, I, mesh):
        P = np.asarray(P, np.float64).reshape(-1, 3)
        N = np.asarray(N, np.float64).reshape(-1, 3)
        I = np.asarray(I, np.int64).reshape(-1, 3)
        self.P.append(P)
        self.N.append(N)
        self.I.append(I + self.nv)
        self.M.append(np.broadcast_to(np.asarray(mesh, np.int64), (I.shape[0],)).copy())
        self.nv += P.shape[0]

    def arrays(self):
        return (np.concatenate(self.P).astype(np.float32), np.concatenate(self.N).astype(np.float32),
                np.concatenate(self.I).astype(np.uint32), np.concatenate(self.M).astype(np.uint32))


def dome(r, n_theta, n_phi, pole=True, theta0=0.0):
    """Hemisphere z >= 0 of radius r, inward normals. pole=True: a pole vertex and a triangle fan at the top
    (n_phi + 2 n_phi (n_theta - 1) triangles); else a quad grid from theta0 (2 n_phi n_theta triangles)."""
    if pole:
        th = np.linspace(0.0, math.pi / 2, n_theta + 1)[1:]
    else:
        th = np.linspace(theta0, math.pi / 2, n_theta + 1)
    ph = np.arange(n_phi) * (2 * math.pi / n_phi)
    T, Ph = np.meshgrid(th, ph, indexing="ij")
    P = np.stack([r * np.sin(T) * np.cos(Ph), r * np.sin(T) * np.sin(Ph), r * np.cos(T)], -1).reshape(-1, 3)
    rows = th.size
    idx = lambda i, j: i * n_phi + (j % n_phi)
    I = []
    if pole:
        P = np.vstack([[0.0, 0.0, r], P])
        off = 1
        j = np.arange(n_phi)
        I.append(np.stack([np.zeros(n_phi, np.int64), off + idx(0, j + 1), off + idx(0, j)], -1))
    else:
        off = 0
    i, j = np.meshgrid(np.arange(rows - 1), np.arange(n_phi), indexing="ij")
    i, j = i.ravel(), j.ravel()
    a, b, c, d = off + idx(i, j), off + idx(i, j + 1), off + idx(i + 1, j), off + idx(i + 1, j + 1)
    I.append(np.stack([a, b, d], -1))
    I.append(np.stack([a, d, c], -1))
    I = np.concatenate(I)
    N = -P / np.linalg.norm(P, axis=1, keepdims=True)
    return P, N, I


def box(lo, hi, g=1, R=None, t=None):
    """Axis box [lo, hi] with each face a g x g grid, flat outward normals; optional rotation R and offset t."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    Ps, Ns, Is = [], [], []
    nv = 0
    s = np.linspace(0.0, 1.0, g + 1)
    U, V = np.meshgrid(s, s, indexing="ij")
    U, V = U.ravel(), V.ravel()
    for ax in range(3):
        for side in (0, 1):
            a1, a2 = (ax + 1) % 3, (ax + 2) % 3
            p = np.zeros((U.size, 3))
            p[:, ax] = hi[ax] if side else lo[ax]
            p[:, a1] = lo[a1] + U * (hi[a1] - lo[a1])
            p[:, a2] = lo[a2] + V * (hi[a2] - lo[a2])
            n = np.zeros(3)
            n[ax] = 1.0 if side else -1.0
            i, j = np.meshgrid(np.arange(g), np.arange(g), indexing="ij")
            i, j = i.ravel(), j.ravel()
            q = lambda ii, jj: ii * (g + 1) + jj
            A, B, C, D = q(i, j), q(i + 1, j), q(i, j + 1), q(i + 1, j + 1)
            tri = np.concatenate([np.stack([A, B, D], -1), np.stack([A, D, C], -1)])
            if side == 0:
                tri = tri[:, [0, 2, 1]]
            Ps.append(p)
            Ns.append(np.broadcast_to(n, p.shape))
            Is.append(tri + nv)
            nv += p.shape[0]
    P, N, I = np.concatenate(Ps), np.concatenate(Ns), np.concatenate(Is)
    if R is not None:
        P = P @ np.asarray(R).T
        N = N @ np.asarray(R).T
    if t is not None:
        P = P + np.asarray(t)
    return P, N, I


def cylinder(radius, height, nseg, nring, R=None, t=None):
    """Closed cylinder along +z from 0 to height, radial normals on the side, flat caps."""
    ph = np.arange(nseg) * (2 * math.pi / nseg)
    z = np.linspace(0.0, height, nring + 1)
    Z, Ph = np.meshgrid(z, ph, indexing="ij")
    side = np.stack([radius * np.cos(Ph), radius * np.sin(Ph), Z], -1).reshape(-1, 3)
    sn = np.stack([np.cos(Ph), np.sin(Ph), np.zeros_like(Ph)], -1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nring), np.arange(nseg), indexing="ij")
    i, j = i.ravel(), j.ravel()
    q = lambda ii, jj: ii * nseg + (jj % nseg)
    A, B, C, D = q(i, j), q(i, j + 1), q(i + 1, j), q(i + 1, j + 1)
    I = [np.stack([A, B, D], -1), np.stack([A, D, C], -1)]
    P, N = [side], [sn]
    nv = side.shape[0]
    for zc, nz in ((0.0, -1.0), (height, 1.0)):
        rim = np.stack([radius * np.cos(ph), radius * np.sin(ph), np.full(nseg, zc)], -1)
        P.append(np.vstack([[0.0, 0.0, zc], rim]))
        N.append(np.broadcast_to([0.0, 0.0, nz], (nseg + 1, 3)))
        jj = np.arange(nseg)
        fan = np.stack([np.zeros(nseg, np.int64), 1 + jj, 1 + (jj + 1) % nseg], -1)
        if nz > 0:
            fan = fan[:, [0, 2, 1]]
        I.append(fan + nv)
        nv += nseg + 1
    P, N, I = np.concatenate(P), np.concatenate(N), np.concatenate(I)
    if R is not None:
        P = P @ np.asarray(R).T
        N = N @ np.asarray(R).T
    if t is not None:
        P = P + np.asarray(t)
    return P, N, I


def quad(center, size, normal, up_hint=(0.0, 1.0, 0.0)):
    n = np.asarray(normal, float)
    n = n / np.linalg.norm(n)
    u = np.cross(up_hint, n)
    if np.linalg.norm(u) < 1e-6:
        u = np.cross((1.0, 0.0, 0.0), n)
    u = u / np.linalg.norm(u)
    v = np.cross(n, u)
    c = np.asarray(center, float)
    h = size / 2
    P = np.array([c - h * u - h * v, c + h * u - h * v, c + h * u + h * v, c - h * u + h * v])
    # winding with geometric normal along n: (b - a) x (c - a) = (2h u) x (2h u + 2h v) ~ u x v = n
    I = np.array([[0, 1, 2], [0, 2, 3]])
    return P, np.broadcast_to(n, (4, 3)), I


def rot_axis(axis, ang):
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def look_at(pos, target, up=(0.0, 0.0, 1.0)):
    """OpenCV camera axes (x right, y down, z forward) as the columns of R_D_C, flattened column-major."""
    f = np.asarray(target, float) - np.asarray(pos, float)
    f /= np.linalg.norm(f)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    return np.concatenate([r, d, f]).astype(np.float32)


# ---------------------------------------------------------------------------- materials, lights, packing

def material_record(kind, c, roughness=0.5, ior=1.5, tau=(1.0, 1.0, 1.0), coat_roughness=0.05):
    """16 floats: [c.rgb, alpha][ior, kind, fbar, roughness][tau.rgb, alpha_c][fbar_computed, -, -, -]
    (gpu/bsdf.js:68-88 prepareMaterial: roughness clamped to >= 0.02, alpha = roughness^2)."""
    r = max(float(roughness), 0.02)
    rec = np.zeros(16, np.float32)
    rec[0:3] = c
    rec[3] = r * r
    rec[4] = ior
    rec[5] = kind
    rec[7] = r
    rec[8:11] = tau
    rc = max(coat_roughness, 0.02)
    rec[11] = rc * rc
    if kind == KIND_DIELECTRIC:
        fb = fresnel_average(ior)
        rec[6] = fb
        rec[12] = fb
    return rec


def fresnel_average(eta, n=4096):
    mu = np.arange(n + 1) / n
    c = np.clip(mu, 0, 1)
    s2 = (1 - c * c) / (eta * eta)
    ct = np.sqrt(np.maximum(0, 1 - s2))
    rs = (c - eta * ct) / (c + eta * ct)
    rp = (eta * c - ct) / (eta * c + ct)
    F = np.where(s2 >= 1, 1.0, 0.5 * (rs * rs + rp * rp))
    w = np.where((np.arange(n + 1) == 0) | (np.arange(n + 1) == n), 1, np.where(np.arange(n + 1) % 2 == 1, 4, 2))
    return float(2 * np.sum(w * F * mu) / (3 * n))


def pack_scene(P, N, I, tri_mesh, meshes, materials, lights, bvh="vt", tri_body=None, body_count=0):
    """Upload arrays (reference layout) with the BVH built natively (bvh='vt') or by the oracle (bvh=callable).
    With posed bodies (tri_body: body per triangle, STATIC_BODY for the static scene) the triangles keep their
    order and no BVH is packed: every group's BVH is built at upload."""
    from . import lib

    if tri_body is not None:
        nodes, order = None, np.arange(I.shape[0])
    elif bvh == "vt":
        nodes, order = lib.build_bvh(P, I)
    else:
        nodes, order = bvh(P, I)
    T = I.shape[0]
    ordI = I[order]
    tris = np.zeros((T, 16), np.float32)
    v0 = P[ordI[:, 0]].astype(np.float64)
    tris[:, 0:3] = P[ordI[:, 0]]
    tris[:, 4:7] = (P[ordI[:, 1]].astype(np.float64) - v0).astype(np.float32)
    tris[:, 8:11] = (P[ordI[:, 2]].astype(np.float64) - v0).astype(np.float32)
    tu = tris.view(np.uint32)
    tu[:, 3], tu[:, 7], tu[:, 11] = ordI[:, 0], ordI[:, 1], ordI[:, 2]
    tu[:, 12] = tri_mesh[order]
    tu[:, 13] = np.asarray(meshes, np.uint32).reshape(-1, 4)[tri_mesh[order], 1]
    verts = np.zeros((P.shape[0], 8), np.float32)
    verts[:, 0:3] = P
    verts[:, 4:7] = N
    # word 3: the vertex's mesh (that of the first triangle using it), as the reference's prepared vertices
    vm = np.zeros(P.shape[0], np.uint32)
    flat = I.reshape(-1)
    tm = np.repeat(np.asarray(tri_mesh, np.uint32), 3)
    _, idx = np.unique(flat, return_index=True)
    vm[flat[idx]] = tm[idx]
    verts.view(np.uint32)[:, 3] = vm
    meshes = np.asarray(meshes, np.uint32).reshape(-1, 4)
    # emitter triangles: every triangle of an emitter mesh, outward normal from the vertex normals
    em_mask = (meshes[tri_mesh, 1] & MESH_EMITTER) != 0
    eI = I[em_mask]
    eM = tri_mesh[em_mask]
    emit = np.zeros((eI.shape[0], 16), np.float32)
    if eI.shape[0]:
        a, b, c = P[eI[:, 0]].astype(np.float64), P[eI[:, 1]].astype(np.float64), P[eI[:, 2]].astype(np.float64)
        e1, e2 = b - a, c - a
        n = np.cross(e1, e2)
        area = 0.5 * np.linalg.norm(n, axis=1)
        vn = N[eI[:, 0]].astype(np.float64) + N[eI[:, 1]] + N[eI[:, 2]]
        flip = np.sum(n * vn, axis=1) < 0
        n[flip] *= -1
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        cdf = np.cumsum(area) / area.sum()
        emit[:, 0:3] = a
        emit[:, 3] = cdf
        emit[:, 4:7] = e1
        emit[:, 7] = area
        emit[:, 8:11] = e2
        emit.view(np.uint32)[:, 11] = meshes[eM, 2]
        emit[:, 12:15] = n
        emit_area = float(area.sum())
    else:
        emit_area = 0.0
    out = dict(nodes=nodes, tris=tris.reshape(-1), verts=verts.reshape(-1), meshes=meshes.reshape(-1),
               materials=np.asarray(materials, np.float32).reshape(-1), lights=np.asarray(lights, np.float32).reshape(-1),
               emit=emit.reshape(-1), emit_area=emit_area, order=order, pos=P, idx=I, tri_mesh=tri_mesh)
    if tri_body is not None:
        out.update(tri_body=np.asarray(tri_body, np.uint32), body_count=int(body_count))
    return out


def kb4_lens(rng, w, h, f):
    """fx = fy = f, centre at the middle pixel, k1..k4 ~ U[-0.05, 0.05], redrawn until the KB4 map stays monotone
    over the image's field (gpu/lens.js:65-83)."""
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    for _ in range(1000):
        k = [rng.uniform(-0.05, 0.05) for _ in range(4)]
        if _monotone_over_image(k, f, cx, cy, w, h):
            return [f, f, cx, cy, k[0], k[1], k[2], k[3], 0.0]
    raise RuntimeError("no monotone KB4 draw")


def _monotone_over_image(k, f, cx, cy, w, h):
    th = np.arange(1, 4097) / 4096 * math.pi
    t2 = th * th
    dd = 1 + t2 * (3 * k[0] + t2 * (5 * k[1] + t2 * (7 * k[2] + t2 * 9 * k[3])))
    bad = np.nonzero(dd <= 0)[0]
    th_mono = math.pi if bad.size == 0 else bad[0] / 4096 * math.pi
    rmax = math.hypot(max(cx + 0.5, w - 0.5 - cx), max(cy + 0.5, h - 0.5 - cy)) / f
    t = np.linspace(0, th_mono, 2048)
    t2 = t * t
    thd = t * (1 + t2 * (k[0] + t2 * (k[1] + t2 * (k[2] + t2 * k[3]))))
    return thd.max() > rmax * 1.0001


def camera(R, pos, lens, w, h, spp, cam_index, depth, light_samples=1, seed=0, mode=0):
    return dict(R=R, pos=np.asarray(pos, np.float32), lens=np.asarray(lens, np.float32), width=w, height=h, spp=spp,
                seed=seed, cam_index=cam_index, light_samples=light_samples, max_depth=depth, mode=mode, scale=1.0)


def window(cam, x0, y0, w, h):
    """The w x h window of a camera's image at (x0, y0) as a camera of its own: same pose and KB4 lens with the
    principal point moved by (-x0, -y0), so its pixel (x, y) sees the rays of the parent's pixel (x0 + x, y0 + y).
    Its sample pattern is its own (the pixel hash takes the window's coordinates)."""
    c = dict(cam)
    lens = np.asarray(cam["lens"], np.float32).copy()
    lens[2] -= np.float32(x0)
    lens[3] -= np.float32(y0)
    c.update(lens=lens, width=int(w), height=int(h))
    return c


def default_sensor(cfa=0, read_noise_dn=2.0, exposure=2e-3, gain=1.0, seed=1006, frame=0, noise=True):
    """gpu/sensor.js:17-84 defaults (pitch 3 um, QE [0.45, 0.55, 0.5] at 600/540/460 nm, f/2, T 0.9, ESF 0.7 px,
    full well 10,000 e-, conversion gain 0.33 DN/e-, 12 bit), read noise set to `read_noise_dn` DN."""
    Hc, Cc = 6.62607015e-34, 299792458.0
    qe, nm = (0.45, 0.55, 0.5), (600, 540, 460)
    g = [3e-6 * 3e-6 * q * (l * 1e-9) / (Hc * Cc) for q, l in zip(qe, nm)]
    cg = 0.33
    return dict(cfa=cfa, frame=frame, seed=seed, optics_k=math.pi * 0.9 / (4 * 2.0 * 2.0),
                esf_sigma=math.sqrt(0.7 * 0.7 - 1.0 / 12.0), exposure=exposure, gain=gain,
                m=np.array([[g[0], 0, 0], [0, g[1], 0], [0, 0, g[2]]], np.float32),
                read_noise=read_noise_dn / (cg * gain), full_well=10000.0, dark=0.0, conv_gain=cg,
                max_dn=4095.0, black=0.0, noise_on=noise, cam_index=0, pupil_exp=1.0)


# ---------------------------------------------------------------------------- the configurations

STATIC_BODY = 0xFFFFFFFF


def quat_axis(axis, ang):
    """Unit quaternion (x, y, z, w) of the rotation by ang about axis."""
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    return np.concatenate([a * math.sin(ang / 2), [math.cos(ang / 2)]])


def posed_parts():
    """The c1 room (dome, a fixed box, four ceiling emitters) and two bodies in their own frames: a box
    (0.5 x 0.3 x 0.4 m, body 0) and a cylinder (r 0.15 m, 0.6 m tall, body 1), 16 Lambertian materials."""
    rs, rm, rl = Rng(2001), Rng(2002), Rng(2003)
    parts = []
    P, N, I = dome(3.0, 40, 48, pole=True)
    parts.append((P, N, I, rs.int(0, 15, n=I.shape[0]), None))
    P, N, I = box((-0.5, -1.0, 0.0), (0.5, 1.0, 0.6))
    parts.append((P, N, I, rs.int(0, 15, n=I.shape[0]), None))
    P, N, I = box((-0.25, -0.15, -0.2), (0.25, 0.15, 0.2), g=2)
    parts.append((P, N, I, rs.int(0, 15, n=I.shape[0]), 0))
    P, N, I = cylinder(0.15, 0.6, 24, 3)
    parts.append((P, N, I, rs.int(0, 15, n=I.shape[0]), 1))
    mats = [material_record(KIND_LAMBERT, rm.uniform(0.2, 0.8, n=3)) for _ in range(16)]
    mats.append(material_record(KIND_LAMBERT, (0.0, 0.0, 0.0)))
    meshes = [[m, MESH_STATIC, 0, 0] for m in range(16)]
    lights = []
    for k, (x, y) in enumerate(((1.0, 1.0), (-1.0, 1.0), (-1.0, -1.0), (1.0, -1.0))):
        L = rl.uniform(5.0, 20.0)
        lights.append([L, L, L, 0.0])
        meshes.append([16, MESH_EMITTER, k, 0])
        P, N, I = quad((x, y, 2.4), 0.6, (0.0, 0.0, -1.0))
        parts.append((P, N, I, np.full(I.shape[0], 16 + k), None))
    return parts, meshes, mats, lights


# the light of posed_light_parts: its angular profile about the body's +x axis (axis, theta knots, p knots)
BODY_LIGHT_PROFILE = ((1.0, 0.0, 0.0), (0.0, 0.6, 1.2, 1.5), (1.0, 0.8, 0.3, 0.0))


def posed_light_parts():
    """posed_parts with the light riding on body 0 (as the reference's outlet rides on the scope): a 0.2 m emitting
    square 2 mm off the box's +x face (mesh 16, light 0, radiance 60) in place of the four ceiling lights."""
    parts, meshes, mats, lights = posed_parts()
    parts = parts[:4]
    P, N, I = quad((0.252, 0.0, 0.0), 0.2, (1.0, 0.0, 0.0))
    parts.append((P, N, I, np.full(I.shape[0], 16), 0))
    return parts, meshes[:17], mats, [[60.0, 60.0, 60.0, 0.0]]


def posed_scene(parts, meshes, mats, lights, place=None, emit_body=None):
    """Scene of posed_parts: bodies kept in their own frames (place=None, uploaded with tri_body), or every body
    moved into the world by place[b] = (R, t) as static geometry (the reference scene for a posed one).
    emit_body: the body the light rides on (bodies kept in their frames only)."""
    B = Builder()
    tb = []
    for P, N, I, M, body in parts:
        if body is not None and place is not None:
            R, t = place[body]
            P = np.asarray(P, float) @ np.asarray(R).T + np.asarray(t)
            N = np.asarray(N, float) @ np.asarray(R).T
        B.add(P, N, I, M)
        tb.append(np.full(np.asarray(I).reshape(-1, 3).shape[0], STATIC_BODY if body is None or place is not None
                          else body, np.uint32))
    P, N, I, M = B.arrays()
    if place is not None:
        return pack_scene(P, N, I, M, meshes, mats, lights)
    nb = 1 + max(b for *_, b in parts if b is not None)
    sc = pack_scene(P, N, I, M, meshes, mats, lights, tri_body=np.concatenate(tb), body_count=nb)
    if emit_body is not None:
        sc["emit_body"] = int(emit_body)
    return sc


def printed_material(paper=(0.85, 0.84, 0.80), ink=(0.04, 0.04, 0.05), rough_paper=0.6, rough_ink=0.35):
    """A printed dielectric (the reference's charuco_print, gpu/scene_gpu.js:180-192, gpu/renderer.js:293): the
    paper's rho_d and alpha in words 0-3, the ink's rho_d and alpha (roughness^2) in words 8-11."""
    rec = material_record(KIND_DIELECTRIC, paper, roughness=rough_paper)
    rec[8:11] = ink
    rec[11] = max(float(rough_ink), 0.02) ** 2
    return rec


def board_mask_pattern(w=96, h=72, seed=2011):
    """A print mask to test with (255 = paper): a 6 x 4 chequer with a horizontal grey ramp over its top sixth and
    scattered single ink texels, so the bilinear filter meets sharp and smooth transitions alike."""
    y, x = np.mgrid[0:h, 0:w]
    m = np.where(((x * 6) // w + (y * 4) // h) % 2 == 0, 255, 0).astype(np.uint8)
    m[: h // 6] = (x[: h // 6] * 255 // (w - 1)).astype(np.uint8)
    m.flat[Rng(seed).int(0, w * h - 1, n=60)] = 0
    return m


def quad_uv(P, size, normal, up_hint=(0.0, 1.0, 0.0)):
    """Texture coordinates of quad()'s corners: u across, v down the image, (0, 0) at the top-left corner."""
    n = np.asarray(normal, float) / np.linalg.norm(normal)
    u = np.cross(up_hint, n)
    if np.linalg.norm(u) < 1e-6:
        u = np.cross((1.0, 0.0, 0.0), n)
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    c = np.mean(np.asarray(P, float), axis=0)
    d = np.asarray(P, float) - c
    return np.stack([d @ u / size + 0.5, 0.5 - d @ v / size], axis=1).astype(np.float32)


def textured_scene(mask=None):
    """posed_parts with two printed boards on the printed material: a static one standing beside the fixed box
    and one on body 0's -y face (bodies placed as set_instances poses them), with texture coordinates per vertex
    and a print mask (board_mask_pattern by default)."""
    parts, meshes, mats, lights = posed_parts()
    mats = list(mats) + [printed_material()]
    m_static, m_body = len(meshes), len(meshes) + 1
    meshes = list(meshes) + [[len(mats) - 1, MESH_STATIC | MESH_TEXTURED, 0, 0], [len(mats) - 1, MESH_TEXTURED, 0, 0]]
    uv_parts = [np.zeros((np.asarray(P).shape[0], 2), np.float32) for P, *_ in parts]
    for center, size, normal, mesh, body in (((0.35, -1.12, 0.95), 0.6, (0.0, -1.0, 0.0), m_static, None),
                                             ((0.0, -0.152, 0.0), 0.28, (0.0, -1.0, 0.0), m_body, 0)):
        P, N, I = quad(center, size, normal, up_hint=(0.0, 0.0, 1.0))
        parts.append((P, N, I, np.full(I.shape[0], mesh), body))
        uv_parts.append(quad_uv(P, size, normal, up_hint=(0.0, 0.0, 1.0)))
    sc = posed_scene(parts, meshes, mats, lights)
    sc["uvs"] = np.concatenate(uv_parts).reshape(-1)
    sc["board_mask"] = board_mask_pattern() if mask is None else mask
    return sc


def posed_camera(w=96, h=72, spp=8, depth=3, shutter=(0.0, 1.0)):
    rc = Rng(2004)
    lens = kb4_lens(rc, w, h, 45.0 * w / 96)
    pos = (0.2, -2.4, 1.3)
    cam = camera(look_at(pos, (0.0, 0.0, 0.5)), pos, lens, w, h, spp, 0, depth, seed=2005)
    cam["shutter"] = tuple(shutter)
    return cam


def c1(bvh="vt"):
    rs, rm, rl, rc = Rng(1001), Rng(1002), Rng(1003), Rng(1004)
    B = Builder()
    P, N, I = dome(3.0, 100, 100, pole=True)
    mesh = rs.int(0, 15, n=I.shape[0])
    B.add(P, N, I, mesh)
    P, N, I = box((-0.5, -1.0, 0.0), (0.5, 1.0, 0.8))
    B.add(P, N, I, rs.int(0, 15, n=I.shape[0]))
    mats = [material_record(KIND_LAMBERT, rm.uniform(0.2, 0.8, n=3)) for _ in range(16)]
    mats.append(material_record(KIND_LAMBERT, (0.0, 0.0, 0.0)))       # emitter surfaces
    meshes = [[m, MESH_STATIC, 0, 0] for m in range(16)]
    lights = []
    for k, (x, y) in enumerate(((1.0, 1.0), (-1.0, 1.0), (-1.0, -1.0), (1.0, -1.0))):
        L = rl.uniform(5.0, 20.0)
        lights.append([L, L, L, 0.0])
        meshes.append([16, MESH_EMITTER, k, 0])
        P, N, I = quad((x, y, 2.4), 0.6, (0.0, 0.0, -1.0))
        B.add(P, N, I, 16 + k)
    P, N, I, M = B.arrays()
    sc = pack_scene(P, N, I, M, meshes, mats, lights, bvh=bvh)
    lens = kb4_lens(rc, 640, 480, 300.0)
    pos = (0.3, -2.3, 1.5)
    cam = camera(look_at(pos, (0.0, 0.0, 0.5)), pos, lens, 640, 480, 16, 0, 3, seed=1005)
    return sc, [cam]


def c1_mixed(bvh="vt"):
    """C1's geometry, lights and camera with its 16 materials replaced by four of each kind: Lambert, GGX
    dielectric (roughness 0.1..0.6), GGX conductor and coated conductor (base roughness 0.2..0.5, clear coat
    0.05), seeded colours. The adjoint's material terms of every kind show up in one tile."""
    sc, cams = c1(bvh)
    rm = Rng(1012)
    mats = []
    for m in range(16):
        kind = (KIND_LAMBERT, KIND_DIELECTRIC, KIND_CONDUCTOR, KIND_COATED)[m % 4]
        col = rm.uniform(0.25, 0.85, n=3)
        rough = rm.uniform(0.1, 0.6)
        if kind == KIND_COATED:
            mats.append(material_record(kind, col, roughness=rough, tau=rm.uniform(0.7, 1.0, n=3)))
        else:
            mats.append(material_record(kind, col, roughness=rough))
    mats.append(material_record(KIND_LAMBERT, (0.0, 0.0, 0.0)))
    sc = dict(sc)
    sc["materials"] = np.concatenate(mats).astype(np.float32)
    return sc, cams


def _objects(rs, B, n_cyl, n_box, mesh_base):
    for k in range(n_cyl):
        r = rs.uniform(0.02, 0.15)
        h = rs.uniform(0.2, 1.2)
        ang = rs.uniform(0, 2 * math.pi)
        rad = rs.uniform(0.2, 2.2)
        tilt = rs.uniform(0, 1.0)
        az = rs.uniform(0, 2 * math.pi)
        R = rot_axis((math.cos(az), math.sin(az), 0.0), tilt)
        t = (rad * math.cos(ang), rad * math.sin(ang), 0.0)
        P, N, I = cylinder(r, h, 64, 72, R=R, t=t)
        B.add(P, N, I, mesh_base + rs.int(0, 255))
    for k in range(n_box):
        sx, sy, sz = rs.uniform(0.1, 0.6), rs.uniform(0.1, 0.6), rs.uniform(0.1, 0.6)
        ang = rs.uniform(0, 2 * math.pi)
        rad = rs.uniform(0.2, 2.2)
        yaw = rs.uniform(0, 2 * math.pi)
        R = rot_axis((0.0, 0.0, 1.0), yaw)
        t = (rad * math.cos(ang), rad * math.sin(ang), 0.0)
        P, N, I = box((-sx / 2, -sy / 2, 0.0), (sx / 2, sy / 2, sz), g=10, R=R, t=t)
        B.add(P, N, I, mesh_base + rs.int(0, 255))


def full_scene(bvh="vt"):
    """The full-scale scene (2,013,432 triangles, 256 materials, 16 lights)."""
    rs, rm, rl = Rng(1001), Rng(1002), Rng(1003)
    B = Builder()
    P, N, I = dome(3.0, 480, 1000, pole=True)
    B.add(P, N, I, rs.int(0, 255, n=I.shape[0]))
    _objects(rs, B, 100, 100, 0)
    mats = []
    for m in range(256):
        u = rm.u()
        rough = rm.uniform(0.05, 0.8)
        col = rm.uniform(0.2, 0.8, n=3)
        if u < 0.5:
            mats.append(material_record(KIND_LAMBERT, col))
        elif u < 0.85:
            mats.append(material_record(KIND_DIELECTRIC, col, roughness=rough, ior=1.5))
        else:
            mats.append(material_record(KIND_CONDUCTOR, rm.uniform(0.5, 0.95, n=3), roughness=rough))
    mats.append(material_record(KIND_LAMBERT, (0.0, 0.0, 0.0)))
    meshes = [[m, MESH_STATIC, 0, 0] for m in range(256)]
    lights = []
    for k in range(16):
        L = rl.uniform(5.0, 20.0)
        lights.append([L, L, L, 0.0])
        meshes.append([256, MESH_EMITTER, k, 0])
        ang = 2 * math.pi * k / 16 + rl.uniform(-0.1, 0.1)
        rad = rl.uniform(0.8, 1.6)
        z = rl.uniform(2.2, 2.5)
        c = (rad * math.cos(ang), rad * math.sin(ang), z)
        n = -np.asarray(c) / np.linalg.norm(c)
        P, N, I = quad(c, 0.4, n)
        B.add(P, N, I, 256 + k)
    P, N, I, M = B.arrays()
    return pack_scene(P, N, I, M, meshes, mats, lights, bvh=bvh)


def ring_cameras(rc, n, sizes, spps, depth, f_at_1920=600.0, first_seed=1005):
    cams = []
    for k in range(n):
        w, h = sizes[k]
        ang = 2 * math.pi * k / n + rc.uniform(-0.05, 0.05)
        z = rc.uniform(0.5, 1.5)
        pos = (2.5 * math.cos(ang), 2.5 * math.sin(ang), z)
        lens = kb4_lens(rc, w, h, f_at_1920 * w / 1920.0)
        cams.append(camera(look_at(pos, (0.0, 0.0, 0.4)), pos, lens, w, h, spps[k], k, depth, seed=first_seed + k))
    return cams


def c2(bvh="vt", scene=None):
    sc = scene if scene is not None else full_scene(bvh)
    rc = Rng(1004)
    cams = ring_cameras(rc, 3, [(1920, 1080)] * 3, [64] * 3, 5)
    return sc, cams


def c3(bvh="vt", scene=None):
    sc = scene if scene is not None else full_scene(bvh)
    rc = Rng(1004)
    choices = [(320, 240), (640, 480), (800, 600), (1024, 768), (1280, 960), (1600, 1200), (2048, 1536)]
    sizes = [choices[rc.int(0, 6)] for _ in range(24)]
    spps = [rc.pick([1, 4, 16, 64]) for _ in range(24)]
    return sc, ring_cameras(rc, 24, sizes, spps, 5)


def c5(bvh="vt", scene=None):
    sc = scene if scene is not None else full_scene(bvh)
    rc = Rng(1004)
    return sc, ring_cameras(rc, 3, [(1280, 960)] * 3, [32] * 3, 5)


def c4_nodes(n_theta=128, n_phi=128):
    """Dome nodes for the dense radiosity: triangle centroids of a 128 x 128 quad grid from 2 degrees off the
    pole (32,768 triangles), 8 floats each (p.xyz, area, inward n.xyz, -), and the scene (dome + table box) that
    masks their visibility."""
    P, N, I = dome(3.0, n_theta, n_phi, pole=False, theta0=math.radians(2.0))
    a, b, c = P[I[:, 0]], P[I[:, 1]], P[I[:, 2]]
    cen = (a + b + c) / 3
    nrm = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(nrm, axis=1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    inward = -cen / np.linalg.norm(cen, axis=1, keepdims=True)
    flip = np.sum(nrm * inward, axis=1) < 0
    nrm[flip] *= -1
    nodes = np.zeros((I.shape[0], 8), np.float32)
    nodes[:, 0:3] = cen
    nodes[:, 3] = area
    nodes[:, 4:7] = nrm
    B = Builder()
    B.add(P, N, I, 0)
    Pb, Nb, Ib = box((-0.5, -1.0, 0.0), (0.5, 1.0, 0.8))
    B.add(Pb, Nb, Ib, 0)
    Pa, Na, Ia, Ma = B.arrays()
    sc = pack_scene(Pa, Na, Ia, Ma, [[0, MESH_STATIC, 0, 0]], [material_record(KIND_LAMBERT, (0.5, 0.5, 0.5))],
                    [[0, 0, 0, 0]])
    return nodes, sc


def c4_emission(n, configs=16, seed=1007):
    """3 channels x `configs` emitter configurations (48 right-hand sides): each configuration lights 64 seeded
    nodes with radiosity U[5, 20] per channel."""
    r = Rng(seed)
    e = np.zeros((n, 3 * configs), np.float32)
    for c in range(configs):
        idx = r.int(0, n - 1, n=64)
        vals = r.uniform(5.0, 20.0, n=64 * 3).reshape(64, 3)
        for ch in range(3):
            e[idx, 3 * c + ch] = vals[:, ch]
    return e
