"""Seeded synthetic scenes for the hot path's workloads.

The reference ships no scene data (its scenes.7z is a Git-LFS pointer), so the
workloads are generated: every generator is a pure function of its seed.

Geometry is returned as a :class:`Scene`: SoA float32 vertices (3, nv),
uint32 index triplets (nt, 3), a PhongMaterial table (nm, 12) laid out as
src/RT/Primitive/Material.hpp:37-84 (kd.xyz, ior+10 flag, ks.xyz, exponent,
ke.xyz, pad) and one material index per triangle.
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Scene:
    verts: np.ndarray            # (3, nv) float32, SoA
    idx: np.ndarray              # (nt, 3) uint32
    materials: np.ndarray        # (nm, 12) float32
    mat_idx: np.ndarray          # (nt,) uint32
    camera: dict = field(default_factory=dict)
    lights: np.ndarray = None    # (nl, 15) float32: position, normal, intensity, edge1, edge2
    name: str = ""

    @property
    def nt(self):
        return int(self.idx.shape[0])

    @property
    def nv(self):
        return int(self.verts.shape[1])


def material(kd, ks=(0.0, 0.0, 0.0), spec_exp=1.0, ke=(0.0, 0.0, 0.0), ior=None):
    """One PhongMaterial row. ior set => refractive (w = ior + 10)."""
    w = 0.0 if ior is None else float(ior) + 10.0
    return np.array([kd[0], kd[1], kd[2], w, ks[0], ks[1], ks[2], spec_exp,
                     ke[0], ke[1], ke[2], 0.0], np.float32)


def area_light(position, edge1, edge2, intensity):
    """AreaLightSource: the normal is normalize(edge1 x edge2) reversed so that
    it faces along -(e1 x e2) when lighting downwards is wanted by the caller."""
    p = np.asarray(position, np.float32)
    e1 = np.asarray(edge1, np.float32)
    e2 = np.asarray(edge2, np.float32)
    n = np.cross(e1, e2).astype(np.float32)
    n = (n / np.linalg.norm(n)).astype(np.float32)
    return np.concatenate([p, n, np.asarray(intensity, np.float32), e1, e2]).astype(np.float32)


# ---------------------------------------------------------------------------
# primitives
# ---------------------------------------------------------------------------

def icosphere(subdiv):
    """Unit icosphere: 10*4^s + 2 vertices, 20*4^s triangles."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts, np.float64), np.array(faces, np.uint32)


def _assemble(parts, materials, name, camera, lights=None):
    vs, ids, mis = [], [], []
    off = 0
    for verts, faces, mi in parts:
        vs.append(np.asarray(verts, np.float32))
        ids.append(np.asarray(faces, np.uint32) + off)
        mis.append(np.full(len(faces), mi, np.uint32))
        off += len(verts)
    V = np.concatenate(vs).astype(np.float32)
    return Scene(verts=np.ascontiguousarray(V.T), idx=np.concatenate(ids).astype(np.uint32),
                 materials=np.stack(materials).astype(np.float32),
                 mat_idx=np.concatenate(mis).astype(np.uint32), camera=camera,
                 lights=None if lights is None else np.stack(lights).astype(np.float32),
                 name=name)


def quad_grid(origin, du, dv, nu, nv_):
    """Regular nu x nv_ quad tessellation of the parallelogram origin + [0,1]du + [0,1]dv."""
    o = np.asarray(origin, np.float64)
    du = np.asarray(du, np.float64)
    dv = np.asarray(dv, np.float64)
    i, j = np.meshgrid(np.arange(nu + 1), np.arange(nv_ + 1), indexing="ij")
    P = o + (i[..., None] / nu) * du + (j[..., None] / nv_) * dv
    verts = P.reshape(-1, 3)
    a = (i[:-1, :-1] * (nv_ + 1) + j[:-1, :-1]).reshape(-1)
    b = a + (nv_ + 1)
    c = b + 1
    d = a + 1
    faces = np.stack([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], 1).reshape(-1, 3)
    return verts, faces.astype(np.uint32)


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

def icosphere_cluster(seed=1, n_spheres=16, subdiv=4, res=256):
    """16 subdivision-4 icospheres, centres U[-10,10]^3, radii U[0.5,2], and a
    2-triangle ground quad at y=-12 spanning the spheres' x/z extent:
    81,922 triangles, 40,996 vertices."""
    rng = np.random.default_rng(seed)
    sv, sf = icosphere(subdiv)
    centers = rng.uniform(-10.0, 10.0, size=(n_spheres, 3))
    radii = rng.uniform(0.5, 2.0, size=n_spheres)
    kds = rng.uniform(0.05, 0.25, size=(n_spheres, 3))
    parts, mats = [], []
    for s in range(n_spheres):
        parts.append((sv * radii[s] + centers[s], sf, s))
        mats.append(material(kds[s]))
    lo = (centers - radii[:, None]).min(0)
    hi = (centers + radii[:, None]).max(0)
    gv = np.array([[lo[0], -12.0, lo[2]], [hi[0], -12.0, lo[2]], [hi[0], -12.0, hi[2]],
                   [lo[0], -12.0, hi[2]]])
    parts.append((gv, np.array([[0, 2, 1], [0, 3, 2]], np.uint32), n_spheres))
    mats.append(material((0.2, 0.2, 0.2)))
    cam = dict(pos=(0.0, 6.0, 38.0), dir=(0.0, -0.25, -1.0), up=(0.0, 1.0, 0.0), fov=45.0,
               res=(res, res))
    light = area_light((-5.0, 25.0, -5.0), (10.0, 0.0, 0.0), (0.0, 0.0, 10.0), (30.0, 30.0, 30.0))
    return _assemble(parts, mats, "icosphere_cluster", cam, [light])


def _room(extent=100.0):
    """Axis-aligned box [-extent/2, extent/2]^3 tessellated into exactly
    100,000 triangles: x and y faces 91x92 quads, z faces 86x96 quads."""
    h = extent / 2.0
    E = extent
    faces = [((-h, -h, -h), (0, E, 0), (0, 0, E), 91, 92),   # x = -h
             ((h, -h, -h), (0, 0, E), (0, E, 0), 92, 91),    # x = +h
             ((-h, -h, -h), (0, 0, E), (E, 0, 0), 91, 92),   # y = -h
             ((-h, h, -h), (E, 0, 0), (0, 0, E), 92, 91),    # y = +h
             ((-h, -h, -h), (E, 0, 0), (0, E, 0), 86, 96),   # z = -h
             ((-h, -h, h), (0, E, 0), (E, 0, 0), 96, 86)]    # z = +h
    vs, fs, off = [], [], 0
    for o, du, dv, nu, nv_ in faces:
        v, f = quad_grid(o, du, dv, nu, nv_)
        vs.append(v)
        fs.append(f + off)
        off += len(v)
    return np.concatenate(vs), np.concatenate(fs)


def soup_in_room(seed=2, n_soup=900_000, res=1024):
    """900k small triangles (vertices = centre U(unit cube) + N(0, 0.005)^3)
    inside a 100k-triangle room box of extent 100: 1,000,000 triangles."""
    rng = np.random.default_rng(seed)
    rv, rf = _room(100.0)
    c = rng.uniform(0.0, 1.0, size=(n_soup, 1, 3))
    sv = (c + rng.normal(0.0, 0.005, size=(n_soup, 3, 3))).reshape(-1, 3)
    sf = np.arange(3 * n_soup, dtype=np.uint32).reshape(-1, 3)
    mats = [material((0.15, 0.15, 0.15)), material((0.22, 0.18, 0.12))]
    cam = dict(pos=(0.5, 1.4, 3.2), dir=(0.0, -0.3, -1.0), up=(0.0, 1.0, 0.0), fov=50.0,
               res=(res, res))
    light = area_light((-2.0, 20.0, -2.0), (4.0, 0.0, 0.0), (0.0, 0.0, 4.0), (50.0, 50.0, 50.0))
    return _assemble([(rv, rf, 0), (sv, sf, 1)], mats, "soup_in_room", cam, [light])


def heightfield_keyframes(seed=3, quads=708, frames=30, extent=100.0, res=(1024, 768)):
    """(quads x quads)-quad heightfield, (quads+1)^2 vertices and 2*quads^2
    triangles (708 -> 502,681 vertices, 1,002,528 triangles), with `frames`
    key frames of y = sum_{i<4} A_i sin(k_i x + l_i z + phi_i + w_i f).
    Returns (scene at frame 0, key-frame vertex array (frames, 3, nv))."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.5, 2.0, 4)
    k = rng.uniform(-0.3, 0.3, 4)
    l_ = rng.uniform(-0.3, 0.3, 4)
    phi = rng.uniform(0.0, 2.0 * np.pi, 4)
    w = rng.uniform(0.1, 0.5, 4)
    h = extent / 2.0
    v, f = quad_grid((-h, 0.0, -h), (extent, 0.0, 0.0), (0.0, 0.0, extent), quads, quads)
    x = v[:, 0]
    z = v[:, 2]
    keys = np.empty((frames, 3, len(v)), np.float32)
    for fr in range(frames):
        y = np.zeros_like(x)
        for i in range(4):
            y += A[i] * np.sin(k[i] * x + l_[i] * z + phi[i] + w[i] * fr)
        keys[fr, 0] = x
        keys[fr, 1] = y
        keys[fr, 2] = z
    mats = [material((0.5 / np.pi, 0.5 / np.pi, 0.5 / np.pi))]
    cam = dict(pos=(0.0, 35.0, 60.0), dir=(0.0, -0.6, -1.0), up=(0.0, 1.0, 0.0), fov=55.0,
               res=res)
    light = area_light((-10.0, 60.0, -10.0), (20.0, 0.0, 0.0), (0.0, 0.0, 20.0),
                       (40.0, 40.0, 40.0))
    sc = _assemble([(keys[0].T, f, 0)], mats, "heightfield", cam, [light])
    return sc, keys


def small_box(seed=0):
    """Unit-cube room (12 triangles) plus one floating triangle: a tiny
    fixture in the spirit of SceneLoader::loadDefaultScene
    (src/Application/SceneLoader.cpp:335-443)."""
    rng = np.random.default_rng(seed)
    verts, faces = [], []
    cube_v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                       [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float64)
    cube_f = np.array([[0, 1, 2], [0, 2, 3], [4, 6, 5], [4, 7, 6], [0, 4, 5], [0, 5, 1],
                       [3, 2, 6], [3, 6, 7], [0, 3, 7], [0, 7, 4], [1, 5, 6], [1, 6, 2]],
                      np.uint32)
    tri = rng.uniform(0.3, 0.7, size=(3, 3))
    mats = [material((0.2, 0.2, 0.2)), material((0.1, 0.15, 0.25))]
    cam = dict(pos=(0.5, 0.5, 0.95), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov=60.0,
               res=(32, 32))
    light = area_light((0.4, 0.99, 0.4), (0.2, 0.0, 0.0), (0.0, 0.0, 0.2), (5.0, 5.0, 5.0))
    return _assemble([(cube_v, cube_f, 0), (tri, np.array([[0, 1, 2]], np.uint32), 1)], mats,
                     "small_box", cam, [light])


def random_soup(seed, n_tris, extent=10.0, size=0.5):
    """n random triangles of edge scale `size` in a cube of side `extent`."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, extent, size=(n_tris, 1, 3))
    v = (c + rng.uniform(-size, size, size=(n_tris, 3, 3))).reshape(-1, 3)
    f = np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3)
    mats = [material((0.2, 0.2, 0.2))]
    cam = dict(pos=(extent / 2, extent / 2, extent * 2.2), dir=(0.0, 0.0, -1.0),
               up=(0.0, 1.0, 0.0), fov=40.0, res=(64, 64))
    return _assemble([(v, f, 0)], mats, "random_soup", cam)


def sphere_cluster_n(seed, n_target, extent=20.0):
    """Icospheres (subdivision 3: 1,280 triangles; 2: 320 for small targets)
    with centres U[0, extent]^3 and radii U[0.5, 2] until about n_target
    triangles."""
    rng = np.random.default_rng(seed)
    sub = 3 if n_target >= 1280 else 2
    sv, sf = icosphere(sub)
    k = max(1, int(round(n_target / len(sf))))
    vs, fs = [], []
    for i in range(k):
        c = rng.uniform(0.0, extent, 3)
        r = rng.uniform(0.5, 2.0)
        vs.append(sv * r + c)
        fs.append(sf + i * len(sv))
    return np.concatenate(vs).astype(np.float32).T.copy(), np.concatenate(fs).astype(np.uint32)


def soup_n(seed, n, extent=20.0):
    """n small random triangles (edge scale extent / n^(1/3)) in a cube."""
    rng = np.random.default_rng(seed)
    size = extent / max(n, 1) ** (1.0 / 3.0)
    c = rng.uniform(0.0, extent, size=(n, 1, 3))
    v = (c + rng.normal(0.0, 0.5 * size, size=(n, 3, 3))).reshape(-1, 3)
    return v.astype(np.float32).T.copy(), np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)


def primary_rays(seed, verts, n, fov_deg=60.0):
    """n pinhole rays from a point outside the bounds (1.5 diagonals from the
    centre, random direction) through random image points aimed at the centre."""
    rng = np.random.default_rng(seed)
    lo, hi = verts.min(1), verts.max(1)
    c = (lo + hi) / 2
    diag = float(np.linalg.norm(hi - lo))
    w = rng.normal(size=3)
    w /= np.linalg.norm(w)
    eye = c + 1.5 * diag * w
    fwd = -w
    up = np.array([0.0, 1.0, 0.0]) if abs(fwd[1]) < 0.9 else np.array([1.0, 0.0, 0.0])
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    h = np.tan(np.radians(fov_deg) / 2)
    sx = rng.uniform(-h, h, n)
    sy = rng.uniform(-h, h, n)
    d = fwd[:, None] + right[:, None] * sx + up[:, None] * sy
    org = np.repeat(eye[:, None], n, 1)
    return org.astype(np.float32), d.astype(np.float32)


def ao_rays(seed, verts, idx, n, dist):
    """n ambient-occlusion rays from random points of random triangles:
    cosine-weighted directions about the face normal (random side), origin
    offset 0.01 * dist along the direction (as the AO generator offsets),
    to be traced as any-hit within dist."""
    rng = np.random.default_rng(seed)
    t = rng.integers(0, idx.shape[0], n)
    a, b, c = (verts[:, idx[t, k]].astype(np.float64) for k in range(3))
    u = rng.uniform(size=n)
    v = rng.uniform(size=n)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    p = a + (b - a) * u + (c - a) * v
    nrm = np.cross((b - a).T, (c - a).T).T
    nrm /= np.maximum(np.linalg.norm(nrm, axis=0), 1e-30)
    nrm *= np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0)
    r1 = rng.uniform(size=n)
    r2 = rng.uniform(size=n)
    phi = 2 * np.pi * r1
    lx, ly, lz = np.cos(phi) * np.sqrt(1 - r2), np.sin(phi) * np.sqrt(1 - r2), np.sqrt(r2)
    helper = np.where(np.abs(nrm[0]) > 0.9, 0.0, 1.0)
    tx = np.stack([helper, 1 - helper, np.zeros(n)])
    tng = np.cross(nrm.T, tx.T).T
    tng /= np.maximum(np.linalg.norm(tng, axis=0), 1e-30)
    bit = np.cross(nrm.T, tng.T).T
    d = tng * lx + bit * ly + nrm * lz
    org = p + 0.01 * dist * d
    return org.astype(np.float32), d.astype(np.float32)


def batched_workload(seed=4, n_scenes=256, tri_range=(1000, 200000), ray_range=(4000, 1000000),
                     scale=1.0):
    """The batched configuration: n_scenes scenes, sizes log-uniform in
    tri_range (even scenes random soups, odd scenes icosphere clusters); per
    scene a ray count log-uniform in ray_range, half primary (closest hit) and
    half ambient occlusion (any hit within 0.05 of the bounds diagonal) from
    random surface points.  `scale` shrinks the ray counts (tests).
    Returns (scenes [(verts, idx)], ray items [(scene, org, dir, anyhit, tmax)])."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]), n_scenes)).astype(np.int64)
    rays = np.exp(rng.uniform(np.log(ray_range[0]), np.log(ray_range[1]), n_scenes)) * scale
    rays = np.maximum(rays.astype(np.int64), 2)
    out_scenes, items = [], []
    for s in range(n_scenes):
        if s % 2 == 0:
            v, f = soup_n(seed * 1000 + s, int(sizes[s]))
        else:
            v, f = sphere_cluster_n(seed * 1000 + s, int(sizes[s]))
        out_scenes.append((v, f))
        diag = float(np.linalg.norm(v.max(1) - v.min(1)))
        n_p = int(rays[s] // 2)
        n_a = int(rays[s] - n_p)
        o, d = primary_rays(seed * 7919 + s, v, n_p)
        items.append((s, o, d, False, 0.0))
        o, d = ao_rays(seed * 104729 + s, v, f, n_a, 0.05 * diag)
        items.append((s, o, d, True, 0.05 * diag))
    return out_scenes, items


def served_workload(seed=5, n_scenes=64, tri_range=(10000, 2000000), n_requests=2000, rate=2000.0,
                    ray_range=(1000, 4000000), pool=262144, rebuild_frac=0.05):
    """The served configuration: n_scenes resident scenes (even: random soups,
    odd: icosphere clusters) of tri_range triangles, log-uniform, each with two
    key frames (the second displaced by a smooth field) and an area light
    above it; per scene a pool of primary rays and one of ambient-occlusion
    rays (any hit within 0.05 of the diagonal); n_requests Poisson arrivals at
    `rate` per second, each a rebuild (probability rebuild_frac) or a closest
    hit, occlusion or depth-4 path request of a log-uniform ray count.
    Returns (scenes [(keys (2,3,nv), idx, light)], pools [(prim_org, prim_dir,
    ao_org, ao_dir, ao_tmax)], requests [(t, kind, scene, n, offset, k1, k2, c)])."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]), n_scenes)).astype(np.int64)
    out_scenes, pools = [], []
    for s in range(n_scenes):
        if s % 2 == 0:
            v, f = soup_n(seed * 1000 + s, int(sizes[s]))
        else:
            v, f = sphere_cluster_n(seed * 1000 + s, int(sizes[s]))
        lo, hi = v.min(1), v.max(1)
        ext = hi - lo
        disp = 0.01 * float(ext.max()) * np.sin(v[0] * (3.0 / max(ext[0], 1e-3)) + s)
        v2 = v.copy()
        v2[1] += disp.astype(np.float32)
        keys = np.stack([v, v2]).astype(np.float32)
        c = (lo + hi) / 2
        light = area_light((c[0] - 0.2 * ext[0], hi[1] + 0.3 * ext[1], c[2] - 0.2 * ext[2]),
                           (0.4 * ext[0], 0.0, 0.0), (0.0, 0.0, 0.4 * ext[2]), (30.0, 30.0, 30.0))
        out_scenes.append((keys, f, light))
        diag = float(np.linalg.norm(ext))
        po, pd = primary_rays(seed * 7919 + s, v, pool)
        ao, ad = ao_rays(seed * 104729 + s, v, f, pool, 0.05 * diag)
        pools.append((po, pd, ao, ad, 0.05 * diag))
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    reqs = []
    for i in range(n_requests):
        s = int(rng.integers(0, n_scenes))
        if rng.uniform() < rebuild_frac:
            reqs.append((float(t[i]), "rebuild", s, 0, 0, 0, 1, float(rng.uniform())))
            continue
        kind = ("closest", "occluded", "path")[int(rng.integers(0, 3))]
        n = int(np.exp(rng.uniform(np.log(ray_range[0]), np.log(ray_range[1]))))
        reqs.append((float(t[i]), kind, s, n, 0, 0, 0, 0.0))
    return out_scenes, pools, reqs


# ---------------------------------------------------------------------------
# instanced scenes (the TLGRIDHIERARCHY build)
# ---------------------------------------------------------------------------

@dataclass
class Instancing:
    obj_first: np.ndarray        # (nobj + 1,) uint32: object k owns triangles [obj_first[k], obj_first[k+1])
    objects: np.ndarray          # (ninst,) uint32
    matrices: np.ndarray         # (ninst, 3, 4) float32: world-to-object rows (m_row_0..2)
    aabb_min: np.ndarray         # (ninst, 3) float32: world box of the instance
    aabb_max: np.ndarray


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def instanced_scene(seed=6, n_inst=64, subdiv=2, n_soup=200, extent=20.0, res=(64, 48), soup_size=0.3):
    """Three objects: 0 a room box of side `extent` centred at the origin
    (8x8 quads per face), 1 a unit icosphere, 2 a triangle soup in [-1, 1]^3
    (corners within +-soup_size of a centre).
    Instance 0 shows the room unmoved; the others place spheres and soups with
    a random rotation, a scale in [0.5, 1.5] and a translation inside the room.
    An instance box is the world bounds of the 8 transformed corners of its
    object's box, padded by 1e-3 (it holds every transformed triangle).  Returns (Scene, Instancing); the mesh holds every object once, in
    object coordinates, so its bounds (the top grid's) are the room's."""
    rng = np.random.default_rng(seed)
    h = extent / 2.0
    E = extent
    room_faces = [((-h, -h, -h), (0, E, 0), (0, 0, E)), ((h, -h, -h), (0, 0, E), (0, E, 0)),
                  ((-h, -h, -h), (0, 0, E), (E, 0, 0)), ((-h, h, -h), (E, 0, 0), (0, 0, E)),
                  ((-h, -h, -h), (E, 0, 0), (0, E, 0)), ((-h, -h, h), (0, E, 0), (E, 0, 0))]
    rv, rf, off = [], [], 0
    for o, du, dv in room_faces:
        v, f = quad_grid(o, du, dv, 8, 8)
        rv.append(v)
        rf.append(f + off)
        off += len(v)
    room_v, room_f = np.concatenate(rv), np.concatenate(rf)
    sph_v, sph_f = icosphere(subdiv)
    c = rng.uniform(-0.7, 0.7, size=(n_soup, 1, 3))
    soup_v = np.clip(c + rng.uniform(-soup_size, soup_size, size=(n_soup, 3, 3)), -1.0, 1.0).reshape(-1, 3)
    soup_f = np.arange(3 * n_soup, dtype=np.uint32).reshape(-1, 3)
    mats = [material((0.5, 0.5, 0.5)), material((0.7, 0.3, 0.2)), material((0.2, 0.4, 0.7))]
    cam = dict(pos=(-0.9 * h, -0.8 * h, 0.85 * h), dir=(1.0, 0.9, -0.8), up=(0.0, 0.0, 1.0), fov=70.0,
               res=res)
    sc = _assemble([(room_v, room_f, 0), (sph_v, sph_f, 1), (soup_v, soup_f, 2)], mats, "instanced", cam)
    counts = [len(room_f), len(sph_f), len(soup_f)]
    obj_first = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    obj_verts = [room_v, sph_v, soup_v]
    objects = [0]
    mats_w2o = [np.hstack([np.eye(3), np.zeros((3, 1))])]
    lo = [room_v.min(0) - 1e-3]
    hi = [room_v.max(0) + 1e-3]
    corners = []
    for ov in obj_verts:
        a, b = ov.min(0), ov.max(0)
        corners.append(np.array([[(a, b)[i][0], (a, b)[j][1], (a, b)[k][2]]
                                 for i in (0, 1) for j in (0, 1) for k in (0, 1)]))
    for i in range(1, n_inst):
        k = 1 + (i % 2)
        R = _rotation(rng)
        s = rng.uniform(0.5, 1.5)
        t = rng.uniform(-0.75 * h, 0.75 * h, size=3)
        A = R.T / s                              # world -> object
        M = np.hstack([A, (-A @ t)[:, None]])
        w = (corners[k] * s) @ R.T + t           # object -> world
        objects.append(k)
        mats_w2o.append(M)
        lo.append(w.min(0) - 1e-3)
        hi.append(w.max(0) + 1e-3)
    inst = Instancing(obj_first=obj_first, objects=np.array(objects, np.uint32),
                      matrices=np.array(mats_w2o, np.float32), aabb_min=np.array(lo, np.float32),
                      aabb_max=np.array(hi, np.float32))
    return sc, inst
