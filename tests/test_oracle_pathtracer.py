"""Pins of the oracle's path tracers (PathTracer.h:41-372 megakernel,
:374-990 multi-kernel) by closed forms that a plausible mistake fails.

* White furnace: a closed unit box whose six walls emit Le and are the six
  area lights (intensity 6 Le each: the light is picked with probability 1/6
  and the estimator does not divide by it), albedo a_c per channel.  Every
  point sees emitters over its whole hemisphere, so one next-event estimate
  has mean a_c Le; the roulette keeps a path with probability A = max a_c and
  scales the throughput by a_c / A, so vertex k contributes a_c^(k+1) Le and
  the primary hit Le: E[L_c] = Le * sum_{k<=d} a_c^k at depth cap d.  A
  dropped term, a roulette weight not divided by A, emission counted after a
  diffuse bounce or a missing last estimate each move the mean by many
  standard errors.  The ray count pins the roulette probability itself (the
  mean radiance does not depend on it): K vertices are reached with
  P(K >= k+1) = A^k, each vertex sends a shadow ray unless the light lies on
  its own wall (probability 5/6), so a path traces K + Binomial(K, 5/6) rays.
* Glass (or_pt_glass): the refracted direction obeys Snell's law, is unit
  and coplanar with the incident ray and the normal; the reflected one obeys
  the mirror law; the reflection frequency over a uniform grid of the random
  number is Schlick's F0 + (1 - F0)(1 - cos)^5 with F0 = ((n1-n2)/(n1+n2))^2;
  beyond the critical angle the ray always reflects and the throughput
  doubles or the path ends (a 1/2 roulette); leaving the material absorbs
  exp(-t kd pi).  A glass slab in front of an emitter transmits the fraction
  (1 - F0)^2 of its emission at normal incidence.
* Lobes: the power-cosine direction about an axis has E[cos] = (e+1)/(e+2)
  and E[cos^2] = (e+1)/(e+3); the cosine sampler 2/3 and 1/2; both through
  getLocalCoordinates, whose tangent is scaled by 1/(c0^2 + z^2) instead of
  its square root (Algebra.hpp:281-298), so the tangential part has squared
  length (1 - cos^2) / (c0^2 + z^2).
* Glossy bounce at normal incidence: never ends and keeps ks / max(ks).
* Mode 1 merge (addIndirectIllumination, PathTracer.h:954-987): the 3x3
  weights read from the kernel -- (1 2 1 / 2 4 2 / 1 2 1) / 16 with the
  centre tap reading the pixel above -- recovered from delta inputs, with
  neighbours clamped at the image border, and the variance threshold
  max(0.25 / image, 0.05).
"""
import numpy as np
import pytest

import pyoracle as O
from happyray_amd import scenes

F = np.float32
N_AIR = 1.000293


# ---------------------------------------------------------------------------
# white furnace
# ---------------------------------------------------------------------------

def furnace_scene():
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                 np.float32).T.copy()
    f = np.array([[0, 1, 2], [0, 2, 3], [4, 6, 5], [4, 7, 6], [0, 4, 5], [0, 5, 1], [3, 2, 6], [3, 6, 7],
                  [0, 3, 7], [0, 7, 4], [1, 5, 6], [1, 6, 2]], np.uint32)
    m = O.OMesh(v[0], v[1], v[2], f)
    bmin, bmax = O.bounds(v)
    return O.OGrid(m, bmin, bmax, O.GRID_UNIFORM_EXACT, 5.0, 1.0)


def furnace_lights(Le):
    # (corner, e1, e2) with e1 x e2 along the inward normal of each wall
    walls = [((0, 0, 0), (0, 0, 1), (1, 0, 0)), ((0, 1, 0), (1, 0, 0), (0, 0, 1)),
             ((0, 0, 0), (0, 1, 0), (0, 0, 1)), ((1, 0, 0), (0, 0, 1), (0, 1, 0)),
             ((0, 0, 0), (1, 0, 0), (0, 1, 0)), ((0, 0, 1), (0, 1, 0), (1, 0, 0))]
    return np.stack([scenes.area_light(p, e1, e2, tuple(6.0 * x for x in Le)) for p, e1, e2 in walls])


def ray_count_moments(A, d, p_shadow=5.0 / 6.0):
    """Mean and variance of K + Binomial(K, p) with P(K >= k + 1) = A^k, K <= d."""
    ks = np.arange(1, d + 1)
    pk = np.array([A ** (k - 1) * (1 - A) if k < d else A ** (d - 1) for k in ks])
    mean = np.sum(pk * ks * (1 + p_shadow))
    second = np.sum(pk * (ks * p_shadow * (1 - p_shadow) + (ks * (1 + p_shadow)) ** 2))
    return mean, second - mean ** 2


@pytest.mark.parametrize("albedo,depth", [((0.5, 0.5, 0.5), 1), ((0.5, 0.5, 0.5), 4), ((0.8, 0.8, 0.8), 4),
                                          ((0.8, 0.4, 0.1), 3), ((0.9, 0.6, 0.3), 6)])
def test_white_furnace(albedo, depth):
    Le = (1.0, 1.5, 0.5)
    g = furnace_scene()
    a = np.array(albedo, np.float64)
    mat = np.stack([scenes.material(tuple(np.float32(x) / F(np.pi) for x in albedo), ke=Le)])
    res = 256
    cam = O.camera((0.5, 0.5, 0.5), (0.1, 0.2, -1.0), (0.0, 1.0, 0.0), 100.0, res, res)
    vals, rays = [], 0
    for seed in ((1, 2), (3, 4), (5, 6)):
        p = O.params(seed=seed, jitter=1, pt_depth=depth, nthreads=8)
        fb, st = O.render_pt(g, cam, p, mat, None, furnace_lights(Le))
        vals.append(fb.astype(np.float64))
        rays += st.rays
    v = np.concatenate(vals)
    n = len(v)
    expect = np.array(Le) * np.array([sum(x ** k for k in range(depth + 1)) for x in a])
    mean = v.mean(0)
    se = v.std(0) / np.sqrt(n)
    z = (mean - expect) / se
    assert np.all(np.abs(z) < 5.0), (mean, expect, z)
    # roulette probability max(kd) * pi: rays per path
    m, var = ray_count_moments(float(a.max()), depth)
    zr = (rays - n * m) / np.sqrt(n * var)
    assert abs(zr) < 5.0, (rays / n, m, zr)


# ---------------------------------------------------------------------------
# glass
# ---------------------------------------------------------------------------

def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def schlick(n1, n2, c):
    f0 = ((n1 - n2) / (n1 + n2)) ** 2
    return f0 + (1 - f0) * (1 - c) ** 5


def incident(rng, n, cos_i):
    """A unit direction making angle acos(cos_i) with -n."""
    t = unit(np.cross(n, rng.normal(size=3)))
    s = np.sqrt(max(0.0, 1 - cos_i ** 2))
    return (-cos_i * n + s * t).astype(np.float32)


@pytest.mark.parametrize("mode1", [False, True])
def test_glass_refraction_obeys_snell(mode1):
    rng = np.random.default_rng(11)
    ior = 1.5
    for i in range(300):
        n = unit(rng.normal(size=3)).astype(np.float32)
        entering = i % 2 == 0
        n1, n2 = (N_AIR, ior) if entering else (ior, N_AIR)
        crit = np.arcsin(min(1.0, n2 / n1))
        th = rng.uniform(0.0, 0.99 * crit if not entering else 1.45)
        ns = n if entering else -n                  # the surface normal on the incident side
        d = incident(rng, ns.astype(np.float64), np.cos(th))
        cos_ne = float(-np.dot(n.astype(np.float64), d))   # negative when leaving
        org = rng.normal(size=3).astype(np.float32)
        st, o2, d2, att = O.pt_glass(org, d, (1, 1, 1), n, cos_ne, ior, 0.0, (0, 0, 0), 1.0, 0.0, mode1)
        assert st == 1, (i, th)
        d2 = d2.astype(np.float64)
        assert abs(np.linalg.norm(d2) - 1) < 2e-6
        sin_i = np.linalg.norm(np.cross(d, ns))
        sin_t = np.linalg.norm(np.cross(d2, ns))
        assert abs(n1 * sin_i - n2 * sin_t) < 5e-6 * n1, (i, n1 * sin_i, n2 * sin_t)
        assert abs(np.dot(np.cross(d, ns), d2)) < 5e-6       # in the plane of incidence
        assert np.dot(d2, ns) < 0                            # continues through the surface
        off = (10e-5 if mode1 else 5e-5)
        np.testing.assert_allclose(o2, org - off * ns, rtol=0, atol=2e-6)
        np.testing.assert_array_equal(att, np.ones(3, np.float32))


def test_glass_reflection_and_schlick_frequency():
    rng = np.random.default_rng(12)
    ior = 1.5
    grid = (np.arange(4000) + 0.5) / 4000
    for entering in (True, False):
        n1, n2 = (N_AIR, ior) if entering else (ior, N_AIR)
        crit = np.arcsin(n2 / n1) if n2 < n1 else np.pi / 2
        for th in np.linspace(0.0, 0.97 * min(crit, 1.5), 7):
            n = unit(rng.normal(size=3)).astype(np.float32)
            ns = n if entering else -n
            d = incident(rng, ns.astype(np.float64), np.cos(th))
            cos_ne = float(-np.dot(n, d))
            c = abs(cos_ne)
            refl = 0
            for w in grid:
                st, o2, d2, _ = O.pt_glass((0, 0, 0), d, (1, 1, 1), n, cos_ne, ior, 0.0, (0, 0, 0), float(w), 0.0)
                if st == 0:
                    refl += 1
                    # mirror law: the normal component flips, the tangential one stays
                    dn = np.dot(d2, ns)
                    assert abs(dn + np.dot(d, ns)) < 3e-6
                    np.testing.assert_allclose(d2 - dn * ns, d - np.dot(d, ns) * ns, atol=3e-6)
                    np.testing.assert_allclose(o2, 5e-5 * ns, atol=1e-9)
            F_ = schlick(n1, n2, c)
            assert abs(refl / len(grid) - F_) < 1.5 / len(grid) + 1e-6, (entering, th, refl / len(grid), F_)


def test_glass_total_internal_reflection_and_absorption():
    rng = np.random.default_rng(13)
    ior = 1.5
    crit = np.arcsin(N_AIR / ior)
    kd = np.array([0.05, 0.1, 0.0], np.float32)
    for i in range(100):
        n = unit(rng.normal(size=3)).astype(np.float32)
        th = rng.uniform(crit * 1.02, 1.5)
        d = incident(rng, -n.astype(np.float64), np.cos(th))      # inside: the ray leaves along +n
        cos_ne = float(-np.dot(n, d))
        assert cos_ne < 0
        t = float(rng.uniform(0.1, 3.0))
        beer = np.exp(-t * kd.astype(np.float64) * np.pi)
        for w_fres in (0.0, 0.5, 0.999999):
            st, _, d2, att = O.pt_glass((0, 0, 0), d, (1, 1, 1), n, cos_ne, ior, t, kd, w_fres, 0.25)
            assert st == 0
            np.testing.assert_allclose(att, 2.0 * beer, rtol=2e-6)
            assert np.dot(d2, n) < 0                                  # stays inside
            st, _, _, att = O.pt_glass((0, 0, 0), d, (1, 1, 1), n, cos_ne, ior, t, kd, w_fres, 0.75)
            assert st == 2
    # below the critical angle a leaving ray refracts (w_fres = 1) and absorbs
    n = np.array([0, 0, 1], np.float32)
    d = incident(rng, -n.astype(np.float64), np.cos(0.5 * crit))
    st, _, _, att = O.pt_glass((0, 0, 0), d, (1, 1, 1), n, float(-np.dot(n, d)), ior, 2.0, kd, 1.0, 0.9)
    assert st == 1
    np.testing.assert_allclose(att, np.exp(-2.0 * kd.astype(np.float64) * np.pi), rtol=2e-6)


def slab_scene(kd_glass):
    """A glass slab (IOR 1.5, |z| <= 0.1) between a camera on the +z axis and
    an emissive backdrop at z = -3 (Ke = (1, 2, 3), opaque, black)."""
    # (the slab's own kd is set by the caller)
    box_v = np.array([[-2, -2, -0.1], [2, -2, -0.1], [2, 2, -0.1], [-2, 2, -0.1],
                      [-2, -2, 0.1], [2, -2, 0.1], [2, 2, 0.1], [-2, 2, 0.1]], np.float64)
    # wound so that the geometric normals point out of the slab: the path
    # tracer tells entering from leaving by the normal's side (PathTracer.h:139-147)
    box_f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 5, 4], [0, 1, 5], [3, 6, 2], [3, 7, 6],
                      [0, 7, 3], [0, 4, 7], [1, 6, 5], [1, 2, 6]], np.uint32)
    bd_v, bd_f = scenes.quad_grid((-10, -10, -3), (20, 0, 0), (0, 20, 0), 2, 2)
    mats = [scenes.material(kd_glass, ior=1.5), scenes.material((0, 0, 0), ke=(1.0, 2.0, 3.0))]
    cam = dict(pos=(0.0, 0.0, 5.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov=2.0, res=(128, 128))
    return scenes._assemble([(box_v, box_f, 0), (bd_v, bd_f, 1)], mats, "slab", cam)


@pytest.mark.parametrize("kd", [(0.0, 0.0, 0.0), (0.3, 0.1, 0.0)])
def test_glass_slab_transmits_one_minus_f0_squared(kd):
    sc = slab_scene(kd)
    m = O.OMesh(sc.verts[0], sc.verts[1], sc.verts[2], sc.idx)
    bmin, bmax = O.bounds(sc.verts)
    g = O.OGrid(m, bmin, bmax, O.GRID_TWO_LEVEL, 1.0, 3.0)
    c = sc.camera
    cam = O.camera(c["pos"], c["dir"], c["up"], c["fov"], *c["res"])
    none = np.zeros((0, 15), np.float32)
    T = (1 - schlick(N_AIR, 1.5, 1.0)) ** 2
    beer = np.exp(-0.2 * np.array(kd) / np.float32(np.pi) * np.pi)   # kd is stored / pi below
    vals = []
    for seed in ((7, 8), (9, 10), (11, 12)):
        p = O.params(seed=seed, jitter=1, pt_depth=3, nthreads=8)
        mats = sc.materials.copy()
        mats[0, :3] = np.array(kd, np.float32) / F(np.pi)
        fb, _ = O.render_pt(g, cam, p, mats, sc.mat_idx, none)
        vals.append(fb.astype(np.float64))
        # depth 2 never reaches the backdrop: glass vertices emit nothing
        p2 = O.params(seed=seed, jitter=1, pt_depth=2, nthreads=8)
        fb2, _ = O.render_pt(g, cam, p2, mats, sc.mat_idx, none)
        assert fb2.max() == 0.0
    v = np.concatenate(vals)
    ke = np.array([1.0, 2.0, 3.0])
    lit = v[:, 0] > 0
    # each pixel is the backdrop's emission times the slab's absorption, or 0
    np.testing.assert_allclose(v[lit], np.broadcast_to(ke * beer, v[lit].shape), rtol=3e-5)
    frac = lit.mean()
    se = np.sqrt(T * (1 - T) / len(v))
    assert abs(frac - T) < 5 * se, (frac, T, se)


# ---------------------------------------------------------------------------
# lobes and the glossy bounce
# ---------------------------------------------------------------------------

def strat(m):
    g = (np.arange(m) + 0.5) / m
    return np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)


def frame_scale(axis):
    a = np.asarray(axis, np.float64)
    c0 = 0 if abs(a[0]) > abs(a[1]) else 1
    return a[c0] ** 2 + a[2] ** 2


@pytest.mark.parametrize("se", [0.0, 1.0, 4.0, 20.0, 200.0])
def test_power_cosine_lobe_moments(se):
    ws = strat(96)
    for axis in (unit((0.3, 0.8, 0.52)), unit((-0.9, 0.1, -0.42)), unit((0.0, 0.0, 1.0))):
        ax = axis.astype(np.float32).astype(np.float64)
        d = np.array([O.power_cosine_dir(float(a), float(b), se, ax) for a, b in ws], np.float64)
        c = d @ ax
        np.testing.assert_allclose(c.mean(), (se + 1) / (se + 2), atol=2e-4)
        np.testing.assert_allclose((c ** 2).mean(), (se + 1) / (se + 3), atol=2e-4)
        perp = d - np.outer(c, ax)
        L = frame_scale(ax)
        np.testing.assert_allclose((perp ** 2).sum(1), (1 - c ** 2) / L, rtol=1e-4, atol=1e-6)
        # azimuth uniform: no preferred tangential direction
        assert np.abs(perp.mean(0)).max() < 2e-3


def test_cosine_lobe_moments():
    ws = strat(128)
    for n in (unit((0.3, 0.8, 0.52)), unit((0.95, -0.2, 0.24))):
        nf = n.astype(np.float32).astype(np.float64)
        d = np.array([O.cosine_dir(float(a), float(b), nf) for a, b in ws], np.float64)
        c = d @ nf
        assert c.min() >= 0
        np.testing.assert_allclose(c.mean(), 2.0 / 3.0, atol=2e-4)
        np.testing.assert_allclose((c ** 2).mean(), 0.5, atol=2e-4)
        perp = d - np.outer(c, nf)
        np.testing.assert_allclose((perp ** 2).sum(1), (1 - c ** 2) / frame_scale(nf), rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("se", [1.0, 8.0, 60.0])
def test_glossy_bounce_conserves_at_normal_incidence(se):
    ks = np.array([0.3, 0.2, 0.05], np.float32)
    n = unit((0.2, -0.5, 0.84)).astype(np.float32)
    d = (-n).astype(np.float32)
    ws = strat(96)
    att = []
    for a, b in ws:
        alive, d2, at = O.pt_glossy(d, (1, 1, 1), n, float(-np.dot(n, d)), ks, se, float(a), float(b))
        assert alive == 1
        att.append(at)
    att = np.array(att, np.float64)
    np.testing.assert_allclose(att.mean(0), ks / ks.max(), rtol=3e-3)
    # at grazing incidence part of the lobe lies below the surface: paths end
    d = incident(np.random.default_rng(3), n.astype(np.float64), np.cos(1.45))
    ended = sum(O.pt_glossy(d, (1, 1, 1), n, float(-np.dot(n, d)), ks, 1.0, float(a), float(b))[0] == 0
                for a, b in strat(40))
    assert ended > 0


# ---------------------------------------------------------------------------
# mode 1 merge
# ---------------------------------------------------------------------------

# weights of the tap at (dx, dy) relative to the pixel, dy = +1 the row above
# (addIndirectIllumination's `top` is threadIdx.y + 1): the centre tap reads
# sharedMem[top] (PathTracer.h:964-968), so the pixel itself has weight 0 and
# the one above 2 + 4
WEIGHTS = {(-1, 1): 1, (0, 1): 6, (1, 1): 1, (-1, 0): 2, (0, 0): 0, (1, 0): 2, (-1, -1): 1, (0, -1): 2, (1, -1): 1}


def merge(nc, fb, image_id, rx, ry):
    n = rx * ry
    return O.pm_merge(np.ones((n, 3), np.float32), nc, np.zeros((n, 3), np.float32), fb, image_id, rx, ry)


def test_merge_blur_weights_from_delta_inputs():
    rx, ry, k = 7, 6, 1                      # image 1: new weight 1/2, threshold max(0.25, 0.05)
    B, v = 100.0, 1e-3                       # bright pixel (blurred branch), dim neighbour (plain branch)
    for (px, py) in ((3, 2), (0, 0), (6, 5), (0, 3)):
        for (dx, dy) in WEIGHTS:
            qx, qy = px + dx, py + dy
            if not (0 <= qx < rx and 0 <= qy < ry) or (dx, dy) == (0, 0):
                continue
            nc = np.zeros((ry, rx, 3), np.float32)
            nc[py, px] = B
            nc[qy, qx] = v
            fb0 = np.zeros((ry * rx, 3), np.float32)
            fb, res = merge(nc.reshape(-1, 3), fb0, k, rx, ry)
            fb = fb.reshape(ry, rx, 3)
            # the dim pixel keeps its plain running mean and no residue
            np.testing.assert_allclose(fb[qy, qx], v * 0.5, rtol=1e-6)
            # taps clamped at the border read the bright pixel itself
            w_self = sum(w for (ex, ey), w in WEIGHTS.items()
                         if (min(max(px + ex, 0), rx - 1), min(max(py + ey, 0), ry - 1)) == (px, py))
            w_q = sum(w for (ex, ey), w in WEIGHTS.items()
                      if (min(max(px + ex, 0), rx - 1), min(max(py + ey, 0), ry - 1)) == (qx, qy))
            blurred = (w_self * 0.5 * B + w_q * 0.5 * v) / 16.0
            val = min(blurred + 0.5 * B, B)
            np.testing.assert_allclose(fb[py, px], 0.5 * val, rtol=2e-6, err_msg=str((px, py, dx, dy)))
            np.testing.assert_allclose(res.reshape(ry, rx, 3)[py, px], B - val, rtol=2e-5)


def test_merge_variance_threshold():
    rx, ry = 3, 3
    for k, thr in ((1, 0.25), (4, 0.0625), (20, 0.05)):
        for s, blurred in ((0.98, False), (1.02, True)):
            c = np.sqrt(thr) / 3.0 * s               # (r + g + b)^2 = s^2 * threshold
            nc = np.full((rx * ry, 3), c, np.float32)
            fb, res = merge(nc, np.zeros((rx * ry, 3), np.float32), k, rx, ry)
            w = 1.0 / (k + 1)
            if blurred:
                # uniform image over a zero running mean: every tap reads w c,
                # so the pixel takes min(w c + c / 2, c) and keeps the rest
                val = min(w + 0.5, 1.0) * c
                np.testing.assert_allclose(fb, np.full_like(nc, w * val), rtol=2e-6)
                np.testing.assert_allclose(res, np.full_like(nc, c - val), rtol=2e-5, atol=1e-7)
            else:
                np.testing.assert_allclose(fb, nc * w, rtol=2e-6)
                assert np.all(res == 0)
            # a dark neighbourhood separates the branches: blurred pixels keep a residue
            nc2 = np.zeros((rx * ry, 3), np.float32)
            nc2[4] = c
            fb2, res2 = merge(nc2, np.zeros((rx * ry, 3), np.float32), k, rx, ry)
            assert (res2[4].sum() > 0) == blurred, (k, s)
