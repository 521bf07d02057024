"""Pins of the oracle's integrators by closed forms.

* AO over an open plane: no AO ray is occluded and every hit pixel equals
  the shading formula of computeAOIllumination (AOIntegrator.h:144-157)
  evaluated on the primary hit; misses are 1;
* AO inside a closed box: every AO ray is occluded, hit pixels are 0;
* direct lighting: a Lambertian plane under a parallel square area light
  (no occluders, one path vertex).  The per-frame estimator is
  kd * I * A * cos_n cos_l / r^2 at a uniform light point, so its mean and
  variance are integrals over the light; both are evaluated by dense
  quadrature in double precision, and the running mean over K frames must lie
  within 5 standard errors at every pixel;
* the running-mean accumulation weight 1/(image+1) (AOIntegrator.h:181-187).
"""
import numpy as np
import pytest

import pyoracle as O
from happyray_amd import scenes

F = np.float32
PI = F(np.pi)


def plane_scene(kd):
    """A 100x100 plane at y = 0 plus one small triangle far off to the side,
    so that y = 0 does not fall on a cell boundary (exact insertion drops a
    triangle lying exactly on the lower face of its cell range, as the
    reference does: its enlarged test box starts at that face)."""
    v, f = scenes.quad_grid((-50, 0, -50), (100, 0, 0), (0, 0, 100), 4, 4)
    tri = np.array([[49.0, 0.37, 49.0], [49.5, 0.37, 49.0], [49.0, 0.9, 49.5]])
    parts = [(v, f, 0), (tri, np.array([[0, 1, 2]], np.uint32), 1)]
    mats = [scenes.material(kd), scenes.material((0.1, 0.1, 0.1))]
    cam = dict(pos=(0.0, 10.0, 0.0), dir=(0.1, -1.0, 0.2), up=(0.0, 0.0, 1.0), fov=60.0, res=(48, 40))
    return scenes._assemble(parts, mats, "plane", cam)


def setup(sc, kind=O.GRID_UNIFORM_EXACT, lt=5.0):
    m = O.OMesh(sc.verts[0], sc.verts[1], sc.verts[2], sc.idx)
    bmin, bmax = O.bounds(sc.verts)
    g = O.OGrid(m, bmin, bmax, kind, lt, 3.0)
    c = sc.camera
    cam = O.camera(c["pos"], c["dir"], c["up"], c["fov"], *c["res"])
    return g, cam


def ao_shade_closed_form(kd, pdir, n, S):
    v = np.ones(3, F) * (kd * PI)
    if np.any(kd * PI < F(0.9)):
        s = max(F(0), np.sqrt(F(1) - n[2]))
        c = abs(F(pdir[0] * n[0] + pdir[1] * n[1]) + F(pdir[2] * n[2]))
        w = np.array([F(0.7) * c + F(0.3) * s, F(0.7) * c + F(0.2) * s, F(0.9) * c], F)
        v = (v * w).astype(F)
    else:
        v = np.ones(3, F)
    rad = np.zeros(3, F)
    for _ in range(S):
        rad = (rad + (v * F(1)) / F(S)).astype(F)
    return rad


@pytest.mark.parametrize("kind", [O.GRID_UNIFORM_EXACT, O.GRID_TWO_LEVEL])
def test_ao_open_plane(kind):
    kd = np.array([0.1, 0.2, 0.25], F)
    sc = plane_scene(kd)
    g, cam = setup(sc, kind, 5.0 if kind != O.GRID_TWO_LEVEL else 0.5)
    p = O.params(jitter=1, ao_samples=4, ao_dist=30.0)
    h, hp, _ = O.primary(g, cam, p)
    fb, occ, _ = O.render_ao(g, cam, p, sc.materials, sc.mat_idx)
    hit = hp < sc.nt
    assert hit.mean() > 0.5
    assert occ.reshape(-1, 4)[hit].sum() == 0
    np.testing.assert_array_equal(fb[~hit], np.ones((np.sum(~hit), 3), F))
    for i in np.nonzero(hit)[0][::7]:
        tri = sc.idx[hp[i]]
        v0, v1, v2 = (sc.verts[:, tri[k]] for k in range(3))
        e1 = (v1 - v0).astype(F)
        e2 = (v2 - v0).astype(F)
        rn = np.array([e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                       e1[0] * e2[1] - e1[1] * e2[0]], F)
        n = (rn * (F(1) / np.sqrt(F(rn[0] * rn[0] + rn[1] * rn[1]) + F(rn[2] * rn[2])))).astype(F)
        np.testing.assert_array_equal(fb[i], ao_shade_closed_form(kd, h[3:6, i], n, 4))


def test_ao_closed_box():
    """Inside the unit cube every AO ray hits a wall within sqrt(3) < 2, except
    those whose start, moved 0.01 * ao_dist along the ray
    (RayGenerators.h:340), leaves the box: only hit points within that
    distance of an edge can do so."""
    sc = scenes.small_box()
    g, cam = setup(sc)
    S, dist = 8, 2.0
    p = O.params(jitter=1, ao_samples=S, ao_dist=dist)
    fb, occ, _ = O.render_ao(g, cam, p, sc.materials, sc.mat_idx)
    h, hp, _ = O.primary(g, cam, p)
    assert np.all(hp < sc.nt)
    free = np.nonzero(occ.reshape(-1, S).min(1) == 0)[0]
    pts = h[0:3, free].T
    near = np.minimum(pts, 1 - pts) < 0.01 * dist + 1e-4
    assert np.all(near.sum(1) >= 2)          # near two faces: an edge
    assert occ.mean() > 0.95
    closed = np.setdiff1d(np.arange(len(hp)), free)
    assert np.all(fb[closed] == 0)


def test_running_mean_weights():
    sc = scenes.small_box()
    g, cam = setup(sc)
    p = O.params(jitter=1, ao_samples=2, ao_dist=0.3)
    f0, _, _ = O.render_ao(g, cam, p, sc.materials, sc.mat_idx, image_id=0)
    f1, _, _ = O.render_ao(g, cam, p, sc.materials, sc.mat_idx, image_id=1, fb=f0)
    n1, _, _ = O.render_ao(g, cam, p, sc.materials, sc.mat_idx, image_id=1)   # fresh buffer
    # fb1 = fb0 * (1 - 1/2) + new * 1/2, and a zero buffer gives new * 1/2
    np.testing.assert_array_equal(f1, (f0 * F(0.5) + (n1 / F(0.5)) * F(0.5)).astype(F))


def irradiance_integrals(p, n, light, m=400):
    """E[f] and E[f^2] of f = I*A*cos_n*cos_l/r^2 at a uniform light point."""
    pos, ln, e1, e2 = (light[0:3].astype(np.float64), light[3:6].astype(np.float64),
                       light[9:12].astype(np.float64), light[12:15].astype(np.float64))
    u = (np.arange(m) + 0.5) / m
    uu, vv = np.meshgrid(u, u, indexing="ij")
    q = pos + uu[..., None] * e1 + vv[..., None] * e2
    d = q - p
    r2 = (d * d).sum(-1)
    dn = d / np.sqrt(r2)[..., None]
    cn = np.maximum(0, dn @ n)
    cl = np.maximum(0, -(dn @ ln))
    area = np.linalg.norm(np.cross(e1, e2))
    f = area * cn * cl / r2
    return f.mean(), (f * f).mean()


def test_direct_lighting_integral():
    kd = np.array([0.3, 0.3, 0.3], F)
    sc = plane_scene(kd)
    light = scenes.area_light((-1.0, 4.0, -1.0), (2.0, 0.0, 0.0), (0.0, 0.0, 2.0), (5.0, 5.0, 5.0))
    assert light[4] < 0                                  # faces down
    g, cam = setup(sc)
    K = 48
    fb = None
    for k in range(K):
        p = O.params(seed=(7, 9), jitter=0, pt_depth=1)
        fb, _ = O.render_pt(g, cam, p, sc.materials, sc.mat_idx, light[None], image_id=k, fb=fb)
    h, hp, _ = O.primary(g, cam, O.params(jitter=0))
    hit = np.nonzero(hp < sc.nt)[0]
    z = []
    for i in hit[::5]:
        pnt = h[0:3, i].astype(np.float64) + 5 * 1e-5 * np.array([0, 1, 0])
        ef, ef2 = irradiance_integrals(pnt, np.array([0, 1.0, 0]), light)
        mean = 0.3 * 5.0 * ef
        sd = 0.3 * 5.0 * np.sqrt(max(ef2 - ef * ef, 1e-30) / K)
        z.append((fb[i, 0] - mean) / max(sd, 1e-6 * mean + 1e-12))
    z = np.array(z)
    assert np.abs(z).max() < 5.0
    assert abs(z.mean()) < 5.0 / np.sqrt(len(z)) + 0.05


def test_shading_normal_interpolation():
    """The path tracer's shading normal (PathTracer.h:109-135): at a corner it
    is that corner's normal; at the centroid the normalized mean of the three;
    at any interior point the weights are the barycentric coordinates, checked
    against barycentrics solved by least squares in double precision.  Without
    normals it is the geometric normal."""
    rng = np.random.default_rng(3)
    v = np.array([[0.2, 0.1, 0.0], [2.0, 0.3, 0.1], [0.4, 1.7, -0.2]], F)
    nrm = rng.normal(size=(3, 3)).astype(F)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    m = O.OMesh(v[:, 0], v[:, 1], v[:, 2], np.array([[0, 1, 2]], np.uint32),
                normals=nrm.T, nidx=np.array([[2, 0, 1]], np.uint32))
    corner_normal = nrm[[2, 0, 1]]          # normal index per corner
    for k in range(3):
        np.testing.assert_allclose(O.shading_normal(m, 0, v[k]), corner_normal[k], atol=2e-6)
    for _ in range(20):
        b = rng.dirichlet([1, 1, 1])
        p = (b @ v.astype(np.float64)).astype(F)
        expect = b @ corner_normal.astype(np.float64)
        expect /= np.linalg.norm(expect)
        np.testing.assert_allclose(O.shading_normal(m, 0, p), expect, atol=5e-5)
    flat = O.OMesh(v[:, 0], v[:, 1], v[:, 2], np.array([[0, 1, 2]], np.uint32))
    g = np.cross(v[1] - v[0], v[2] - v[0]).astype(np.float64)
    np.testing.assert_allclose(O.shading_normal(flat, 0, v.mean(0)), g / np.linalg.norm(g), atol=1e-6)


def test_paths_from_camera_rays_equal_the_render():
    """Paths from given rays, fed the camera's own pixel rays, reproduce the
    path-traced image 0 exactly: the hit record of a given ray is the primary
    hit record, and ray i takes the streams of pixel i."""
    sc = scenes.small_box()
    g, cam = setup(sc, O.GRID_TWO_LEVEL, 1.2)
    p = O.params(seed=(3, 4), jitter=1, pt_depth=4, nthreads=4)
    fb, _ = O.render_pt(g, cam, p, sc.materials, sc.mat_idx, sc.lights)
    d = O.primary_dirs(cam, p)
    org = np.repeat(np.asarray(cam.center[:], np.float32)[:, None], d.shape[1], 1)
    rgb, _ = O.path_trace_rays(g, org, d, p, sc.materials, sc.mat_idx, sc.lights)
    assert fb.max() > 0
    np.testing.assert_array_equal(rgb, fb)


# ---------------------------------------------------------------------------
# renderFrame mode 0: the simple integrator
# ---------------------------------------------------------------------------

def facing_triangle_scene(res=(40, 32)):
    """One large triangle in the plane z = 5 facing a camera at the origin
    that looks along +z; the triangle does not cover the whole image."""
    v = np.array([[-4.0, -3.0, 5.0], [4.5, -2.5, 5.0], [0.5, 3.5, 5.0]])
    parts = [(v, np.array([[0, 1, 2]], np.uint32), 0)]
    cam = dict(pos=(0.0, 0.0, 0.0), dir=(0.0, 0.0, 1.0), up=(0.0, 1.0, 0.0), fov=70.0, res=res)
    return scenes._assemble(parts, [scenes.material((0.5, 0.5, 0.5))], "facing", cam)


def test_simple_sampler_offsets():
    """RegularPixelSampler<2,2> (Camera.h:82-97) for images 0-3: the sample id
    picks the sub-pixel offsets (1/4, 1/4), (3/4, 1/4), (1/4, 3/4), (3/4, 3/4);
    from image 4 on mode 0 uses the Gaussian pixel filter (RTEngine.cu:414-423)."""
    sc = facing_triangle_scene()
    g, cam = setup(sc)
    rx, ry = sc.camera["res"]
    tl, stx, sty = (np.array(a[:], np.float64) for a in (cam.top_left, cam.step_x, cam.step_y))
    pix = np.arange(rx * ry)
    x, y = pix % rx, pix // rx
    offsets = [(0.25, 0.25), (0.75, 0.25), (0.25, 0.75), (0.75, 0.75)]
    for s, (ox, oy) in enumerate(offsets):
        d = O.primary_dirs(cam, O.params(jitter=0, simple_sampler=1), s)
        expect = tl[:, None] + (x + ox)[None] * stx[:, None] + (y + oy)[None] * sty[:, None]
        np.testing.assert_allclose(d, expect, rtol=0, atol=2e-5)
    for s in (4, 9):
        np.testing.assert_array_equal(O.primary_dirs(cam, O.params(jitter=0, simple_sampler=1), s),
                                      O.primary_dirs(cam, O.params(jitter=1), s))


def test_simple_shade_barycentric():
    """simpleShade (SimpleIntegrator.h:72-116): a hit pixel is the hit point's
    barycentric coordinates (weights of v0, v1, v2) times |cos| between the
    unit ray and the face normal; a miss is (0.15, 0.2, 0.3); the running mean
    weights are 1/(image+1)."""
    sc = facing_triangle_scene()
    g, cam = setup(sc)
    p = O.params()
    fb, _ = O.render_simple(g, cam, p, image_id=0)
    h, hp, _ = O.primary(g, cam, O.params(jitter=0, simple_sampler=1), 0)
    hit = hp < sc.nt
    assert 0.2 < hit.mean() < 0.9
    np.testing.assert_array_equal(fb[~hit], np.tile(np.array([0.15, 0.2, 0.3], F), (int((~hit).sum()), 1)))
    V = sc.verts.T.astype(np.float64)
    n = np.cross(V[1] - V[0], V[2] - V[0])
    n /= np.linalg.norm(n)
    for i in np.nonzero(hit)[0]:
        q = h[0:3, i].astype(np.float64)
        A = np.stack([V[0] - V[2], V[1] - V[2]], 1)
        lam = np.linalg.lstsq(A, q - V[2], rcond=None)[0]
        d = h[3:6, i].astype(np.float64)
        c = abs(np.dot(n, d / np.linalg.norm(d)))
        np.testing.assert_allclose(fb[i], np.array([lam[0], lam[1], 1 - lam[0] - lam[1]]) * c, atol=2e-4)
    f1, _ = O.render_simple(g, cam, p, image_id=1, fb=fb)
    n1, _ = O.render_simple(g, cam, p, image_id=1)
    np.testing.assert_array_equal(f1, (fb * F(0.5) + (n1 / F(0.5)) * F(0.5)).astype(F))


# ---------------------------------------------------------------------------
# renderFrame mode 1: the multi-kernel path tracer
# ---------------------------------------------------------------------------

def up_plane_scene(kd):
    """plane_scene with the plane's faces wound to face +y (towards the
    light), so that computeDirectIllumination's unflipped normal sees it."""
    v, f = scenes.quad_grid((-50, 0, -50), (100, 0, 0), (0, 0, 100), 4, 4)
    f = f[:, [0, 2, 1]]
    tri = np.array([[49.0, 0.37, 49.0], [49.5, 0.37, 49.0], [49.0, 0.9, 49.5]])
    parts = [(v, f, 0), (tri, np.array([[0, 1, 2]], np.uint32), 1)]
    mats = [scenes.material(kd), scenes.material((0.1, 0.1, 0.1))]
    cam = dict(pos=(0.0, 10.0, 0.0), dir=(0.1, -1.0, 0.2), up=(0.0, 0.0, 1.0), fov=60.0, res=(40, 32))
    return scenes._assemble(parts, mats, "up_plane", cam)


def test_path_multi_direct_light_at_absorbed_vertices():
    """Image 0 of the multi-kernel path tracer over a Lambertian plane under a
    downward area light (PathTracer.h:374-1215).  At the primary hit the
    roulette absorbs the path with probability 1 - pi*kd and divides the
    throughput by that; a path that bounces leaves the scene and adds
    nothing.  The end vertex of an absorbed path is the primary hit, where the
    shadow ray carries I*A*cos_l/r^2*0.5 from a uniform light point, the direct
    light is kd*that*cos_n, and image 0 stores throughput^2 x direct (the
    throughput multiplies both the new-radiance buffer and the merge).  So
    every hit pixel is 0 or att^2*kd*0.5*I*f(q) with f = A*cos_n*cos_l/r^2 at a
    light point q: f lies within its range over the light, the mean of
    f / E[f] is 1 within 5 standard errors, and the absorbed fraction is
    1 - pi*kd within 5 binomial standard errors."""
    kd = 0.05
    sc = up_plane_scene((kd, kd, kd))
    light = scenes.area_light((-1.0, 4.0, -1.0), (2.0, 0.0, 0.0), (0.0, 0.0, 2.0), (5.0, 5.0, 5.0))
    g, cam = setup(sc)
    p = O.params(seed=(11, 12), nthreads=4)
    fb, state, end, endp, imp, _ = O.render_pt_multi(g, cam, p, sc.materials, sc.mat_idx, light[None])
    h, hp, _ = O.primary(g, cam, O.params(seed=(11, 12), jitter=1))
    on_plane = np.nonzero(hp < sc.nt - 1)[0]
    assert len(on_plane) > 600
    albedo = F(kd) * PI
    att = F(1) / (F(1) - F(0) - albedo)
    absorbed = fb[on_plane, 0] > 0
    frac = absorbed.mean()
    q = 1 - float(albedo)
    assert abs(frac - q) < 5 * np.sqrt(q * (1 - q) / len(on_plane))
    # absorbed paths end at the primary hit with throughput att; bounced ones are black
    ia = on_plane[absorbed]
    np.testing.assert_array_equal(endp[ia], hp[ia])
    np.testing.assert_array_equal(end[6, ia], h[6, ia])
    np.testing.assert_array_equal(imp[ia], np.full((len(ia), 3), att, F))
    np.testing.assert_array_equal(fb[on_plane[~absorbed]], 0)
    ratios = []
    for i in ia[::3]:
        pnt = h[0:3, i].astype(np.float64)
        ef, ef2 = irradiance_integrals(pnt, np.array([0, 1.0, 0]), light, m=100)
        f = fb[i, 0] / (float(att) ** 2 * kd * 0.5 * 5.0)
        sd = np.sqrt(max(ef2 - ef * ef, 0.0))
        assert abs(f - ef) < 6 * sd + 1e-4 * ef
        ratios.append(f / ef)
    ratios = np.array(ratios)
    assert abs(ratios.mean() - 1) < 5 * ratios.std() / np.sqrt(len(ratios)) + 1e-3


def test_path_multi_merge_conserves_the_contribution():
    """addIndirectIllumination (PathTracer.h:889-990) over several images:
    each pixel's new contribution is importance x new radiance + the old
    residue; the blurred value replaces part of it only where the squared
    intensity reaches max(0.25/image, 0.05), so fb_new + residue_new / (k+1)
    equals fb_old * k/(k+1) + contribution / (k+1), the residue is never
    negative, and it is zero wherever the intensity stays under the threshold."""
    sc = scenes.small_box()
    g, cam = setup(sc, O.GRID_TWO_LEVEL, 1.2)
    p = O.params(seed=(5, 6), nthreads=4)
    fb = np.zeros((cam.res_x * cam.res_y, 3), F)
    state = None
    blurred_any = False
    for k in range(6):
        res_old = np.zeros_like(fb) if state is None else state[1].copy()
        fb_new, state, _, _, imp, _ = O.render_pt_multi(g, cam, p, sc.materials, sc.mat_idx, sc.lights,
                                                       image_id=k, fb=fb, state=state)
        nw = F(1) / F(k + 1)
        ow = F(1) - nw
        nc = (imp * state[0] + res_old).astype(F)
        lhs = fb_new.astype(np.float64) + state[1].astype(np.float64) * nw
        rhs = fb.astype(np.float64) * ow + nc.astype(np.float64) * nw
        np.testing.assert_allclose(lhs, rhs, rtol=2e-6, atol=1e-7)
        assert (state[1] >= 0).all()
        var = nc.astype(np.float64).sum(1) ** 2
        thr = max(0.25 / k, 0.05) if k > 0 else np.inf
        low = var < thr * (1 - 1e-5)
        assert (state[1][low] == 0).all()
        blurred_any |= bool((state[1] > 0).any())
        fb = fb_new
    assert fb.max() > 0 and blurred_any


def test_path_multi_without_lights_is_a_no_op():
    """PathTracer::integrate returns at once without light sources
    (PathTracer.h:1049-1053)."""
    sc = scenes.small_box()
    g, cam = setup(sc)
    fb0 = np.random.default_rng(1).random((cam.res_x * cam.res_y, 3)).astype(F)
    st0 = np.random.default_rng(2).random((2, cam.res_x * cam.res_y, 3)).astype(F)
    fb, st, _, _, _, _ = O.render_pt_multi(g, cam, O.params(), sc.materials, sc.mat_idx,
                                           np.zeros((0, 15), F), image_id=3, fb=fb0, state=st0)
    np.testing.assert_array_equal(fb, fb0)
    np.testing.assert_array_equal(st, st0)
