"""Shared helpers of the GPU parity tests: run one scene through the product
(C ABI) and through the CPU oracle on identical inputs."""
import numpy as np

import pyoracle as O
from happyray_amd import Engine, GRID_TWO_LEVEL, GRID_UNIFORM, GRID_UNIFORM_EXACT

KIND_TO_ORACLE = {GRID_UNIFORM: O.GRID_UNIFORM, GRID_UNIFORM_EXACT: O.GRID_UNIFORM_EXACT,
                  GRID_TWO_LEVEL: O.GRID_TWO_LEVEL}

SEED = (0x1234, 0x5678)


def oracle_grid(scene, kind, lam_top, lam_leaf=2.2, verts=None):
    v = scene.verts if verts is None else verts
    m = O.OMesh(v[0], v[1], v[2], scene.idx)
    bmin, bmax = O.bounds(v)
    return O.OGrid(m, bmin, bmax, KIND_TO_ORACLE[kind], lam_top, lam_leaf)


def gpu_engine(scene, kind, lam_top, lam_leaf=2.2):
    e = Engine(seed=SEED)
    e.upload_scene(scene)
    e.set_grid(kind, lam_top, lam_leaf)
    e.build_acc_struct()
    return e


def assert_grid_equal(e, og):
    info, cells, leaves, prims = e.grid()
    p = og.params()
    assert tuple(info.res) == p["res"], (tuple(info.res), p["res"])
    np.testing.assert_array_equal(np.array(info.bmin[:], np.float32), p["bmin"])
    np.testing.assert_array_equal(np.array(info.bmax[:], np.float32), p["bmax"])
    np.testing.assert_array_equal(np.array(info.cell_size[:], np.float32), p["cell_size"])
    np.testing.assert_array_equal(np.array(info.cell_size_rcp[:], np.float32), p["cell_size_rcp"])
    assert info.num_cells == og.c.num_cells
    assert info.num_refs == og.c.num_refs, (info.num_refs, og.c.num_refs)
    np.testing.assert_array_equal(cells, og.cells())
    if og.c.kind == O.GRID_TWO_LEVEL:
        assert info.num_leaves == og.c.num_leaves
        assert info.num_top_pairs == og.c.num_top_pairs
        np.testing.assert_array_equal(leaves, og.leaves())
    np.testing.assert_array_equal(prims, og.prims())
    return info


def random_rays(scene, n, seed, inside_frac=0.3):
    """Rays aimed at random points of the scene from outside and inside its bounds."""
    rng = np.random.default_rng(seed)
    lo = scene.verts.min(1)
    hi = scene.verts.max(1)
    c = (lo + hi) / 2
    ext = (hi - lo).max()
    tgt = scene.verts[:, rng.integers(0, scene.nv, n)] + rng.normal(0, 0.01 * ext, (3, n))
    org = c[:, None] + rng.normal(0, ext, (3, n))
    k = int(n * inside_frac)
    org[:, :k] = lo[:, None] + rng.uniform(0, 1, (3, k)) * (hi - lo)[:, None]
    d = tgt - org
    # a few axis-parallel rays
    m = n // 16
    d[:, -m:] = 0.0
    ax = rng.integers(0, 3, m)
    d[ax, np.arange(n - m, n)] = rng.choice([-1.0, 1.0], m)
    return org.astype(np.float32), d.astype(np.float32)
