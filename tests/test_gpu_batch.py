"""Batches of scenes (hr_batch_*): the segmented build gives every scene the
grid a single-scene build gives it (bit for bit, read back scene-local), the
mixed persistent traces give every ray the single-scene result, buffers grow
after an overflow, and the autotuner's choice keeps both properties."""
import numpy as np
import pytest

import pyoracle as O
from gpu_util import SEED, assert_grid_equal, random_rays
from happyray_amd import GRID_TWO_LEVEL, Batch, Engine, scenes

pytestmark = pytest.mark.gpu


def scene_set():
    out = [scenes.icosphere_cluster(res=32)]
    for i, (n, size) in enumerate([(1, 0.5), (40, 8.0), (700, 0.3), (3000, 1.5), (12000, 0.2), (25000, 0.6)]):
        out.append(scenes.random_soup(100 + i, n, extent=10.0, size=size))
    out.append(scenes.small_box())
    return out


def single_engine(sc, lt, ll):
    e = Engine(seed=SEED)
    e.upload_scene(sc)
    e.set_grid(GRID_TWO_LEVEL, float(lt), float(ll))
    e.build_acc_struct()
    return e


def batch_grid_equal(b, s, e):
    gb, cb, lb, pb = b.grid(s)
    ge, ce, le, pe = e.grid()
    assert tuple(gb.res) == tuple(ge.res)
    np.testing.assert_array_equal(np.array(gb.bmin[:]), np.array(ge.bmin[:]))
    np.testing.assert_array_equal(np.array(gb.cell_size[:]), np.array(ge.cell_size[:]))
    assert (gb.num_cells, gb.num_leaves, gb.num_refs, gb.num_top_pairs) == \
        (ge.num_cells, ge.num_leaves, ge.num_refs, ge.num_top_pairs)
    np.testing.assert_array_equal(cb, ce)
    np.testing.assert_array_equal(lb, le)
    np.testing.assert_array_equal(pb, pe)


def ray_items(scs, seed):
    items = []
    rng = np.random.default_rng(seed)
    for s, sc in enumerate(scs):
        n = int(rng.integers(1, 3000))
        org, d = random_rays(sc, n, seed=seed + s)
        items.append((s, org, d, False, 0.0))
        ext = float((sc.verts.max(1) - sc.verts.min(1)).max())
        items.append((s, org, d, True, 0.3 * ext))
    items.append((0, np.zeros((3, 0), np.float32), np.zeros((3, 0), np.float32), False, 0.0))
    return items


def check_rays(items, res, engines):
    for (s, org, d, anyhit, tmax), r in zip(items, res):
        if org.shape[1] == 0:
            continue
        e = engines[s]
        if anyhit:
            np.testing.assert_array_equal(r[0], e.occluded(org, d, tmax))
        else:
            t, p = e.trace(org, d)
            np.testing.assert_array_equal(r[0], t)
            np.testing.assert_array_equal(r[1], p)


def test_batch_grids_and_traces_match_single_scenes():
    scs = scene_set()
    rng = np.random.default_rng(5)
    lt = rng.choice([1.0 / 16, 0.25, 1.2, 3.0], len(scs)).astype(np.float32)
    ll = rng.choice([1.0, 2.2, 3.0], len(scs)).astype(np.float32)
    b = Batch()
    b.set_scenes([(sc.verts, sc.idx) for sc in scs])
    b.set_densities(lt, ll)
    print(f"batch build of {len(scs)} scenes", flush=True)
    b.build()
    engines = []
    for s, sc in enumerate(scs):
        print(f"batch grid scene {s} ({sc.nt} tris, lt={lt[s]}, ll={ll[s]})", flush=True)
        e = single_engine(sc, lt[s], ll[s])
        batch_grid_equal(b, s, e)
        engines.append(e)
    # one scene against the oracle as well
    sc = scs[4]
    m = O.OMesh(sc.verts[0], sc.verts[1], sc.verts[2], sc.idx)
    bmin, bmax = O.bounds(sc.verts)
    og = O.OGrid(m, bmin, bmax, O.GRID_TWO_LEVEL, float(lt[4]), float(ll[4]))
    gi, cells, leaves, prims = b.grid(4)
    np.testing.assert_array_equal(cells, og.cells())
    np.testing.assert_array_equal(leaves, og.leaves())
    np.testing.assert_array_equal(prims, og.prims())
    items = ray_items(scs, 77)
    rays = b.make_rays(items)
    for refill, bpc in ((48, 0), (1, 2), (64, 1)):
        print(f"batch trace refill={refill} blocks_per_cu={bpc}", flush=True)
        b.set_tuning(refill, bpc)
        b.trace(rays)
        b.synchronize()
        check_rays(items, rays.results(), engines)
    bms, tms, n = b.timing()
    assert bms > 0 and tms > 0 and n >= rays.rays
    rays.free()
    for e in engines:
        e.close()
    b.close()


def test_batch_regrows_after_overflow():
    """Large triangles need far more pairs than the first capacity guess."""
    scs = [scenes.random_soup(300 + i, 60, extent=10.0, size=9.0) for i in range(3)] + \
          [scenes.random_soup(310, 2000, extent=10.0, size=0.1)]
    b = Batch()
    b.set_scenes([(sc.verts, sc.idx) for sc in scs])
    b.set_densities(40.0, 30.0)
    print("batch regrowth build", flush=True)
    b.build()
    engines = []
    for s, sc in enumerate(scs):
        e = single_engine(sc, 40.0, 30.0)
        batch_grid_equal(b, s, e)
        engines.append(e)
    items = ray_items(scs, 91)
    rays = b.make_rays(items)
    b.trace(rays)
    b.synchronize()
    check_rays(items, rays.results(), engines)
    rays.free()
    for e in engines:
        e.close()
    b.close()


def test_autotune_choice_keeps_results():
    scs = scene_set()[:6]
    b = Batch()
    b.set_scenes([(sc.verts, sc.idx) for sc in scs])
    items = ray_items(scs, 123)
    rays = b.make_rays(items)
    print("batch autotune", flush=True)
    r = b.autotune(rays)
    assert r.evaluations > 10
    assert 1 <= r.tuning.refill <= 64 and 0 <= r.tuning.blocks_per_cu <= 32
    lt, ll, tun = b.densities()
    assert tun == (r.tuning.refill, r.tuning.blocks_per_cu)
    b.trace(rays)
    b.synchronize()
    engines = [single_engine(sc, lt[s], ll[s]) for s, sc in enumerate(scs)]
    for s in range(len(scs)):
        batch_grid_equal(b, s, engines[s])
    check_rays(items, rays.results(), engines)
    rays.free()
    for e in engines:
        e.close()
    b.close()
