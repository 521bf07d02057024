"""Captured frames (hr_render_frame_graph): interpolation, rebuild, primary
rays and the integrator replayed from recorded HIP graphs give the same grids
and images, bit for bit, as the same frames run call by call; eager work and
synchronizations run between the replays, and a replay whose rebuild outgrows
its buffers falls back and re-records."""
import numpy as np
import pytest

import pyoracle as O
from gpu_util import SEED, random_rays
from happyray_amd import (GRID_TWO_LEVEL, GRID_UNIFORM_EXACT, MODE_AO, MODE_PATH, MODE_PATH_MULTI, MODE_SIMPLE,
                          Engine, scenes)

pytestmark = pytest.mark.gpu


def make_engine(sc, keys, kind, lt, ll):
    e = Engine(seed=SEED)
    e.set_keyframes(keys, sc.idx)
    e.upload_materials(sc.materials, sc.mat_idx)
    e.set_grid(kind, lt, ll)
    e.set_camera_dict(sc.camera)
    e.set_lights(sc.lights)
    return e


def grids_equal(a, b):
    ia, ca, la, pa = a.grid()
    ib, cb, lb, pb = b.grid()
    assert tuple(ia.res) == tuple(ib.res)
    assert ia.num_refs == ib.num_refs and ia.num_cells == ib.num_cells
    np.testing.assert_array_equal(ca, cb)
    if la is not None or lb is not None:
        np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(pa, pb)


@pytest.mark.parametrize("lanes", [2, 3, 1])
@pytest.mark.parametrize("mode", [MODE_PATH, MODE_AO, MODE_SIMPLE, MODE_PATH_MULTI])
def test_frame_graph_matches_uncaptured_frames(mode, lanes):
    sc, keys = scenes.heightfield_keyframes(quads=64, frames=5, res=(64, 48))
    g = make_engine(sc, keys, GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    assert g.tuning().frame_lanes == 2
    g.set_tuning(frame_lanes=lanes)
    e = make_engine(sc, keys, GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    org, d = random_rays(sc, 4000, seed=9)
    kw = dict(jitter=1, pt_depth=3) if mode != MODE_AO else dict(jitter=1, ao_samples=4, ao_dist=5.0)
    for i, cur in enumerate([0.0, 0.4, 1.7, 3.9, 2.2, 0.1, 4.5]):
        k1, k2, c = O.frame_ids(cur, 0.4, 5)
        image = i % 3
        print(f"graph frame mode={mode} t={cur} keys=({k1},{k2}) image={image}", flush=True)
        fb_g = g.render_frame_graph(k1, k2, c, image_id=image, mode=mode, **kw)
        e.set_frame(k1, k2, c)
        e.build_acc_struct()
        fb_e = e.render_frame(image, mode, **kw)
        np.testing.assert_array_equal(fb_g, fb_e)
        grids_equal(g, e)
        # eager work on the captured context between replays
        t_g, p_g = g.trace(org, d)
        t_e, p_e = e.trace(org, d)
        np.testing.assert_array_equal(t_g, t_e)
        np.testing.assert_array_equal(p_g, p_e)
        tm = g.timing()
        assert tm.build_ms > 0 and tm.render_ms > 0
    g.close()
    e.close()


def growing_soup():
    """Key frame 1 turns the small triangles of key frame 0 into large ones, so
    a rebuild of key frame 1 outgrows buffers sized for key frame 0."""
    rng = np.random.default_rng(21)
    n = 3000
    c = rng.uniform(0, 10, (n, 1, 3))
    small = c + rng.normal(0, 0.02, (n, 3, 3))
    big = c + rng.normal(0, 2.5, (n, 3, 3))
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    keys = np.stack([small.reshape(-1, 3).T, big.reshape(-1, 3).T]).astype(np.float32)
    sc = scenes._assemble([(small.reshape(-1, 3), idx, 0)], [scenes.material((0.3, 0.3, 0.3))], "soup",
                          dict(pos=(5.0, 5.0, 30.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov=50.0,
                               res=(48, 40)),
                          [scenes.area_light((2.0, 14.0, 2.0), (6.0, 0.0, 0.0), (0.0, 0.0, 6.0),
                                             (20.0, 20.0, 20.0))])
    return sc, keys


# frames of growing_soup: the third outgrows the buffers the graphs were recorded with
GROWING_FRAMES = [(0, 0, 0.0), (0, 0, 0.0), (1, 1, 0.0), (1, 1, 0.0), (0, 1, 0.5), (0, 0, 0.0)]


def test_frame_graph_regrows_after_overflow():
    """The replayed rebuild of key frame 1 overflows the pair buffers sized
    for key frame 0; the frame is rendered again without graphs and the graphs
    are recorded anew."""
    sc, keys = growing_soup()
    for kind, lt, ll in ((GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_TWO_LEVEL, 0.5, 3.0)):
        g = make_engine(sc, keys, kind, lt, ll)
        e = make_engine(sc, keys, kind, lt, ll)
        for i, (k1, k2, cf) in enumerate(GROWING_FRAMES):
            print(f"graph regrowth kind={kind} frame={i}", flush=True)
            fb_g = g.render_frame_graph(k1, k2, cf, image_id=0, mode=MODE_PATH, pt_depth=2)
            e.set_frame(k1, k2, cf)
            e.build_acc_struct()
            fb_e = e.render_frame(0, MODE_PATH, pt_depth=2)
            np.testing.assert_array_equal(fb_g, fb_e)
            grids_equal(g, e)
        g.close()
        e.close()


@pytest.mark.parametrize("lanes", [2, 3, 1])
@pytest.mark.parametrize("mode", [MODE_PATH, MODE_AO, MODE_PATH_MULTI])
def test_overflowed_replay_keeps_running_mean(mode, lanes):
    """A replay whose rebuild overflows is skipped on the device, so the
    running mean over images 0..5 and the traced-ray count equal those of the
    same frames run call by call."""
    sc, keys = growing_soup()
    kw = dict(jitter=1, pt_depth=2) if mode != MODE_AO else dict(jitter=1, ao_samples=4, ao_dist=3.0)
    for kind, lt, ll in ((GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_TWO_LEVEL, 0.5, 3.0)):
        g = make_engine(sc, keys, kind, lt, ll)
        g.set_tuning(frame_lanes=lanes)
        e = make_engine(sc, keys, kind, lt, ll)
        for i, (k1, k2, cf) in enumerate(GROWING_FRAMES):
            print(f"running mean across overflow mode={mode} lanes={lanes} kind={kind} image={i}", flush=True)
            fb_g = g.render_frame_graph(k1, k2, cf, image_id=i, mode=mode, **kw)
            e.set_frame(k1, k2, cf)
            e.build_acc_struct()
            fb_e = e.render_frame(i, mode, **kw)
            np.testing.assert_array_equal(fb_g, fb_e)
            grids_equal(g, e)
        assert g.timing().rays_traced == e.timing().rays_traced
        g.close()
        e.close()


def run_submitted(g, frames, mode, kw, check_at=()):
    """Submits the frames without waiting; returns the framebuffers read at
    the frame indices in check_at (each read completes the frames in flight)."""
    got = {}
    for i, (k1, k2, cf) in enumerate(frames):
        g.submit_frame_graph(k1, k2, cf, image_id=i, mode=mode, **kw)
        if i in check_at:
            got[i] = g.framebuffer()
    return got


def run_called(e, frames, mode, kw):
    fbs = []
    for i, (k1, k2, cf) in enumerate(frames):
        e.set_frame(k1, k2, cf)
        e.build_acc_struct()
        fbs.append(e.render_frame(i, mode, **kw))
    return fbs


@pytest.mark.parametrize("lanes", [2, 3, 1])
@pytest.mark.parametrize("mode", [MODE_PATH, MODE_AO, MODE_SIMPLE])
def test_submitted_frames_match_called_frames(mode, lanes):
    """Frames queued by hr_submit_frame_graph (two in flight) give the same
    running mean, grid and ray count as frames run call by call, read back in
    the middle of the sequence and at its end."""
    sc, keys = scenes.heightfield_keyframes(quads=64, frames=5, res=(64, 48))
    kw = dict(jitter=1, pt_depth=3) if mode != MODE_AO else dict(jitter=1, ao_samples=4, ao_dist=5.0)
    frames = [O.frame_ids(t, 0.4, 5) for t in (0.0, 0.4, 1.7, 3.9, 2.2, 0.1, 4.5, 1.1, 2.9, 3.3)]
    g = make_engine(sc, keys, GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    g.set_tuning(frame_lanes=lanes)
    e = make_engine(sc, keys, GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    print(f"submitted frames mode={mode} lanes={lanes}", flush=True)
    got = run_submitted(g, frames, mode, kw, check_at=(4,))
    g.synchronize()
    last = g.framebuffer()
    want = run_called(e, frames, mode, kw)
    np.testing.assert_array_equal(got[4], want[4])
    np.testing.assert_array_equal(last, want[-1])
    grids_equal(g, e)
    # the hit buffers are those of the last frame, whichever lane ran it
    for a, b in zip(g.primary_hits(), e.primary_hits()):
        np.testing.assert_array_equal(a, b)
    if mode == MODE_AO:
        np.testing.assert_array_equal(g.ao_occlusion(4), e.ao_occlusion(4))
    assert g.timing().rays_traced == e.timing().rays_traced
    tm = g.timing()
    assert tm.build_ms > 0 and tm.render_ms > 0
    # replays after each slot's first (and after any re-recording) are timed inside their graphs
    assert 1 <= tm.frames_replayed <= len(frames) - 1
    assert tm.replay_build_ms > 0 and tm.replay_primary_ms > 0 and tm.replay_render_ms >= tm.replay_primary_ms
    g.close()
    e.close()


@pytest.mark.parametrize("lanes", [2, 3, 1])
def test_submitted_frames_recover_from_overflow(lanes):
    """An overflowing replay with another frame queued behind it: both are
    skipped on the device and rendered again when the host completes them
    (in a read of the framebuffer, or in the next submission), so the running
    mean equals that of the frames run call by call."""
    sc, keys = growing_soup()
    kw = dict(jitter=1, pt_depth=2)
    frames = GROWING_FRAMES + [(1, 1, 0.0), (0, 1, 0.25)]
    last = len(frames) - 1
    cases = [(kind, lt, ll, at) for kind, lt, ll in ((GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_TWO_LEVEL, 0.5, 3.0))
             for at in ((3, last), (last,))]
    for kind, lt, ll, at in cases:
        print(f"submitted frames across overflow lanes={lanes} kind={kind} reads at {at}", flush=True)
        g = make_engine(sc, keys, kind, lt, ll)
        g.set_tuning(frame_lanes=lanes)
        e = make_engine(sc, keys, kind, lt, ll)
        got = run_submitted(g, frames, MODE_PATH, kw, check_at=at)
        want = run_called(e, frames, MODE_PATH, kw)
        for i in at:
            np.testing.assert_array_equal(got[i], want[i])
        grids_equal(g, e)
        assert g.timing().rays_traced == e.timing().rays_traced
        g.close()
        e.close()
