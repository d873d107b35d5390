"""rgh_analyse_frames_noalloc (no allocation by the library) against rgh_analyse_frames.

Two calls in a row through the C ABI, the second with fewer streams, frames and a smaller size:
each call's decisions, levels and final deblocked reconstructions must equal those of the
allocating entry point on the same inputs, bit for bit. The second call reuses the scratch
array the first one wrote (its progress records must be cleared again).
"""
import numpy as np
import pytest

from ragdoll_amd import DECISION_DTYPE, LEVELS_PER_MB, StreamAnalyzer, analyse_frames
from ragdoll_amd.hip import Hip
from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed

pytestmark = pytest.mark.gpu


def streams(name, n_streams, n_frames, seed):
    kw = dict(CONFIGS[name])
    w, h = kw.pop("width"), kw.pop("height")
    out = []
    for s in range(n_streams):
        g = SceneGenerator(w, h, seed=config_seed(seed, s), **kw)
        out.append([g.next_frame() for _ in range(n_frames)])
    return w, h, out


def run(hip, w, h, frames, qps, profile, me, lag, groups, noalloc):
    """Every stream's GOP (IDR first) in one call; returns per stream the decisions and levels of
    every frame and the final deblocked picture."""
    ans, srcs, decs, lvs, bufs = [], [], [], [], []
    try:
        for fr, qp in zip(frames, qps):
            a = StreamAnalyzer(w, h, qp, profile=profile, me_range=me, keep_undeblocked=False, ring=len(fr) + 1)
            ans.append(a)
            n = a.mb_w * a.mb_h
            s, d, l = [], [], []
            for y, u, v in fr:
                dy, du, dv = hip.upload(y), hip.upload(u), hip.upload(v)
                s.append((dy, du, dv, y.shape[1], u.shape[1], v.shape[1]))
                d.append(hip.alloc(n * 32))
                l.append(hip.alloc(n * LEVELS_PER_MB * 2))
                hip.memset(d[-1], 0xEE, n * 32)
                hip.memset(l[-1], 0xEE, n * LEVELS_PER_MB * 2)
                bufs += [dy, du, dv, d[-1], l[-1]]
            srcs.append(s)
            decs.append(d)
            lvs.append(l)
        analyse_frames(ans, srcs, decs, lvs, lag=lag, groups=groups, first_idr=True, noalloc=noalloc)
        out = []
        for a, d, l in zip(ans, decs, lvs):
            n = a.mb_w * a.mb_h
            out.append(([hip.download(x, np.zeros(n, DECISION_DTYPE)) for x in d],
                        [hip.download(x, np.zeros((n, LEVELS_PER_MB), np.int16)) for x in l],
                        a.outputs(which_recon=1, want_levels=False)[2]))
        return out
    finally:
        for a in ans:
            a.close()
        for b in bufs:
            hip.free(b)


def compare(want, got, label):
    for s, ((wd, wl, wr), (gd, gl, gr)) in enumerate(zip(want, got)):
        for k in range(len(wd)):
            bad = np.nonzero(wd[k].view(np.uint8).reshape(-1, 32) != gd[k].view(np.uint8).reshape(-1, 32))[0]
            assert len(bad) == 0, "%s stream %d frame %d: decision of MB %d differs" % (label, s, k, bad[0])
            assert np.array_equal(wl[k], gl[k]), "%s stream %d frame %d: levels differ" % (label, s, k)
        for p in range(3):
            assert np.array_equal(wr[p], gr[p]), "%s stream %d: final picture plane %d differs" % (label, s, p)


def test_noalloc_matches_allocating_entry_point_twice():
    hip = Hip()
    # Call 1: three 720p streams of 5 frames (IDR + 4 P), two concurrent groups, +-32.
    w, h, fr = streams("720p", 3, 5, 41)
    print("noalloc call 1: 3 x 720p x 5 frames", flush=True)
    args = (hip, w, h, fr, [24, 30, 36], "high", 32, 6, 2)
    compare(run(*args, noalloc=False), run(*args, noalloc=True), "call 1")
    # Call 2, smaller: one CIF stream of 3 frames, one group.
    w, h, fr = streams("cif", 1, 3, 42)
    print("noalloc call 2: 1 x CIF x 3 frames", flush=True)
    args = (hip, w, h, fr, [28], "baseline", 16, 3, 1)
    compare(run(*args, noalloc=False), run(*args, noalloc=True), "call 2")


def test_noalloc_refuses_more_frames_than_its_scratch_holds():
    from ragdoll_amd.abi import RghError
    a = StreamAnalyzer(32, 32, 26, keep_undeblocked=False)
    try:
        hip = Hip()
        n = a.mb_w * a.mb_h
        y = np.zeros((32, 32), np.uint8)
        c = np.zeros((16, 16), np.uint8)
        dy, dc = hip.upload(y), hip.upload(c)
        d, l = hip.alloc(n * 32), hip.alloc(n * LEVELS_PER_MB * 2)
        nf = 2049  # RGH_NOALLOC_MAX_FRAMES + 1, checked before any device work
        with pytest.raises(RghError) as e:
            analyse_frames([a], [[(dy, dc, dc, 32, 16, 16)] * nf], [[d] * nf], [[l] * nf], first_idr=True,
                           noalloc=True)
        assert e.value.status == 1
        for b in (dy, dc, d, l):
            hip.free(b)
    finally:
        a.close()
