"""The pipelined schedule (rgh_analyse_frames) vs the CPU oracle, bit for bit.

Several frames of a stream are in flight at once: frame k + 1 starts `lag` wavefront steps after
frame k and its macroblocks wait (by deferral) for the reference samples they can read. A small
lag makes most macroblocks defer; a ring of 2 pictures serialises the frames through the
picture-reuse gate. Every frame's decisions and levels, and the last frame's deblocked
reconstruction, must equal the oracle's.
"""
import numpy as np
import pytest

from oracle_lib import OracleEncoder
from ragdoll_amd import DECISION_DTYPE, LEVELS_PER_MB, StreamAnalyzer, analyse_frames
from ragdoll_amd.hip import Hip
from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed, fixture_frame

pytestmark = pytest.mark.gpu


def scene(name, n, seed_extra=0):
    kw = dict(CONFIGS[name])
    w, h = kw.pop("width"), kw.pop("height")
    g = SceneGenerator(w, h, seed=config_seed({"cif": 0, "720p": 1, "1080p": 2}[name], seed_extra), **kw)
    return w, h, [g.next_frame() for _ in range(n)]


def oracle_run(frames, w, h, qp, profile, me_range=16, speed="balanced"):
    ora = OracleEncoder(w, h, qp, profile=profile, keyint=1 << 20, me_range=me_range, speed=speed)
    out = []
    for i, f in enumerate(frames):
        ora.encode(*f)
        if i == 0:
            idr = ora.recon(deblocked=True)
        else:
            out.append((ora.decisions(), ora.levels()))
    return idr, out, ora.recon(deblocked=True)


def run_pipeline(streams, lag, ring, calls=1, groups=0, label="", persistent=False, graph=False):
    """streams: list of (frames, w, h, qp, profile[, me_range[, speed]]). Runs frames 1.. of every stream
    through rgh_analyse_frames (split into `calls` consecutive calls; the per-step launch schedule,
    or one persistent launch per call) and checks them against the oracle."""
    hip = Hip()
    ans, refs, bufs = [], [], []
    try:
        for st in streams:
            frames, w, h, qp, profile = st[:5]
            me_range = st[5] if len(st) > 5 else 16
            speed = st[6] if len(st) > 6 else "balanced"
            idr, want, final = oracle_run(frames, w, h, qp, profile, me_range, speed)
            a = StreamAnalyzer(w, h, qp, profile=profile, me_range=me_range, keep_undeblocked=False, ring=ring,
                               speed=speed)
            a.set_reference(*idr)
            n = a.mb_w * a.mb_h
            srcs, decs, lvs = [], [], []
            for y, u, v in frames[1:]:
                dy, du, dv = hip.upload(y), hip.upload(u), hip.upload(v)
                bufs += [dy, du, dv]
                srcs.append((dy, du, dv, y.shape[1], u.shape[1], v.shape[1]))
                d, l = hip.alloc(n * 32), hip.alloc(n * LEVELS_PER_MB * 2)
                hip.memset(d, 0xEE, n * 32)
                bufs += [d, l]
                decs.append(d)
                lvs.append(l)
            ans.append(a)
            refs.append((want, final, srcs, decs, lvs, n))
        nf = len(streams[0][0]) - 1
        cuts = [round(i * nf / calls) for i in range(calls + 1)]
        for c in range(calls):
            sl = slice(cuts[c], cuts[c + 1])
            analyse_frames(ans, [r[2][sl] for r in refs], [r[3][sl] for r in refs], [r[4][sl] for r in refs], lag=lag,
                           groups=groups, persistent=persistent, graph=graph)
        for s, (a, (want, final, srcs, decs, lvs, n)) in enumerate(zip(ans, refs)):
            for k in range(nf):
                dec = hip.download(decs[k], np.zeros(n, DECISION_DTYPE))
                lv = hip.download(lvs[k], np.zeros((n, LEVELS_PER_MB), np.int16))
                odec, olv = want[k]
                bad = np.nonzero(dec.view(np.uint8).reshape(-1, 32) != odec.view(np.uint8).reshape(-1, 32))[0]
                assert len(bad) == 0, "%s stream %d frame %d: decision of MB %d differs\n gpu %s\n ora %s" % (
                    label, s, k + 1, bad[0], dec[bad[0]], odec[bad[0]])
                badl = np.nonzero((lv != olv).any(axis=1))[0]
                assert len(badl) == 0, "%s stream %d frame %d: levels of MB %d differ" % (label, s, k + 1, badl[0])
            _, _, rec = a.outputs(which_recon=1, want_levels=False)
            for p, (g, o) in enumerate(zip(rec, final)):
                assert np.array_equal(g, o), "%s stream %d: final deblocked plane %d differs" % (label, s, p)
    finally:
        for a in ans:
            a.close()
        for b in bufs:
            hip.free(b)


# (12, 2): frames fall far enough behind to leave the planned launch range and rejoin later.
@pytest.mark.parametrize("lag,ring", [(20, 8), (3, 8), (1, 4), (1, 2), (12, 2)])
def test_cif_pipeline(lag, ring):
    w, h, frames = scene("cif", 8)
    print("pipeline cif lag %d ring %d" % (lag, ring), flush=True)
    run_pipeline([(frames, w, h, 28, "baseline")], lag, ring, label="cif lag %d ring %d" % (lag, ring))


@pytest.mark.parametrize("groups", [1, 2, 3])
def test_multi_stream_two_calls(groups):
    # Three streams of different sizes and profiles, split over two calls, in one schedule or
    # in concurrent schedules on separate streams.
    print("pipeline multi-stream groups %d" % groups, flush=True)
    w, h, cif = scene("cif", 6, 3)
    fx = [fixture_frame(160, 90, i) for i in range(6)]
    w7, h7, f720 = scene("720p", 6, 5)
    run_pipeline([(cif, w, h, 26, "high"), (fx, 160, 90, 20, "baseline"), (f720, w7, h7, 30, "main")],
                 lag=6, ring=6, calls=2, groups=groups, label="multi groups %d" % groups)


def test_batched_uneven_streams():
    # The batched configuration's shape at reduced length: one stream of each size class
    # (640x360, 1280x720, 1920x1080) with its seeded QP and width-scaled search range, High
    # decisions, all in one schedule of 2 concurrent groups.
    from ragdoll_amd.synth import batched_stream
    print("pipeline batched", flush=True)
    streams = []
    for sid in (3, 12, 17):
        kw, seed, qp, me_range = batched_stream(sid)
        kw = dict(kw)
        w, h = kw.pop("width"), kw.pop("height")
        g = SceneGenerator(w, h, seed=seed, **kw)
        streams.append(([g.next_frame() for _ in range(4)], w, h, qp, "high", me_range))
    run_pipeline(streams, lag=12, ring=4, groups=2, label="batched")


# RGH_FRAMES_GRAPH: the planned launches captured once and replayed. Eight P frames over four
# calls of two frames each: the first call captures, the next three replay the graph (between
# them each call uploads its descriptors, finishes deferred frames with launches of its own and
# synchronises). Lag 1 defers most macroblocks, so completion launches follow every replay; two
# groups capture across two streams. Every frame is checked against the oracle.
@pytest.mark.parametrize("lag,ring,groups", [(20, 8, 1), (1, 4, 1), (6, 6, 2)])
def test_graph_replays(lag, ring, groups):
    w, h, cif = scene("cif", 9, 31)
    fx = [fixture_frame(160, 90, i) for i in range(9)]
    print("graph lag %d ring %d groups %d" % (lag, ring, groups), flush=True)
    run_pipeline([(cif, w, h, 26, "high"), (fx, 160, 90, 20, "baseline")], lag, ring, calls=4, groups=groups,
                 graph=True, label="graph lag %d ring %d groups %d" % (lag, ring, groups))


def test_graph_recaptures_when_the_launches_change():
    # Calls of 3, 2 and 3 frames: each differs from the graph kept by the call before it.
    w, h, cif = scene("cif", 9, 32)
    print("graph recapture", flush=True)
    run_pipeline([(cif, w, h, 30, "main")], 8, 6, calls=3, graph=True, label="graph recapture")


def test_graph_is_refused_by_the_noalloc_entry_point():
    from ragdoll_amd.abi import RghError
    w, h, frames = scene("cif", 2, 33)
    idr, _, _ = oracle_run(frames[:1], w, h, 28, "baseline")
    hip = Hip()
    a = StreamAnalyzer(w, h, 28, keep_undeblocked=False, ring=4)
    bufs = []
    try:
        a.set_reference(*idr)
        n = a.mb_w * a.mb_h
        y, u, v = frames[1]
        planes = [hip.upload(p) for p in (y, u, v)]
        d, l = hip.alloc(n * 32), hip.alloc(n * LEVELS_PER_MB * 2)
        bufs += planes + [d, l]
        src = tuple(planes) + (y.shape[1], u.shape[1], v.shape[1])
        with pytest.raises(RghError, match="RGH_FRAMES_GRAPH"):
            analyse_frames([a], [[src]], [[d]], [[l]], noalloc=True, graph=True)
    finally:
        a.close()
        for b in bufs:
            hip.free(b)


# The persistent schedule (RGH_FRAMES_PERSISTENT): one launch whose items wait for each other.
# Small lags make macroblocks need reference items planned after them (they defer, and the
# per-step schedule finishes the call); ring 2 exercises the picture-slot gate.
@pytest.mark.parametrize("lag,ring", [(20, 8), (3, 8), (1, 2), (12, 2)])
def test_cif_persistent(lag, ring):
    w, h, frames = scene("cif", 8)
    print("persistent cif lag %d ring %d" % (lag, ring), flush=True)
    run_pipeline([(frames, w, h, 28, "baseline")], lag, ring, label="persistent cif lag %d ring %d" % (lag, ring),
                 persistent=True)


def test_multi_stream_persistent():
    print("persistent multi-stream", flush=True)
    w, h, cif = scene("cif", 6, 3)
    fx = [fixture_frame(160, 90, i) for i in range(6)]
    w7, h7, f720 = scene("720p", 6, 5)
    run_pipeline([(cif, w, h, 26, "high"), (fx, 160, 90, 20, "baseline"), (f720, w7, h7, 30, "main")],
                 lag=6, ring=6, calls=2, label="persistent multi", persistent=True)


@pytest.mark.parametrize("env", [{"RGH_PERSIST_JITTER": "48"}, {"RGH_PERSIST_SPIN_LIMIT": "0"}],
                         ids=["jitter", "gives-up"])
def test_persistent_under_stress(env, monkeypatch):
    # Jitter: items finish far out of ticket order. A wait bound of zero: the first wait gives up
    # and the per-step schedule finishes the call. Both must give the oracle's results.
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    w, h, cif = scene("cif", 6, 11)
    w7, h7, f720 = scene("720p", 4, 12)
    print("persistent stress %s" % env, flush=True)
    run_pipeline([(cif, w, h, 30, "high"), (cif, w, h, 24, "baseline")], lag=16, ring=6,
                 label="persistent %s" % env, persistent=True)
    run_pipeline([(f720, w7, h7, 26, "high", 32)], lag=20, ring=4, label="persistent 720p %s" % env, persistent=True)


def test_failed_call_leaves_context_without_reference(monkeypatch):
    # A call that fails after its launches started (forced from the host) leaves its contexts
    # without a reference: the next P frame and the outputs are refused until a new reference.
    from ragdoll_amd.abi import RghError
    w, h, frames = scene("cif", 3, 21)
    idr, _, _ = oracle_run(frames[:1], w, h, 28, "baseline")
    hip = Hip()
    a = StreamAnalyzer(w, h, 28, keep_undeblocked=False, ring=4)
    bufs = []
    try:
        a.set_reference(*idr)
        n = a.mb_w * a.mb_h
        y, u, v = frames[1]
        dy, du, dv = hip.upload(y), hip.upload(u), hip.upload(v)
        d, l = hip.alloc(n * 32), hip.alloc(n * LEVELS_PER_MB * 2)
        bufs = [dy, du, dv, d, l]
        monkeypatch.setenv("RGH_TEST_FAIL_AFTER_LAUNCH", "1")
        with pytest.raises(RghError) as e:
            analyse_frames([a], [[(dy, du, dv, y.shape[1], u.shape[1], v.shape[1])]], [[d]], [[l]])
        assert e.value.status == 5  # RGH_ERR_DEVICE
        monkeypatch.delenv("RGH_TEST_FAIL_AFTER_LAUNCH")
        with pytest.raises(RghError) as e:
            a.analyse(*frames[2])
        assert e.value.status == 4  # RGH_ERR_NO_REFERENCE
        with pytest.raises(RghError) as e:
            a.outputs()
        assert e.value.status == 4
        # A new reference makes it usable again.
        a.set_reference(*idr)
        a.analyse(*frames[1])
    finally:
        a.close()
        for b in bufs:
            hip.free(b)


def run_gops(streams, keyint, lag, ring, groups=0, label=""):
    """streams: (frames, w, h, qp, profile, me_range, speed). Every stream's frames run in one
    rgh_analyse_frames call with an IDR every `keyint` frames (RGH_FRAMES_KEYINT); each GOP must
    equal the oracle coding that GOP alone from its IDR (a fresh encoder per GOP), and the last
    deblocked picture the oracle's last."""
    hip = Hip()
    ans, refs, bufs = [], [], []
    try:
        for frames, w, h, qp, profile, me_range, speed in streams:
            want, final = [], None
            for g0 in range(0, len(frames), keyint):
                ora = OracleEncoder(w, h, qp, profile=profile, keyint=1 << 20, me_range=me_range, speed=speed)
                for f in frames[g0:g0 + keyint]:
                    ora.encode(*f)
                    want.append((ora.decisions(), ora.levels()))
                final = ora.recon(deblocked=True)
            a = StreamAnalyzer(w, h, qp, profile=profile, me_range=me_range, keep_undeblocked=False, ring=ring,
                               speed=speed)
            n = a.mb_w * a.mb_h
            srcs, decs, lvs = [], [], []
            for y, u, v in frames:
                dy, du, dv = hip.upload(y), hip.upload(u), hip.upload(v)
                bufs += [dy, du, dv]
                srcs.append((dy, du, dv, y.shape[1], u.shape[1], v.shape[1]))
                d, l = hip.alloc(n * 32), hip.alloc(n * LEVELS_PER_MB * 2)
                hip.memset(d, 0xEE, n * 32)
                bufs += [d, l]
                decs.append(d)
                lvs.append(l)
            ans.append(a)
            refs.append((want, final, srcs, decs, lvs, n))
        print("keyint %d %s" % (keyint, label), flush=True)
        analyse_frames(ans, [r[2] for r in refs], [r[3] for r in refs], [r[4] for r in refs], lag=lag, groups=groups,
                       first_idr=True, keyint=keyint)
        for s, (a, (want, final, srcs, decs, lvs, n)) in enumerate(zip(ans, refs)):
            for k in range(len(want)):
                dec = hip.download(decs[k], np.zeros(n, DECISION_DTYPE))
                lv = hip.download(lvs[k], np.zeros((n, LEVELS_PER_MB), np.int16))
                odec, olv = want[k]
                bad = np.nonzero(dec.view(np.uint8).reshape(-1, 32) != odec.view(np.uint8).reshape(-1, 32))[0]
                assert len(bad) == 0, "%s stream %d frame %d: decision of MB %d differs\n gpu %s\n ora %s" % (
                    label, s, k, bad[0], dec[bad[0]], odec[bad[0]])
                badl = np.nonzero((lv != olv).any(axis=1))[0]
                assert len(badl) == 0, "%s stream %d frame %d: levels of MB %d differ" % (label, s, k, badl[0])
            _, _, rec = a.outputs(which_recon=1, want_levels=False)
            for p, (g, o) in enumerate(zip(rec, final)):
                assert np.array_equal(g, o), "%s stream %d: final deblocked plane %d differs" % (label, s, p)
    finally:
        for a in ans:
            a.close()
        for b in bufs:
            hip.free(b)


@pytest.mark.parametrize("lag,groups", [(20, 1), (3, 2)])
def test_gops_back_to_back(lag, groups):
    # Three GOPs of 4 frames per stream in one call (IDRs at frames 0, 4, 8), two streams of
    # different profiles; a small lag makes frames of consecutive GOPs share launches with
    # deferrals, so the mixed IDR + P kernel runs IDR frames in the middle of the schedule.
    w, h, cif = scene("cif", 12, 21)
    fx = [fixture_frame(160, 90, i) for i in range(12)]
    run_gops([(cif, w, h, 26, "high", 16, "balanced"), (fx, 160, 90, 20, "baseline", 16, "balanced")], keyint=4,
             lag=lag, ring=6, groups=groups, label="lag %d groups %d" % (lag, groups))


@pytest.mark.parametrize("speed,profile", [("fast", "high"), ("partitions", "main"), ("compact", "baseline")])
def test_gops_back_to_back_presets(speed, profile):
    # The other presets' kernels across mid-call IDRs (Compact runs its IDR and P frames as separate
    # kernels within a launch).
    w, h, cif = scene("cif", 9, 22)
    run_gops([(cif, w, h, 24, profile, 32, speed)], keyint=3, lag=6, ring=5, label=speed)
