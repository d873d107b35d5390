"""The measured configurations at their stated sizes, GPU vs the CPU oracle bit for bit.

* 1080p High at QP 24/30/36 with a +-64 search (64 diamond iterations): four frames of the
  headline content whose last is a scene cut (new background), single-frame calls and the
  pipelined schedule. Walks longer than 16 iterations stage new search regions as they go.
* 1280x720 High at QP 26 with a +-32 search (32 iterations): the 720p content with its global pan.
"""
import pytest

from parity import run_parity
from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed
from test_gpu_pipeline import run_pipeline

pytestmark = pytest.mark.gpu

# Scene cuts of the headline content (seed config_seed(2)) fall at frames 6, 11, 18, 23, ...:
# frames 3..6 are an IDR, two P frames and a cut frame.
CUT_SLICE = slice(3, 7)


def headline_frames(n=7):
    kw = dict(CONFIGS["1080p"])
    w, h = kw.pop("width"), kw.pop("height")
    g = SceneGenerator(w, h, seed=config_seed(2), **kw)
    return w, h, [g.next_frame() for _ in range(n)]


@pytest.mark.parametrize("qp", [24, 30, 36])
def test_1080p_scene_cut_me64(qp):
    w, h, frames = headline_frames()
    s = run_parity(frames[CUT_SLICE], w, h, qp, profile="high", me_range=64, label="1080p qp%d me64" % qp)
    assert s["frames"] == 3
    assert s["intra"] >= 8000  # the cut frame codes (nearly) every macroblock intra


def test_1080p_pipelined_me64():
    # Two streams (QP 24 and 36) of the same four frames through one pipelined schedule at a
    # small lag, so that macroblocks whose walks re-stage defer while the reference is in flight.
    w, h, frames = headline_frames()
    f = frames[CUT_SLICE]
    print("pipeline 1080p me64", flush=True)
    run_pipeline([(f, w, h, 24, "high", 64), (f, w, h, 36, "high", 64)], lag=6, ring=4, groups=2,
                 label="1080p me64")


def test_1080p_persistent_me64():
    # The same streams through the persistent schedule.
    w, h, frames = headline_frames()
    f = frames[CUT_SLICE]
    print("persistent 1080p me64", flush=True)
    run_pipeline([(f, w, h, 24, "high", 64), (f, w, h, 30, "high", 64), (f, w, h, 36, "high", 64)], lag=20, ring=5,
                 label="persistent 1080p me64", persistent=True)


def test_720p_high_qp26_me32():
    kw = dict(CONFIGS["720p"])
    w, h = kw.pop("width"), kw.pop("height")
    g = SceneGenerator(w, h, seed=config_seed(1, 26), **kw)
    frames = [g.next_frame() for _ in range(5)]
    s = run_parity(frames, w, h, 26, profile="high", me_range=32, label="720p qp26 me32")
    assert s["frames"] == 4 and s["inter"] > 0
