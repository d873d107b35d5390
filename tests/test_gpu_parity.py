"""GPU P-frame analysis vs the CPU oracle, bit for bit (decisions, levels, reconstructions)."""
import itertools

import numpy as np
import pytest

from parity import run_parity
from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed, fixture_frame

pytestmark = pytest.mark.gpu


def fixture_frames(w, h, n):
    return [fixture_frame(w, h, i) for i in range(n)]


def scene_frames(w, h, n, seed, **kw):
    g = SceneGenerator(w, h, seed=seed, **kw)
    return [g.next_frame() for _ in range(n)]


@pytest.mark.parametrize("w,h", [(160, 90), (320, 180)])
def test_golden_clip_content(w, h):
    # The content and settings of the reference's golden clips (QP 20, Baseline).
    s = run_parity(fixture_frames(w, h, 15), w, h, 20, label="fixture %dx%d" % (w, h))
    assert s["frames"] == 14


def test_cif_qp28_baseline():
    # Plain CPU-runnable configuration: 352x288, QP 28, Baseline, 6 objects, |v| <= 8, sigma 2.
    kw = dict(CONFIGS["cif"])
    w, h = kw.pop("width"), kw.pop("height")
    s = run_parity(scene_frames(w, h, 8, config_seed(0), **kw), w, h, 28, label="cif")
    assert s["inter"] > 0


@pytest.mark.parametrize("qp", [10, 22, 30, 40, 51])
def test_cif_qp_sweep(qp):
    kw = dict(CONFIGS["cif"])
    w, h = kw.pop("width"), kw.pop("height")
    run_parity(scene_frames(w, h, 4, config_seed(0, qp), **kw), w, h, qp, label="cif qp%d" % qp)


@pytest.mark.parametrize("profile", ["main", "high"])
def test_cif_cabac_profiles(profile):
    kw = dict(CONFIGS["cif"])
    w, h = kw.pop("width"), kw.pop("height")
    s = run_parity(scene_frames(w, h, 5, config_seed(0, 7), **kw), w, h, 26, profile=profile,
                   label="cif " + profile)
    assert s["frames"] == 4


def test_720p_high():
    kw = dict(CONFIGS["720p"])
    w, h = kw.pop("width"), kw.pop("height")
    run_parity(scene_frames(w, h, 3, config_seed(1), **kw), w, h, 26, profile="high", label="720p")


@pytest.mark.parametrize("w,h", [(64, 48), (50, 30), (18, 18), (130, 66)])
def test_odd_sizes(w, h):
    run_parity(scene_frames(w, h, 4, 99, n_objects=2, max_speed=6, noise_sigma=2), w, h, 24,
               label="%dx%d" % (w, h))


@pytest.mark.parametrize("me_range", [1, 4, 32])
def test_me_range_extension(me_range):
    kw = dict(CONFIGS["cif"])
    w, h = kw.pop("width"), kw.pop("height")
    run_parity(scene_frames(w, h, 3, config_seed(0, 100 + me_range), **kw), w, h, 28, me_range=me_range,
               label="range %d" % me_range)
