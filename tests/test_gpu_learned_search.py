"""The learned-proposal search on the GPU (rgh_analyse_frames_learned) with the shipped network.

What the network gains over the same search without it: the learned form seeds the integer
search with the predictor and four proposed cells and stops the diamond early; with proposals
that carry no information (every cell the zero vector, uniform mode probabilities) the same
search falls back to the plain search far more often and, where it does not, ends at higher
motion costs. The trained network must beat that on held-out content (seeds the training never
saw) on both counts, and its rate at each QP must stay within 3% of the plain search's at no
worse PSNR (the benchmark's BD-rate sample measures the full curve).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 640, 368


@pytest.fixture(scope="module")
def frames():
    from ragdoll_amd.synth import SceneGenerator, config_seed
    g = SceneGenerator(W, H, n_objects=8, max_speed=24, noise_sigma=3, seed=config_seed(9, 1))
    return [g.next_frame() for _ in range(12)]


class NullProposer:
    """Proposals with no information: every cell (0, 0), uniform mode probabilities."""

    def __init__(self, hip, n):
        from ragdoll_amd.learned import PROPOSAL_DTYPE
        p = np.zeros(n, PROPOSAL_DTYPE)
        p["mode_prob"] = 85
        self.hip = hip
        self.buf = hip.upload(p)

    def run(self, pictures):
        return [self.buf for _ in pictures]

    def close(self):
        self.hip.free(self.buf)


@pytest.fixture(scope="module")
def results(frames):
    from ragdoll_amd.hip import Hip
    from ragdoll_amd.learned import Proposer, load_trained
    from ragdoll_amd.rd import code_gop
    net = load_trained(max_batch=1024)
    n = (W // 16) * (H // 16)
    prop = Proposer(net, n)
    null = NullProposer(Hip(), n)
    cfg = dict(iters=4, skip_gate=0, intra_gate=0, fallback_pct=105)
    out = {}
    for qp in (27, 32):
        print("coding QP %d" % qp, flush=True)
        out[qp] = {"plain": code_gop(frames, W, H, qp),
                   "learned": code_gop(frames, W, H, qp, proposer=prop, learned=cfg),
                   "null": code_gop(frames, W, H, qp, proposer=null, learned=cfg)}
        for k, r in out[qp].items():
            print(qp, k, r["bytes"], round(r["psnr_y"], 3), "plain search", r["plain_search"],
                  "inter cost/MB %.1f" % r["inter_cost_per_mb"], r["mb_types"][:4], flush=True)
    prop.close()
    null.close()
    net.close()
    return out


def test_learned_search_is_deterministic(frames):
    from ragdoll_amd.learned import Proposer, load_trained
    from ragdoll_amd.rd import code_gop
    net = load_trained(max_batch=1024)
    prop = Proposer(net, (W // 16) * (H // 16))
    a = code_gop(frames[:4], W, H, 30, proposer=prop, learned=dict(iters=2))
    b = code_gop(frames[:4], W, H, 30, proposer=prop, learned=dict(iters=2))
    prop.close()
    net.close()
    assert a["frame_bytes"] == b["frame_bytes"] and a["frame_psnr_y"] == b["frame_psnr_y"]


def test_trained_network_beats_proposals_without_information(results):
    for qp, r in results.items():
        assert r["learned"]["plain_search"] < 0.5 * r["null"]["plain_search"], qp
        assert r["learned"]["inter_cost_per_mb"] <= r["null"]["inter_cost_per_mb"], qp


def test_learned_rate_stays_near_the_plain_search(results):
    for qp, r in results.items():
        assert r["learned"]["bytes"] <= 1.03 * r["plain"]["bytes"], qp
        assert r["learned"]["psnr_y"] >= r["plain"]["psnr_y"] - 0.05, qp
