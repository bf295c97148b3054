"""GPU tests of the correction stage (ngap_set_stage) and of the learned
initial guess on the correlated right-hand-side sequence (128^3 periodic).

The stage never changes what a solve converges to:
- every solve meets the same true-residual tolerance with the exact operator;
- solutions agree with the stage-free solve to 1e-8 (mean-free relative L2).
"""
import numpy as np
import pytest
import torch

from nga2_amd import _lib, learned
from nga2_amd.unet import UNet

pytestmark = pytest.mark.gpu

RCVG = 1e-10


@pytest.fixture(scope="module")
def seq_solver():
    S = learned.SequenceSolver(n=128, rcvg=RCVG)
    yield S
    S.close()


def _mean_free_rel(a, b):
    a = a - a.mean()
    b = b - b.mean()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def test_stage_with_untrained_net_keeps_the_solution(seq_solver):
    print("\n[stage: untrained net as guess and as PCG stage]", flush=True)
    S = seq_solver
    net = UNet((128, 128, 128), seed=11)
    net.set_kappa(1e-4)
    x0 = torch.zeros(S.N, dtype=torch.float64, device=S.dev)
    b = S.rhs(3)
    xa, sa = S.solve(b, x0)
    assert sa["true_rerr"] <= RCVG and sa["stage_calls"] == 0
    for flags, iters in [(_lib.STAGE_GUESS, 0), (_lib.STAGE_PRECOND, 3), (_lib.STAGE_GUESS | _lib.STAGE_PRECOND, 2)]:
        xb, sb = S.solve(b, x0, stage=net.stage(), flags=flags, precond_iters=iters)
        print(f"  flags {flags} iters {iters}: it {sb['it']} (plain {sa['it']}), stage calls {sb['stage_calls']}, "
              f"fallback {sb['fallback']}, true {sb['true_rerr']:.2e}", flush=True)
        assert sb["status"] == 0 and sb["true_rerr"] <= RCVG
        assert sb["stage_calls"] >= (1 if flags & _lib.STAGE_GUESS else 0) + (min(iters, sb["it"]) > 0)
        assert _mean_free_rel(xb, xa) <= 1e-8
    # removed again
    xc, sc = S.solve(b, x0)
    assert sc["stage_calls"] == 0 and sc["it"] == sa["it"]
    net.close()


def test_stage_shape_mismatch_is_an_error(seq_solver):
    S = seq_solver
    net = UNet((128, 64, 64), seed=1)
    b = S.rhs(0)
    x0 = torch.zeros(S.N, dtype=torch.float64, device=S.dev)
    with pytest.raises(_lib.NgapError) as ei:
        S.solve(b, x0, stage=net.stage(), flags=_lib.STAGE_GUESS)
    assert "stage" in str(ei.value)
    net.close()


def test_learned_guess_beats_previous_step_guess(seq_solver):
    """Trained on 48 solves of the sequence, the learned guess starts the next
    8 solves from a far smaller residual than the previous solution does, and
    they need fewer iterations.

    - The temporal predictor (m = 6) is fitted on the solutions.
    - The network stage is trained with the residual loss on the residuals of
      the extrapolated guesses.
    - The network's correction lowers the residual further, or is rejected by
      the stage."""
    print("\n[learned guess: train 48 steps, evaluate 8]", flush=True)
    S = seq_solver
    sols, st = learned.collect(S, range(48))
    alpha = learned.fit_predictor(sols, 6)
    pairs = learned.stage_pairs(S, sols, alpha, range(7, 48))
    kappa = learned.fit_kappa(pairs)
    net = UNet((128, 128, 128), seed=0)
    hist = learned.train(net, pairs, kappa, S=S, epochs=4, lr=5e-4, log=lambda m: print("  " + m, flush=True))
    del pairs
    recs = learned.evaluate(S, net, alpha, sols, range(48, 56))
    s = learned.summary(recs)
    print("  " + ", ".join(f"{k} {v:.3g}" if isinstance(v, float) else f"{k} {v}" for k, v in s.items()), flush=True)
    assert hist[-1] < hist[0]
    assert max(s["max_true_prev"], s["max_true_extrap"], s["max_true_learned"]) <= RCVG
    assert s["max_solution_diff"] <= 1e-8
    assert s["r0_extrap"] < 1e-2 * s["r0_prev"]
    assert s["r0_learned"] < s["r0_extrap"]  # the network's correction lowers the residual
    assert s["it_extrap"] <= s["it_prev"] - 3
    assert s["it_learned"] <= s["it_extrap"] + 0.5
    net.close()
