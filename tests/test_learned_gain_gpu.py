"""What the trained networks gain, on held-out problems (GPU).

The dam-break network (scihpc_amd/weights/unet_dambreak_1024x512_level3.npz)
was trained by python -m scihpc_amd.learned --data dambreak on snapshots of
steps 10-1040 of the solver's own 1024x512 dam break; the fixture
(dambreak_heldout_1024x512.npz) holds the right-hand side and density of
two later steps the training never saw: step 655 is an ordinary step, and
step 1310 is one of the hard steps before the flow breaks down, where the
mixed FCG needs about 100 iterations to 1e-8.  Started from the network's
guess (use_learned = 2) that solve reaches the same FP64 tolerance in
about two thirds of the iterations and less time; on the ordinary step the
guess saves at most a couple of iterations.  Every solve is checked to
reach the tolerance with the FP64 residual b - A p.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from scihpc_amd import PREC_MIXED, Multigrid
from scihpc_amd import learned as LRN

WEIGHTS = os.path.join(LRN.WEIGHTS_DIR, "unet_dambreak_1024x512_level3.npz")


def _best(mg, reps=3, **kw):
    best = None
    for _ in range(reps):
        st = mg.solve(rel_tol=1e-8, max_cycles=200, precision=PREC_MIXED, accel=1, **kw)
        assert st["converged"] and st["rel_residual"] <= 1e-8, st
        best = st if best is None or st["ms"] < best["ms"] else best
    return best


def test_dambreak_guess_gains_on_hard_step():
    assert os.path.exists(WEIGHTS) and os.path.exists(LRN.DAMBREAK_FIXTURE)
    net, meta = LRN.load_net(WEIGHTS)
    probs = LRN.dambreak_fixture_problems()
    mg = Multigrid(2, 1024, 512, 1, 4 / 1024, 2 / 512, 1.0, "neumann")
    mg.attach_learned(net, meta["level"])
    out = {}
    for step, w in sorted(probs.items()):
        mg.set_density(w["rho"])
        mg.set_rhs(w["rhs"])
        plain = _best(mg)
        forced = _best(mg, use_learned=2)
        judged = _best(mg, use_learned=3)
        print(f"dam-break step {step}: plain {plain['cycles']} it {plain['ms']:.2f} ms; from the guess "
              f"{forced['cycles']} it {forced['ms']:.2f} ms; judged {judged['cycles']} it {judged['ms']:.2f} ms "
              f"(kept {judged['learned_accepted']})", flush=True)
        out[step] = (plain, forced, judged)
        # a judged guess never costs iterations
        assert judged["cycles"] <= plain["cycles"]
    plain, forced, _ = out[1310]
    assert plain["cycles"] >= 50  # the hard step
    assert forced["cycles"] <= 0.8 * plain["cycles"]
    assert forced["ms"] < plain["ms"]
    plain, forced, _ = out[655]
    assert forced["cycles"] <= plain["cycles"] + 1
