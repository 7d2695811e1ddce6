"""The learned initial guess inside the time step: pairs recorded by the
solver itself, training by the project's kernels, then the MAC solve started
from the network's guess.  Two identical simulations run in lockstep, one
with the learned guess and one without; the fields must agree to the solve
tolerance, the fallback rule must be applied at every solve, and the test
reports what the network gains (MAC V-cycles per step, initial residual)."""
import sys

import numpy as np
import pytest

from vdropimpact_amd.params import PhysicalConfig

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def mac_cycles(sim):
    return sim.stats().total_cycles[0]


def test_gpu_learned_guess_trains_and_keeps_the_solution(vdi):
    p = PhysicalConfig(m=256, n=128).params()
    a, b = vdi.Fluid2D(p), vdi.Fluid2D(p)
    a.init_fields()
    b.init_fields()
    a.guess_attach(capacity=128, seed=5)
    a.guess_mode(1)
    n_cap = 130
    for _ in range(n_cap):
        a.step_forward(p.dt_reg)
        b.step_forward(p.dt_reg)
    st = a.guess_stats()
    say("pairs recorded", st["pairs"])
    assert st["pairs"] == 128
    loss = a.guess_train(epochs=25, batch=4, lr=1e-3)
    st = a.guess_stats()
    say("training loss (scaled target, last epoch)", loss, "kappa", st["kappa"])
    assert np.isfinite(loss) and loss < 1.0
    a.guess_mode(2)
    # direct path for both: every solve runs to the tolerance with exact
    # cycle counts (no captured budgets, no redone steps)
    a.set_graphs(False)
    b.set_graphs(False)
    a.reset_stats()
    b.reset_stats()
    K = 40
    for _ in range(K):
        a.step_forward(p.dt_reg)
        b.step_forward(p.dt_reg)
    st = a.guess_stats()
    ca, cb = mac_cycles(a) / K, mac_cycles(b) / K
    say(f"GAIN of the learned guess: MAC V-cycles per step {ca:.3f} (learned) vs {cb:.3f} (extrapolated), "
        f"{cb - ca:+.3f} saved; network guess kept {st['accepted']} of {K} times")
    # the fallback rule ran at every solve, and the solution is the same to
    # the solve tolerance whichever guess the solve started from
    assert st["accepted"] + st["rejected"] == K
    # The network does not gain here (its guesses have the smaller initial
    # residual but leave low-frequency error that costs cycles: measured
    # 3.900 against 3.725 MAC cycles per step).  The bound fails a network
    # that costs half a cycle per step or more.
    assert ca <= cb + 0.5
    for aid in (vdi.U, vdi.V, vdi.P, vdi.PG, vdi.H):
        e = rel(a.get(aid), b.get(aid))
        assert e < 1e-8, (aid, e)
    a.close()
    b.close()
