"""The autotuner changes speed, not results: a run after autotune() (which
times candidate skins, cell sizes and launch shapes from a saved point and
returns the solver to it) equals a plain run bit for bit, for a
non-periodic system (where the cell size is tuned too) and a periodic one."""
import numpy as np
import pytest

from fundem_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("periodic", [(False, False), (True, True)])
def test_autotune_keeps_results(periodic):
    from fundem_amd import Solver
    scn = W.cluster(n=27, periodic=periodic, speed=1.0)
    plain = Solver()
    W.apply(scn, plain)
    plain.step(150)
    ref = plain.state()
    tuned = Solver()
    W.apply(scn, tuned)
    choice = tuned.autotune(trial_steps=6)
    for line in choice["trials"]:
        print(line)
    print({k: v for k, v in choice.items() if k != "trials"})
    assert len(choice["trials"]) >= 16
    assert choice["law_block"] in (64, 128, 256) and choice["sweep_waves"] in (1, 2, 4)
    tuned.step(150)
    got = tuned.state()
    assert tuned.status() == 0
    for key in ("x", "v", "w", "q", "F", "T"):
        assert np.array_equal(got[key], ref[key]), key
