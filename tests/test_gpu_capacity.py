"""Capacity overflows are reported, not hidden: a solver whose candidate-pair
or contact-record capacity is too small for a contact-dense packing sets the
device status word (FDM_STATUS_PAIR_OVERFLOW or FDM_STATUS_GHOST_OVERFLOW;
FDM_STATUS_RECORD_OVERFLOW or FDM_STATUS_STAGING_OVERFLOW), its step raises,
and the kernels keep every write inside the buffers (the run ends normally
and the solver can still be read and closed). The reference has fixed-size
arrays and no such check.

The packing is a 32^3 slice of the pour compressed until its layers touch
(well over 100k particle node contacts). With one record per particle the
solver sizes a set to 32768 + 64 x 1024 records (each of the 64 record pools
keeps 1024 spare), and with one candidate pair per particle to 32768 pairs,
which a 1 mm candidate-list skin overflows."""
import numpy as np
import pytest

from fundem_amd import workloads as W

pytestmark = pytest.mark.gpu

PAIR, RECORD, GHOST, STAGING = 1, 2, 8, 16


@pytest.fixture(scope="module")
def dense():
    from fundem_amd import Solver

    def make():
        s = Solver()
        s.set_capacity(12, 64)
        return s

    scn, done, c = W.compressed_pour(W.pour((32, 32, 32)), make)
    print(f"32^3 slice compressed for {done} steps: {c} particle node contacts", flush=True)
    assert c > 32768 + 64 * 1024
    return scn


@pytest.mark.parametrize("caps,skin,bits", [((12, 1), -1.0, RECORD | STAGING),
                                             ((1, 64), 1e-3, PAIR | GHOST)],
                         ids=["records", "pairs"])
def test_capacity_overflow_is_reported(dense, caps, skin, bits):
    from fundem_amd import Solver
    s = Solver()
    s.set_capacity(*caps)
    s.set_skin(skin)  # a 1 mm skin keeps several candidate pairs per particle
    with pytest.raises(RuntimeError, match="capacity exceeded"):
        W.apply(dense, s)
        s.step(20)
    status = s.status()
    print(f"capacities {caps}: status {status}", flush=True)
    assert status & bits, status
    assert s.state()["x"].shape == (dense.n, 3)
    s.close()
    # the same packing with enough room runs clean
    g = Solver()
    g.set_capacity(12, 64)
    W.apply(dense, g)
    g.step(20)
    assert g.status() == 0 and np.isfinite(g.state()["x"]).all()
    g.close()
