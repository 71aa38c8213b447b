"""The 1M-particle pour at its full size (100 x 100 x 100 superellipsoids,
16 templates, 32^3 SDF, 1000 surface nodes, periodic in x and y with a
floor wall), checked through properties that hold at any size. The oracle
would need hours for one step of it; the small pours and clusters of
test_gpu_parity.py compare every bit with it.

* Candidate-list reuse is exact: with the default skin the lists are kept
  for many steps, and the states equal those of a run that rebuilds every
  list every step (skin 0), bit for bit.
* Launches are deterministic: a second run, and a run replaying captured
  graphs, give the same bits.
* The same reuse property in the contact-dense state (the lattice compressed
  until its layers touch: about a million particle pairs and several million
  particle node contacts, so the particle-particle cull and sweep, the contact
  law and the mirror reduction run at size)."""
import numpy as np
import pytest

from fundem_amd import workloads as W

pytestmark = pytest.mark.gpu

STEPS = 40


@pytest.fixture(scope="module")
def pour():
    return W.pour()


def run(scn, skin=None, graph=False):
    from fundem_amd import Solver
    g = Solver()
    try:
        if skin is not None:
            g.set_skin(skin)
        W.apply(scn, g)
        g.use_graph(graph)
        g.step(1)  # the first step runs eagerly
        g.step(STEPS - 1)
        assert g.status() == 0
        return g.state(), g.list_builds()
    finally:
        g.close()


def assert_same(a, b, where):
    for key in ("x", "v", "w", "q", "F", "T"):
        if not np.array_equal(a[key], b[key]):
            bad = np.argwhere(a[key] != b[key])[0]
            raise AssertionError(f"{where}: {key} differs first at {tuple(bad)}")


def test_full_pour_list_reuse_equals_rebuilding_every_step(pour):
    print(f"pour of {len(pour.x)} particles, {STEPS} steps", flush=True)
    kept, (builds, ghost_builds, skin) = run(pour)
    rebuilt, (builds0, _, _) = run(pour, skin=0.0)
    print(f"default skin {skin:.3g} m: {builds} list builds, {ghost_builds} ghost-list builds; "
          f"skin 0: {builds0}", flush=True)
    assert skin > 0 and builds < STEPS // 2, "lists were not reused"
    assert builds0 >= STEPS
    # the floor carries the bottom layer from the first step
    touching = np.count_nonzero(np.abs(kept["F"]).sum(axis=1))
    assert touching > 5000, touching
    assert_same(kept, rebuilt, "skin vs rebuild every step")


def test_full_pour_repeat_and_graph_replay(pour):
    first, _ = run(pour)
    again, _ = run(pour)
    replay, _ = run(pour, graph=True)
    assert_same(again, first, "repeat")
    assert_same(replay, first, "graph replay")


def test_full_dense_pour_list_reuse_equals_rebuilding_every_step(pour):
    from fundem_amd import Solver

    def make_solver():
        g = Solver()
        g.set_capacity(12, 64)
        return g

    dense, done, contacts = W.compressed_pour(pour, make_solver)
    print(f"compressed for {done} steps: {contacts} particle node contacts", flush=True)
    runs = []
    for skin in (None, 0.0):
        g = make_solver()
        if skin is not None:
            g.set_skin(skin)
        W.apply(dense, g)
        g.step(1)
        g.step(19)
        assert g.status() == 0
        pp = g.contacts(0)["A"].size
        runs.append((g.state(), g.list_builds(), pp))
        g.close()
    (a, (builds, _, skin), pp), (b, (builds0, _, _), _) = runs
    print(f"skin {skin:.3g} m: {builds} list builds in 20 steps, skin 0: {builds0}; "
          f"{pp} particle node contacts at the end", flush=True)
    assert pp > 1_000_000 and builds0 >= 20 and builds < builds0
    assert_same(a, b, "dense: kept lists vs rebuilds")
