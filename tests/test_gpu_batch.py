"""Several independent systems advanced by one solver: each system must
evolve exactly as the oracle evolves it alone (same templates, its own
domain, periodic axes, materials and walls)."""
import numpy as np
import pytest

import pyoracle as po
from fundem_amd import workloads as W

pytestmark = pytest.mark.gpu

SETS = (po.SET_PP, po.SET_PW, po.SET_GX, po.SET_GY, po.SET_GXY)


def systems():
    base = W.cluster(n=27, seed=1, periodic=(True, True), speed=1.0)
    tm = base.templates
    return [
        base,
        W.cluster(n=20, seed=2, periodic=(False, False), speed=1.0, kn=2e4, mu=0.3, templates=tm),
        W.cluster(n=33, seed=3, periodic=(True, False), speed=1.5, templates=tm),
        W.cluster(n=8, seed=4, periodic=(False, True), speed=1.0, zeta=0.5, templates=tm),
        W.cluster(n=27, seed=5, periodic=(True, True), speed=0.5, walls=False, templates=tm),
    ]


def split_contacts(c, a_off, a_n, b_off):
    sel = (c["A"] >= a_off) & (c["A"] < a_off + a_n)
    return {"A": c["A"][sel] - a_off, "B": c["B"][sel] - b_off, "k": c["k"][sel],
            "spring": c["spring"][sel]}


def rule_cell(oracle, n_templates):
    """The grid rule's cell: twice the largest template radius."""
    return 2.0 * max(oracle.geometry(k)["radius"] for k in range(n_templates))


@pytest.mark.parametrize("skin,cells", [(None, None), (0.0, None), (None, 1.5)])
def test_batch_systems_equal_their_plain_runs(skin, cells):
    """With cells=1.5 every instance grid of the batch has cells 1.5 times
    the rule's (as the autotuner may choose) while each oracle grids by the
    rule."""
    from fundem_amd import Solver
    scns = systems()
    oracles = []
    for s in scns:
        o = po.OracleSim(True, True)
        W.apply(s, o)
        oracles.append(o)
    g = Solver()
    if skin is not None:
        g.set_skin(skin)
    if cells is not None:
        g.set_cell_size(cells * rule_cell(oracles[0], len(scns[0].templates)))
    W.apply_batch(scns, g)
    p_off = np.cumsum([0] + [s.n for s in scns])
    w_off = np.cumsum([0] + [len(s.walls) for s in scns])
    total = 0
    for step in range(1, 301):
        g.step(1)
        for o in oracles:
            o.step(1)
        if step % 50 and step != 300:
            continue
        gs = g.state()
        gc = {k: g.contacts(k) for k in SETS}
        for i, (s, o) in enumerate(zip(scns, oracles)):
            os_ = o.state()
            for key in ("x", "v", "w", "q", "F", "T"):
                a = gs[key][p_off[i]:p_off[i + 1]]
                assert np.array_equal(a, os_[key]), f"step {step} system {i}: {key} differs"
            for k in SETS:
                b_off = w_off[i] if k == po.SET_PW else p_off[i]
                c = split_contacts(gc[k], p_off[i], s.n, b_off)
                oc = o.contacts(k)
                for key in ("A", "B", "k", "spring"):
                    assert np.array_equal(c[key], oc[key]), \
                        f"step {step} system {i} set {k}: {key} differs ({c['A'].size} vs {oc['A'].size})"
                total += oc["A"].size
    assert total > 0
    assert g.status() == 0


def test_batch_of_packings_matches_plain_solvers():
    """Packings of the batched sweep (shared templates, per-packing
    materials, periodic boxes with floors): the batch equals plain solver runs
    of the same packings, bit for bit, including the wall contacts."""
    from fundem_amd import Solver
    tm = W.superellipsoid_family(4, 9, n_grid=16, n_nodes=128)
    scns = W.batch_packings(count=5, seed=3, n_range=(60, 300), templates=tm)
    g = Solver()
    W.apply_batch(scns, g)
    plain = []
    for s in scns:
        p = Solver()
        W.apply(s, p)
        plain.append(p)
    p_off = np.cumsum([0] + [s.n for s in scns])
    w_off = np.cumsum([0] + [len(s.walls) for s in scns])
    g.step(400)
    for p in plain:
        p.step(400)
    gs = g.state()
    wall_contacts = 0
    for i, (s, p) in enumerate(zip(scns, plain)):
        ps = p.state()
        for key in ("x", "v", "w", "q", "F", "T"):
            assert np.array_equal(gs[key][p_off[i]:p_off[i + 1]], ps[key]), f"system {i}: {key}"
        for k in SETS:
            b_off = w_off[i] if k == po.SET_PW else p_off[i]
            c = split_contacts(g.contacts(k), p_off[i], s.n, b_off)
            pc = p.contacts(k)
            for key in ("A", "B", "k", "spring"):
                assert np.array_equal(c[key], pc[key]), f"system {i} set {k}: {key}"
            if k == po.SET_PW:
                wall_contacts += pc["A"].size
    assert wall_contacts > 0
    assert g.status() == 0
