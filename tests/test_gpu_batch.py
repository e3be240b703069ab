"""GPU: batched specimens (one launch per RK4 stage over every specimen's tiles).

A batch run with a fixed tile layout must give every specimen bit-for-bit the
state, clock and energies of the same specimen run alone with that layout (the
per-tile arithmetic and the per-specimen energy sums are the same), and match the
oracle to the single-specimen tolerance."""

import numpy as np
import pytest

import pdmtel_amd as pm
import pyoracle
from pdmtel_amd.deck import specimen_specs

pytestmark = pytest.mark.gpu

LAYOUT = dict(tile_edge=4, waves_per_tile=8, families_per_wave=2)


def small_specs():
    specs = specimen_specs(6, seed=5, nodes=(300, 3000), steps=40)
    specs[2]["total_step"] = 4  # this specimen's run ends inside the batch run
    return specs


@pytest.mark.parametrize("precision", [pm.PDM_FP64, pm.PDM_FP32C])
def test_batch_equals_single(precision):
    specs = small_specs()
    decks = [pm.deck_text(s) for s in specs]
    b = pm.Batch(decks, precision=precision, graph_chunk=5, **LAYOUT)
    print(f"[batch] precision={precision} nodes={b.n} entries={sum(b.E)}", flush=True)
    assert b.step(7) == 7  # a captured chunk of 5 steps, then one of 2
    for k, d in enumerate(decks):
        p = pm.Problem(d, precision=precision, graph_chunk=5, **LAYOUT)
        done = p.step(7)
        assert done == (4 if k == 2 else 7)
        U, V = p.state()
        Ub, Vb = b.state(k)
        assert np.array_equal(U, Ub) and np.array_equal(V, Vb), k
        assert p.clock() == b.clock(k), k
        assert p.energy() == b.energy(k), k
        assert p.E_live == b.E_live[k] and 0 < p.E_live < p.E, k
        p.close()
    b.close()


def test_batch_matches_oracle_and_set_state():
    specs = small_specs()[:3]
    decks = [pm.deck_text(s) for s in specs]
    b = pm.Batch(decks, precision=pm.PDM_FP64, graph_chunk=5, **LAYOUT)
    rng = np.random.default_rng(3)
    orcs = [pyoracle.Oracle(s, bond_method=1) for s in specs]
    for k, o in enumerate(orcs):
        V0 = rng.uniform(-1e-3, 1e-3, size=(o.n, 3))
        o.set_state(None, V0)
        b.set_state(k, None, V0)
    pyoracle.set_square_mode(True)
    try:
        b.step(3)
        for k, o in enumerate(orcs):
            o.run(3)
            U, V = b.state(k)
            Uo, Vo = o.state()
            eu = np.abs(U - Uo).max() / np.abs(Uo).max()
            ev = np.abs(V - Vo).max() / np.abs(Vo).max()
            print(f"[batch vs oracle] specimen {k} eU={eu:.2e} eV={ev:.2e}", flush=True)
            assert eu <= 1e-11 and ev <= 1e-11
    finally:
        pyoracle.set_square_mode(False)
    b.close()
