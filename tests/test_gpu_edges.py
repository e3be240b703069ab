"""GPU edge cases against the CPU oracle: tiny specimens, families that are empty once the
entries with beta = 0 are skipped, and nodes without any bond.

Two nodes at lattice offset (3, 2, 1) dx are bonded (distance to the neighbour's cell 2.96 dx
< delta = 3.015 dx) but their beta is 0 (r = sqrt(14) dx > delta + dx/2): their rows hold one
entry each in the reference and none in the gathers, so the force is 0 and gravity still acts
once per entry.  A single node has no bond at all.  A fixed time step keeps dt independent of
the (degenerate) stable-step estimate, and bdryEff is off (a family without a beta > 0 entry has
no energy to compensate).
"""

import numpy as np
import pytest

import pdmtel_amd as pm
import pyoracle

pytestmark = pytest.mark.gpu

DX = 1e-3
BIG = 1.0


def lattice_spec(nx, ny, nz, deletions=(), gravity=(2.0e4, -1.0e4, 5.0e3), steps=20):
    spec = pm.bar_spec(nx, ny, nz, dx=DX, steps=steps, bdryeff_on=0)
    spec["materials"][0]["gravity"] = gravity
    spec["fixed_time_step"] = 1e-8
    spec["total_time"] = steps * 1e-8
    # no BC nodes: every node moves under gravity and its bonds
    spec["bcs"] = [dict(type=1, flags=(1, 1, 1), value=(0.0, 0.0, 0.0), t_end=1.0, load=2)]
    spec["boundary_spaces"] = [dict(geo=1, ref=1, box=(-BIG, -0.5 * BIG, -BIG, BIG, -BIG, BIG))]
    spec["deletion_spaces"] = [dict(geo=1, box=b) for b in deletions]
    return spec


def two_nodes_zero_beta():
    """Keep lattice nodes (0,0,0) and (3,2,1) of a 4x3x2 block."""
    d = DX
    dels = [(1 * d, 3 * d, -BIG, BIG, -BIG, BIG),          # layers i = 1, 2
            (-BIG, 1 * d, 1 * d, BIG, -BIG, BIG),           # i = 0: j >= 1
            (-BIG, 1 * d, -BIG, BIG, 1 * d, BIG),           # i = 0: k >= 1
            (3 * d, BIG, -BIG, 2 * d, -BIG, BIG),           # i = 3: j < 2
            (3 * d, BIG, -BIG, BIG, -BIG, 1 * d)]           # i = 3: k < 1
    return lattice_spec(4, 3, 2, dels)


def rel(a, b):
    """max |a - b| / max |b|; max |a| where b is 0 (a node that never moves)."""
    d, m = np.abs(a - b).max(), np.abs(b).max()
    return d / m if m > 0 else d


def run_pair(spec, precision):
    prob = pm.Problem(pm.deck_text(spec), precision=precision, use_graph=0)
    orc = pyoracle.Oracle(spec, bond_method=1)
    return prob, orc


@pytest.mark.parametrize("precision", [pm.PDM_FP64, pm.PDM_FP32C])
@pytest.mark.parametrize("case", ["zero_beta_pair", "single_node", "pair_dx"])
def test_tiny_specimens(case, precision):
    if case == "zero_beta_pair":
        spec = two_nodes_zero_beta()
    elif case == "single_node":
        spec = lattice_spec(1, 1, 1)
    else:
        spec = lattice_spec(2, 1, 1)
    prob, orc = run_pair(spec, precision)
    rp, col, cut = prob.csr()
    orp, ocol, _ = orc.csr()
    print(f"[edges] {case} prec={precision} n={prob.n} E={prob.E} live={prob.E_live}", flush=True)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    want_live = {"zero_beta_pair": 0, "single_node": 0, "pair_dx": 2}[case]
    want_n = {"zero_beta_pair": 2, "single_node": 1, "pair_dx": 2}[case]
    assert prob.n == want_n and prob.E_live == want_live
    if case == "zero_beta_pair":
        assert prob.E == 2
    pyoracle.set_square_mode(True)
    try:
        for chunk in (1, 4):
            prob.step(chunk)
            orc.run(chunk)
            U, V = prob.state()
            Uo, Vo = orc.state()
            tol = 1e-12 if precision == pm.PDM_FP64 else 1e-6
            eu, ev = rel(U, Uo), rel(V, Vo)
            print(f"[edges] {case} step={prob.clock()[1]} eU={eu:.3e} eV={ev:.3e}", flush=True)
            assert eu <= tol and ev <= tol
        if case == "single_node":
            # no family entry, so no body force either (it acts once per entry)
            assert np.all(U == 0.0) and np.all(V == 0.0)
        else:
            # no bond force acts (beta = 0, or no relative motion): both nodes move alike,
            # driven by gravity once per entry
            assert np.array_equal(U[0], U[1]) and np.abs(U).max() > 0
    finally:
        pyoracle.set_square_mode(False)
