"""GPU parity of adaptive integration (ai_on 1) against the CPU oracle.

Points and weights bit-identical to the oracle's (both generate them with the same FP64
operations in the same order).  With MLS, the coefficients come from a Cholesky solve here and
a full-pivot Gauss-Jordan solve in the oracle, which agree to ~3e-14 (as for fixed Gaussian
integration): forces within 1e-11 normwise, trajectories within 1e-10.  Without MLS (relative
displacement G xi from DISPLGRAD) no coefficient enters: 1e-12.
"""

import numpy as np
import pytest

import pdmtel_amd as pm
import pyoracle

pytestmark = pytest.mark.gpu


def ai_spec(n, mls=1, nx=8, ny=5, nz=5, gradient=None, **kw):
    s = pm.bar_spec(nx, ny, nz, **kw)
    s["cci"] = None
    s["ai"] = (1, mls, n)
    if gradient is not None:
        s["bcs"] = [dict(type=3, flags=(1, 1, 1), value=tuple(np.asarray(gradient).reshape(-1)), t_end=1.0, load=2)]
        s["boundary_spaces"] = [dict(geo=1, ref=1, box=(-1e3, 1e3, -1e3, 1e3, -1e3, 1e3))]
        s["traction_body"] = 0
    return s


def pair(spec):
    prob = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64)
    orc = pyoracle.Oracle(spec, bond_method=1)
    return prob, orc


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def oracle_points(orc):
    """(node, nbr, xw) of the oracle's points in reference entry order."""
    ep, xw = orc.ai_points()
    rp, col, _ = orc.csr()
    cnt = np.diff(ep)
    rows = np.repeat(np.arange(orc.n), np.diff(rp))
    return np.repeat(rows, cnt), np.repeat(col, cnt), xw


@pytest.mark.parametrize("n", [1, 2, 3])
def test_points_bit_identical(n):
    spec = ai_spec(n, jitter=0.01, seed=n)
    spec["slits"] = [pm.mid_slit(spec, 30.0)]
    prob, orc = pair(spec)
    print(f"[ai points] n={n} nodes={prob.n} points={prob.ai_points()[2].shape[0]}", flush=True)
    ni, nj, xw = prob.ai_points()
    oi, oj, oxw = oracle_points(orc)
    assert len(xw) == len(oxw)
    order = np.lexsort((nj, ni))  # stable: an entry's points keep their order
    assert np.array_equal(ni[order], oi) and np.array_equal(nj[order], oj)
    assert np.array_equal(xw[order], oxw)


@pytest.mark.parametrize("n", [1, 2])
def test_forces_mls(n):
    spec = ai_spec(n, jitter=0.01, seed=3)
    spec["slits"] = [pm.mid_slit(spec, 30.0)]
    prob, orc = pair(spec)
    pyoracle.set_square_mode(True)
    try:
        rng = np.random.default_rng(5)
        X, _, _ = prob.nodes()
        for name, u in (("random", rng.uniform(-1e-6, 1e-6, size=(prob.n, 3))),
                        ("affine", X @ np.array([[1.4e-5, 0, 0], [0, -3.5e-6, 0], [0, 2e-6, -3.5e-6]]).T)):
            F, el = prob.forces(u, with_elast=True)
            Fo, elo = orc.forces(u, with_elast=True)
            ef, ee = rel(F, Fo), rel(el, elo)
            print(f"[ai forces] n={n} {name} eF={ef:.3e} eE={ee:.3e}", flush=True)
            assert ef <= 1e-11 and ee <= 1e-11
    finally:
        pyoracle.set_square_mode(False)


def test_forces_prescribed_gradient():
    G = np.array([[1e-4, 2e-5, 0.0], [2e-5, -3e-5, 1e-5], [0.0, 1e-5, -3e-5]])
    spec = ai_spec(2, mls=0, gradient=G, jitter=0.01, seed=4)
    prob, orc = pair(spec)
    pyoracle.set_square_mode(True)
    try:
        u = np.zeros((prob.n, 3))
        F, el = prob.forces(u, with_elast=True)
        Fo, elo = orc.forces(u, with_elast=True)
        ef, ee = rel(F, Fo), rel(el, elo)
        print(f"[ai forces G] eF={ef:.3e} eE={ee:.3e}", flush=True)
        assert ef <= 1e-12 and ee <= 1e-12
    finally:
        pyoracle.set_square_mode(False)


def test_ai_without_gradient_everywhere_is_refused():
    spec = ai_spec(1, mls=0)
    with pytest.raises(pm.PdmError) as e:
        pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64)
    assert "DISPLGRAD" in str(e.value)


@pytest.mark.parametrize("case", ["displ", "displgrad_ramp"])
def test_rk4_trajectory(case):
    """Steps with the points interpolating each stage's trial displacements; BC nodes' points
    prescribed (DISPL, or DISPLGRAD at the point's own position with a ramp)."""
    spec = ai_spec(2, jitter=0.01, seed=6, steps=10)
    if case == "displgrad_ramp":
        spec["bcs"][0] = dict(type=3, flags=(1, 1, 0), value=(2e-3, 0, 0, 0, -5e-4, 0, 0, 0, 0), t_end=2e-6,
                              load=1.0)
    prob, orc = pair(spec)
    pyoracle.set_square_mode(True)
    try:
        prob.step(3)
        orc.run(3)
        U, V = prob.state()
        Uo, Vo = orc.state()
        eu, ev = rel(U, Uo), rel(V, Vo)
        ke, ee = prob.energy()
        print(f"[ai rk4] {case} eU={eu:.3e} eV={ev:.3e} KE={ke:.3e}/{orc.ke:.3e} EE={ee:.3e}/{orc.ee:.3e}",
              flush=True)
        assert eu <= 1e-10 and ev <= 1e-10
        assert abs(ke - orc.ke) <= 1e-10 * abs(orc.ke) and abs(ee - orc.ee) <= 1e-10 * abs(orc.ee)
    finally:
        pyoracle.set_square_mode(False)
