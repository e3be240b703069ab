"""GPU parity of fixed Gaussian integration (fgi_on 1) against the CPU oracle.

Gauss-point positions within 1e-15 relative.  One Gauss point: forces and energy densities
within 1e-12 normwise, trajectories within 1e-11, as for CCI.  MLS (n > 1): the coefficients
come from a Cholesky solve here and a full-pivot Gauss-Jordan solve in the oracle (both in
node-local coordinates, condition number ~230), which agree to ~3e-14; a Gauss point's
displacement sum_k phi_k u_k carries that difference times the field's offset |u|, against
bond stretches eta ~ |u| dx / L, so forces agree to ~3e-12 on a stretched bar: the bound is
1e-11, and 1e-10 for trajectories, whose step load amplifies differences ~1e3-fold in a few
steps (the same amplification bounds CCI at 1e-11 from 1e-16 per step).
"""

import numpy as np
import pytest

import pdmtel_amd as pm
import pyoracle

pytestmark = pytest.mark.gpu


def fgi_spec(n, nx=10, ny=6, nz=6, **kw):
    s = pm.bar_spec(nx, ny, nz, **kw)
    s["cci"] = None
    s["fgi"] = (1, n)
    return s


def pair(spec, **opts):
    prob = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64, **opts)
    orc = pyoracle.Oracle(spec, bond_method=1)
    return prob, orc


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_gauss_points_and_forces(n):
    spec = fgi_spec(n, jitter=0.01, seed=n)
    prob, orc = pair(spec)
    pyoracle.set_square_mode(True)
    try:
        print(f"[fgi forces] n={n} nodes={prob.n}", flush=True)
        Xg, Ug = prob.gauss()
        Xo, Uo = orc.gauss()
        assert Xg.shape == (prob.n, n ** 3, 3)
        assert np.abs(Xg - Xo).max() <= 1e-15 * np.abs(Xo).max()
        assert np.all(Ug == 0.0) and np.all(Uo == 0.0)
        rng = np.random.default_rng(11)
        X, _, _ = prob.nodes()
        for name, u in (("random", rng.uniform(-1e-6, 1e-6, size=(prob.n, 3))),
                        ("affine", X @ np.array([[1.4e-5, 0, 0], [0, -3.5e-6, 0], [0, 2e-6, -3.5e-6]]).T)):
            F, el = prob.forces(u, with_elast=True)
            Fo, elo = orc.forces(u, with_elast=True)
            ef, ee = rel(F, Fo), rel(el, elo)
            print(f"[fgi forces] n={n} {name} eF={ef:.3e} eE={ee:.3e}", flush=True)
            tol = 1e-12 if n == 1 else 1e-11
            assert ef <= tol and ee <= tol
    finally:
        pyoracle.set_square_mode(False)


@pytest.mark.parametrize("n,load", [(1, 2), (2, 1), (3, 1)])
def test_rk4_matches_oracle(n, load):
    """Stepping, including the BC points' prescribed values (a ramp load shows the time at which
    each stage sets them) and the free points' interpolation of every stage's displacement."""
    spec = fgi_spec(n, steps=30, jitter=0.01, seed=5)
    spec["bcs"][0] = dict(type=1, flags=(1, 0, 1), value=(2e-7, 0.0, -1e-7), t_end=1.0, load=load)
    prob, orc = pair(spec, use_graph=0)
    pyoracle.set_square_mode(True)
    try:
        for chunk in (1, 2, 5):
            prob.step(chunk)
            orc.run(chunk)
            U, V = prob.state()
            Uo, Vo = orc.state()
            _, G = prob.gauss()
            _, Go = orc.gauss()
            eu, ev, eg = rel(U, Uo), rel(V, Vo), rel(G, Go)
            t, step, _ = prob.clock()
            print(f"[fgi rk4] n={n} load={load} step={step} eU={eu:.3e} eV={ev:.3e} eG={eg:.3e}", flush=True)
            assert step == orc.steps_done and t == orc.t
            tol = 1e-11 if n == 1 else 1e-10
            assert eu <= tol and ev <= tol and eg <= tol
        ke, ee = prob.energy()
        assert abs(ke - orc.ke) <= tol * abs(orc.ke)
        assert abs(ee - orc.ee) <= tol * abs(orc.ee)
    finally:
        pyoracle.set_square_mode(False)


def test_graph_matches_eager_and_set_state():
    spec = fgi_spec(2, steps=40, jitter=0.01, seed=8)
    a = pm.Problem(pm.deck_text(spec), use_graph=1, graph_chunk=4)
    b = pm.Problem(pm.deck_text(spec), use_graph=0)
    for _ in range(2):
        a.step(4)
        b.step(4)
    Ua, Va = a.state()
    Ub, Vb = b.state()
    assert np.array_equal(Ua, Ub) and np.array_equal(Va, Vb)
    # set_state re-interpolates the free nodes' points as the end of a step does
    orc = pyoracle.Oracle(spec, bond_method=1)
    pyoracle.set_square_mode(True)
    try:
        orc.run(8)
        rng = np.random.default_rng(2)
        U = rng.uniform(-1e-7, 1e-7, size=(a.n, 3))
        a.set_state(U, None)
        orc.set_state(U, None)
        _, G = a.gauss()
        _, Go = orc.gauss()
        assert rel(G, Go) <= 1e-11
    finally:
        pyoracle.set_square_mode(False)


def test_live_entries_hold_a_point_with_nonzero_beta():
    """The gathers step through the (entry, Gauss point) pairs of uncut entries whose point
    has beta > 0 (cell size dx/2 at n = 2, so beta = 0 from r = delta + dx/4 on);
    pdm_num_live_entries counts the entries holding at least one such pair.  r is formed as
    the library forms it: (offset + X_j) - X_i."""
    spec = fgi_spec(2, jitter=0.01, seed=4)
    spec["slits"] = [pm.mid_slit(spec, 30.0)]
    prob = pm.Problem(pm.deck_text(spec))
    rp, col, cut = prob.csr()
    X, _, _ = prob.nodes()
    X = X.reshape(-1, 3)
    Xg, _ = prob.gauss()
    row = np.repeat(np.arange(prob.n), np.diff(rp))
    d = Xg[col] - X[row][:, None, :]
    r = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2])
    delta, cell = spec["horizon"], spec["grid_spacing"] / 2.0
    inner, outer = delta - 0.5 * cell, delta + 0.5 * cell
    live = (r <= inner) | ((r <= outer) & ((delta + 0.5 * cell - r) / cell != 0.0))
    live &= (cut == 0)[:, None]
    want = int(live.any(axis=1).sum())
    print(f"[fgi live] E={prob.E} cut={int(cut.sum())} live entries={prob.E_live} want={want} "
          f"live pairs={int(live.sum())} of {live.size}", flush=True)
    assert cut.sum() > 0 and want < prob.E - int(cut.sum())
    assert prob.E_live == want
    assert prob.live_gauss_pairs == int(live.sum())


@pytest.mark.parametrize("prec", [pm.PDM_FP32C, pm.PDM_FP32C2])
def test_fp32_bond_arithmetic(prec):
    """FP32 bond forces over the Gauss points with FP64 Gauss-point displacements and RK4
    state: forces within 1e-6 of the oracle's FP64 forces, and after 200 steps the displacement
    within 1e-6 of the FP64 path's (max |dU| / max |U|)."""
    spec = fgi_spec(2, nx=16, ny=6, nz=6, steps=400, jitter=0.01, seed=9)
    spec["slits"] = [pm.mid_slit(spec, 30.0)]
    p32 = pm.Problem(pm.deck_text(spec), precision=prec)
    p64 = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64)
    orc = pyoracle.Oracle(spec, bond_method=1)
    rng = np.random.default_rng(3)
    u = rng.uniform(-1e-6, 1e-6, size=(p32.n, 3))
    F, el = p32.forces(u, with_elast=True)
    Fo, elo = orc.forces(u, with_elast=True)
    ef, ee = rel(F, Fo), rel(el, elo)
    p32.step(200)
    p64.step(200)
    eu = rel(p32.state()[0], p64.state()[0])
    print(f"[fgi fp32] prec={prec} eF={ef:.3e} eE={ee:.3e} eU(200 steps)={eu:.3e}", flush=True)
    assert ef <= 1e-6 and ee <= 1e-6 and eu <= 1e-6


def test_unsupported_combinations_fail_loudly():
    spec = fgi_spec(2)
    text = pm.deck_text(spec)
    prob = pm.Problem(text)
    with pytest.raises(pm.PdmError) as e:
        prob.solve_static()
    assert e.value.code == pm.PDM_E_MODEL
    with pytest.raises(pm.PdmError):
        pm.Batch([text, text], precision=pm.PDM_FP64)
