"""GPU parity: the HIP library (through the C ABI) against the CPU oracle.

Bar: bit-exact for the bond set (row_ptr, col, cut), bdryEff and node data;
forces within 1e-12 normwise (max |dF| / max |F|) in FP64 with the oracle's
pow(x, 2.0) evaluated as the correctly rounded square; RK4 trajectories within
1e-11 normwise after a few steps (only the family summation order differs).
"""

import numpy as np
import pytest

import pdmtel_amd as pm
import pyoracle

pytestmark = pytest.mark.gpu

CASES = {
    "bar_12x6x6": dict(nx=12, ny=6, nz=6),
    "bar_20x10x10": dict(nx=20, ny=10, nz=10),
    "bar_jitter": dict(nx=16, ny=8, nz=8, jitter=0.01, seed=0),
    "bar_delta4": dict(nx=14, ny=10, nz=10, delta_over_dx=4.015, jitter=0.01, seed=3),
    "bar_cci_original": dict(nx=12, ny=8, nz=8, cci_modified=0),
}


def make(case, **over):
    kw = dict(CASES[case])
    kw.update(over)
    return pm.bar_spec(kw.pop("nx"), kw.pop("ny"), kw.pop("nz"), **kw)


def pair(spec, precision=pm.PDM_FP64, **opts):
    prob = pm.Problem(pm.deck_text(spec), precision=precision, **opts)
    orc = pyoracle.Oracle(spec, bond_method=1)
    return prob, orc


@pytest.mark.parametrize("case", list(CASES))
def test_bond_set_bit_exact(case):
    spec = make(case)
    prob, orc = pair(spec)
    print(f"[bond set] {case} n={prob.n}", flush=True)
    rp, col, cut = prob.csr()
    orp, ocol, ocut = orc.csr()
    assert prob.n == orc.n
    assert np.array_equal(rp, orp)
    assert np.array_equal(col, ocol)
    assert np.array_equal(cut, ocut)
    assert prob.max_family == orc.max_family
    X, mat, bc = prob.nodes()
    Xo, mato, bco = orc.nodes()
    assert np.array_equal(X, Xo) and np.array_equal(mat, mato) and np.array_equal(bc, bco)
    assert np.array_equal(prob.bdryeff(), orc.bdryeff())


def test_bond_set_with_slits():
    spec = make("bar_jitter")
    spec["slits"] = [pm.mid_slit(spec, 30.0),
                     dict(type=1, box=(0.0, 1.0, 0.004, 0.004, 0.0, 1.0))]
    prob, orc = pair(spec)
    rp, col, cut = prob.csr()
    orp, ocol, ocut = orc.csr()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.array_equal(cut, ocut)
    assert cut.sum() > 0
    assert np.array_equal(prob.bdryeff(), orc.bdryeff())


@pytest.mark.parametrize("case", ["bar_jitter", "bar_delta4", "bar_cci_original"])
def test_live_entries(case):
    """The gathers evaluate the uncut entries whose CCI beta is nonzero: r < delta + dx/2
    (modified beta), r <= delta (original), with r formed as the library forms it."""
    spec = make(case)
    spec["slits"] = [pm.mid_slit(spec, 30.0)]
    prob = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64)
    rp, col, cut = prob.csr()
    X, _, _ = prob.nodes()
    X = X.reshape(-1, 3)
    row = np.repeat(np.arange(prob.n), np.diff(rp))
    d = X[col] - X[row]
    r = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    dx, delta = spec["grid_spacing"], spec["horizon"]
    outer = delta if spec["cci"][1] == 0 else delta + 0.5 * dx
    beta_pos = (r <= outer) & (delta + 0.5 * dx - r > 0.0)
    want = int(((cut == 0) & beta_pos).sum())
    print(f"[live entries] {case} E={prob.E} live={prob.E_live} want={want}", flush=True)
    assert cut.sum() > 0 and want < prob.E
    assert prob.E_live == want


def test_rk4_gravity_with_slits():
    """Gravity acts once per unbroken family entry (pdDataManager.cpp:1767-1992), entries with
    beta = 0 included: the count survives the gathers skipping those entries."""
    spec = make("bar_jitter", steps=30)
    spec["slits"] = [pm.mid_slit(spec, 30.0)]
    spec["materials"][0]["gravity"] = (3.0e4, -2.0e4, 1.0e4)
    prob, orc = pair(spec, use_graph=0)
    pyoracle.set_square_mode(True)
    try:
        for chunk in (2, 5):
            prob.step(chunk)
            orc.run(chunk)
            U, V = prob.state()
            Uo, Vo = orc.state()
            eu = np.abs(U - Uo).max() / np.abs(Uo).max()
            ev = np.abs(V - Vo).max() / np.abs(Vo).max()
            print(f"[rk4 gravity] step={prob.clock()[1]} eU={eu:.3e} eV={ev:.3e}", flush=True)
            assert eu <= 1e-11 and ev <= 1e-11
    finally:
        pyoracle.set_square_mode(False)


@pytest.mark.parametrize("case", ["bar_12x6x6", "bar_jitter", "bar_delta4", "bar_cci_original"])
def test_forces_fp64(case):
    spec = make(case)
    prob, orc = pair(spec)
    pyoracle.set_square_mode(True)
    try:
        rng = np.random.default_rng(7)
        X, _, _ = prob.nodes()
        for name, u in (("random", rng.uniform(-1e-6, 1e-6, size=(prob.n, 3))),
                        ("affine", X @ np.array([[1.4e-5, 0, 0], [0, -3.5e-6, 0], [0, 2e-6, -3.5e-6]]).T)):
            F, el = prob.forces(u, with_elast=True)
            Fo, elo = orc.forces(u, with_elast=True)
            err = np.abs(F - Fo).max() / np.abs(Fo).max()
            print(f"[forces fp64] {case} {name} err={err:.3e}", flush=True)
            assert err <= 1e-12, (name, err)
            eerr = np.abs(el - elo).max() / np.abs(elo).max()
            assert eerr <= 1e-12, (name, eerr)
    finally:
        pyoracle.set_square_mode(False)


@pytest.mark.parametrize("case", ["bar_12x6x6", "bar_jitter"])
def test_rk4_fp64(case):
    spec = make(case, steps=40)
    prob, orc = pair(spec, use_graph=0)
    pyoracle.set_square_mode(True)
    try:
        dt, mb = prob.timestep()
        assert abs(dt - orc.dt) <= 4e-16 * orc.dt
        for chunk in (1, 3, 6):
            prob.step(chunk)
            orc.run(chunk)
            U, V = prob.state()
            Uo, Vo = orc.state()
            eu = np.abs(U - Uo).max() / np.abs(Uo).max()
            ev = np.abs(V - Vo).max() / np.abs(Vo).max()
            t, step, _ = prob.clock()
            print(f"[rk4 fp64] {case} step={step} eU={eu:.3e} eV={ev:.3e}", flush=True)
            assert step == orc.steps_done and t == orc.t
            assert eu <= 1e-11 and ev <= 1e-11
        ke, ee = prob.energy()
        assert abs(ke - orc.ke) <= 1e-11 * abs(orc.ke)
        assert abs(ee - orc.ee) <= 1e-11 * abs(orc.ee)
    finally:
        pyoracle.set_square_mode(False)


def test_rk4_graph_matches_eager():
    spec = make("bar_jitter", steps=40)
    a = pm.Problem(pm.deck_text(spec), use_graph=1, graph_chunk=5)
    b = pm.Problem(pm.deck_text(spec), use_graph=0)
    for _ in range(3):
        a.step(5)
        b.step(5)
        Ua, Va = a.state()
        Ub, Vb = b.state()
        assert np.array_equal(Ua, Ub) and np.array_equal(Va, Vb)


def test_rk4_prepared_graph_matches_eager():
    """pdm_prepare_steps captures a 7-step chunk before any step runs; chunks of 3 (captured on
    the way) and 7 (replayed) then give what eager launches give, bit for bit, and the clock
    and energies agree."""
    spec = make("bar_jitter", steps=40)
    a = pm.Problem(pm.deck_text(spec), use_graph=1, graph_chunk=7, energies=1)
    b = pm.Problem(pm.deck_text(spec), use_graph=0, energies=1)
    a.prepare(7)
    with pytest.raises(pm.PdmError):
        a.prepare(0)
    for n in (3, 7, 7, 3):
        assert a.step(n) == n and b.step(n) == n
        Ua, Va = a.state()
        Ub, Vb = b.state()
        assert np.array_equal(Ua, Ub) and np.array_equal(Va, Vb)
        assert a.clock() == b.clock() and a.energy() == b.energy()


@pytest.mark.parametrize("case", ["bar_jitter", "bar_delta4", "bar_cci_original", "slits"])
@pytest.mark.parametrize("precision", [pm.PDM_FP32C, pm.PDM_FP32C2])
def test_forces_fp32(case, precision):
    """FP32 bond forces (coefficients formed from the records and the per-entry correction
    bytes; cut and zero-beta entries left out of the rows) against the FP64 oracle."""
    if case == "slits":
        spec = make("bar_jitter")
        spec["slits"] = [pm.mid_slit(spec, 30.0), dict(type=1, box=(0.0, 1.0, 0.004, 0.004, 0.0, 1.0))]
    else:
        spec = make(case)
    prob, orc = pair(spec, precision=precision)
    rng = np.random.default_rng(9)
    X, _, _ = prob.nodes()
    for name, u in (("random", rng.uniform(-1e-6, 1e-6, size=(prob.n, 3))),
                    ("affine", X @ np.array([[1.4e-5, 0, 0], [0, -3.5e-6, 0], [0, 2e-6, -3.5e-6]]).T)):
        F, el = prob.forces(u, with_elast=True)
        Fo, elo = orc.forces(u, with_elast=True)
        err = np.abs(F - Fo).max() / np.abs(Fo).max()
        eerr = np.abs(el - elo).max() / np.abs(elo).max()
        print(f"[forces fp32] {case} precision={precision} {name} err={err:.3e} energy={eerr:.3e}", flush=True)
        assert err <= 1e-6 and eerr <= 1e-6, (name, err, eerr)


@pytest.mark.parametrize("layout", [(2, 4, 1), (3, 8, 4), (5, 4, 2), (7, 8, 1)])
@pytest.mark.parametrize("precision", [pm.PDM_FP64, pm.PDM_FP32C])
def test_forces_every_layout(layout, precision):
    """Tile edge, wavefronts per tile and families per wavefront change the blocks, the
    lane mapping and the LDS footprint, not the forces."""
    edge, waves, fams = layout
    spec = make("bar_delta4")
    spec["slits"] = [pm.mid_slit(spec, 30.0)]
    prob, orc = pair(spec, precision=precision, tile_edge=edge, waves_per_tile=waves, families_per_wave=fams)
    assert prob.tuning()["tile_edge"] == edge
    pyoracle.set_square_mode(True)
    try:
        u = np.random.default_rng(4).uniform(-1e-6, 1e-6, size=(prob.n, 3))
        F = prob.forces(u)
        Fo = orc.forces(u)
        err = np.abs(F - Fo).max() / np.abs(Fo).max()
        print(f"[forces layout] {layout} precision={precision} err={err:.3e}", flush=True)
        assert err <= (1e-12 if precision == pm.PDM_FP64 else 1e-6)
    finally:
        pyoracle.set_square_mode(False)


@pytest.mark.parametrize("layout", [(4, 8, 4), (3, 4, 4), (5, 8, 2), (4, 4, 1)])
@pytest.mark.parametrize("precision", [pm.PDM_FP64, pm.PDM_FP32C, pm.PDM_FP32C2])
def test_rk4_every_layout(layout, precision):
    """The stage kernel in each lane mapping (16 to 64 lanes per family, 4 or 8 wavefronts
    per tile) steps the slit bar as the oracle does: 1e-11 in FP64, 1e-6 in the FP32 modes,
    after 12 steps."""
    edge, waves, fams = layout
    spec = make("bar_delta4", steps=40)
    spec["slits"] = [pm.mid_slit(spec, 30.0)]
    prob, orc = pair(spec, precision=precision, use_graph=0, tile_edge=edge, waves_per_tile=waves,
                     families_per_wave=fams)
    assert prob.tuning()["families_per_wave"] == fams
    pyoracle.set_square_mode(True)
    try:
        prob.step(12)
        orc.run(12)
        U, V = prob.state()
        Uo, Vo = orc.state()
        eu = np.abs(U - Uo).max() / np.abs(Uo).max()
        ev = np.abs(V - Vo).max() / np.abs(Vo).max()
        print(f"[rk4 layout] {layout} precision={precision} eU={eu:.3e} eV={ev:.3e}", flush=True)
        tol = 1e-11 if precision == pm.PDM_FP64 else 1e-6
        assert eu <= tol and ev <= tol
    finally:
        pyoracle.set_square_mode(False)


def test_rk4_fp32c_error():
    spec = make("bar_jitter", steps=40)
    p64 = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64)
    p32 = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP32C)
    p64.step(40)
    p32.step(40)
    U64, _ = p64.state()
    U32, _ = p32.state()
    err = np.abs(U32 - U64).max() / np.abs(U64).max()
    print(f"[rk4 fp32c] max rel displacement error vs fp64 after 40 steps: {err:.3e}", flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("precision", [pm.PDM_FP64, pm.PDM_FP32C, pm.PDM_FP32C2])
def test_static_solve_matches_dense(precision):
    """Matrix-free PCG (FP64, or FP32 / FP32 + 16-bit weights with FP64 residual replacement)
    vs a dense LAPACK solve of the same masked tangent system assembled from the oracle."""
    spec = make("bar_jitter")
    prob, orc = pair(spec, precision=precision)
    st = prob.solve_static(rel_tol=1e-10)
    U, V = prob.state()
    Ud = orc.solve_static_dense()
    err = np.abs(U - Ud).max() / np.abs(Ud).max()
    print(f"[static] precision={precision} iters={st.iterations} sweeps={st.refinements} "
          f"res={st.rel_residual:.2e} err={err:.2e} t={st.seconds:.3f}s", flush=True)
    assert st.rel_residual <= 1e-10
    assert err <= 1e-7
    assert np.all(V == 0.0)


def test_static_solve_edge_crack_fp32c():
    """Static solve of a bar with an edge crack (cut bonds): FP32C refinement vs the dense solve."""
    spec = make("bar_jitter")
    spec["slits"] = [pm.mid_slit(spec, 30.0, edge=True)]
    prob, orc = pair(spec, precision=pm.PDM_FP32C)
    assert np.array_equal(prob.csr()[2], orc.csr()[2]) and prob.csr()[2].sum() > 0
    st = prob.solve_static(rel_tol=1e-10)
    U, _ = prob.state()
    Ud = orc.solve_static_dense()
    err = np.abs(U - Ud).max() / np.abs(Ud).max()
    print(f"[static edge crack] iters={st.iterations} res={st.rel_residual:.2e} err={err:.2e}", flush=True)
    assert st.rel_residual <= 1e-10 and err <= 1e-7


@pytest.mark.parametrize("precision", [pm.PDM_FP64, pm.PDM_FP32C, pm.PDM_FP32C2])
def test_tangent_operator(precision):
    """pdm_apply_tangent (per-entry weights coef / r^3 precomputed once) against the oracle's
    tangent, on a jittered, slit bar: FP64 within 1e-12, FP32 weights within 1e-6, FP32 weights
    with their 16-bit corrections (within 2^-39 of the FP64 weights) within 1e-10; the operator
    is symmetric."""
    spec = make("bar_jitter")
    spec["slits"] = [pm.mid_slit(spec, 30.0, edge=True)]
    prob, orc = pair(spec, precision=precision)
    rng = np.random.default_rng(17)
    u = rng.uniform(-1e-6, 1e-6, size=(prob.n, 3))
    v = rng.uniform(-1e-6, 1e-6, size=(prob.n, 3))
    Ku = prob.tangent(u)
    Kuo = orc.tangent(u)
    err = np.abs(Ku - Kuo).max() / np.abs(Kuo).max()
    sym = abs(np.vdot(v, Ku) - np.vdot(u, prob.tangent(v))) / abs(np.vdot(v, Ku))
    print(f"[tangent] precision={precision} err={err:.3e} symmetry={sym:.3e}", flush=True)
    tol = {pm.PDM_FP64: 1e-12, pm.PDM_FP32C: 1e-6, pm.PDM_FP32C2: 1e-10}[precision]
    assert err <= tol and sym <= tol
