"""GPU parity at full size: the 320,000-node bar of the FP64 and precision workloads.

Workloads (pdmtel_amd.deck.bar_spec): a 200 x 40 x 40 bar, dx = 1 mm, delta = 3.015 dx,
node jitter 0.01 dx from seed 0, x = 0 layer fixed, 1 MPa pulling the x = L layer, 2,000
RK4 steps.  "plain" is the bar as it is (73.7M family entries); "slit" adds one planar
slit through the whole section at the bar centre, normal at 30 degrees to x in the x-y
plane.  Layout options are left at 0, so the autotuner picks the tile layout, as in
production runs.

Bar, against the oracle (the reference's CPU algorithm, OpenMP over nodes):
  - row_ptr / col / cut, node data and bdryEff bit-exact;
  - FP64 forces and energy densities within 1e-12 normwise for a random and an affine field;
  - FP32C and FP32C2 forces within 1e-6 normwise;
  - FP64 RK4 state within 1e-11 after 3 steps;
then the FP32C and FP32C2 runs of all 2,000 steps within 1e-6 of the FP64 run
(max |U32 - U64| / max |U64|).
"""

import numpy as np
import pytest

import pdmtel_amd as pm
import pyoracle

pytestmark = pytest.mark.gpu

STEPS = 2000


def bar(slit: bool):
    spec = pm.bar_spec(200, 40, 40, jitter=0.01, seed=0, steps=STEPS)
    if slit:
        spec["slits"] = [pm.mid_slit(spec, 30.0)]
    return spec


@pytest.fixture(scope="module")
def plain():
    spec = bar(False)
    print("[full] oracle init: bar 200x40x40", flush=True)
    return spec, pyoracle.Oracle(spec, bond_method=1)


@pytest.fixture(scope="module")
def slit():
    spec = bar(True)
    print("[full] oracle init: bar 200x40x40 with a 30 deg slit", flush=True)
    return spec, pyoracle.Oracle(spec, bond_method=1)


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def check_bond_set(prob, orc, name):
    rp, col, cut = prob.csr()
    orp, ocol, ocut = orc.csr()
    print(f"[full] {name}: n={prob.n} E={prob.E} live={prob.E_live} cut={int(cut.sum())} "
          f"tuning={prob.tuning()}", flush=True)
    assert prob.n == orc.n == 320000
    assert np.array_equal(rp, orp)
    assert np.array_equal(col, ocol)
    assert np.array_equal(cut, ocut)
    X, mat, bc = prob.nodes()
    Xo, mato, bco = orc.nodes()
    assert np.array_equal(X, Xo) and np.array_equal(mat, mato) and np.array_equal(bc, bco)
    assert np.array_equal(prob.bdryeff(), orc.bdryeff())
    assert prob.max_family == orc.max_family
    return X


def fields(X, seed):
    rng = np.random.default_rng(seed)
    grad = np.array([[1.4e-5, 0, 0], [0, -3.5e-6, 0], [0, 2e-6, -3.5e-6]])
    return (("random", rng.uniform(-1e-6, 1e-6, size=X.shape)), ("affine", X @ grad.T))


def check_forces(prob, orc, X, bar_, name):
    for fname, u in fields(X, 11):
        F, el = prob.forces(u, with_elast=True)
        Fo, elo = orc.forces(u, with_elast=True)
        ef, ee = rel(F, Fo), rel(el, elo)
        print(f"[full] {name} forces {fname}: err={ef:.3e} energy={ee:.3e} (bar {bar_:g})", flush=True)
        assert ef <= bar_ and ee <= bar_, (fname, ef, ee)


def check_rk4_oracle(prob, orc, name, steps=3):
    pyoracle.set_square_mode(True)
    try:
        prob.step(steps)
        orc.run(steps)
    finally:
        pyoracle.set_square_mode(False)
    U, V = prob.state()
    Uo, Vo = orc.state()
    eu, ev = rel(U, Uo), rel(V, Vo)
    t, step, _ = prob.clock()
    print(f"[full] {name} rk4 fp64 vs oracle after {step} steps: eU={eu:.3e} eV={ev:.3e}", flush=True)
    assert step == orc.steps_done == steps and t == orc.t
    assert eu <= 1e-11 and ev <= 1e-11
    ke, ee = prob.energy()
    assert abs(ke - orc.ke) <= 1e-11 * abs(orc.ke) and abs(ee - orc.ee) <= 1e-11 * abs(orc.ee)


def run_precisions(spec, p64, name):
    """Finish the FP64 run and step FP32C and FP32C2 over the same steps."""
    done = p64.clock()[1]
    p64.step(STEPS - done)
    U64, _ = p64.state()
    assert p64.clock()[1] == STEPS and np.isfinite(U64).all()
    for prec, label in ((pm.PDM_FP32C, "fp32c"), (pm.PDM_FP32C2, "fp32c2")):
        p = pm.Problem(pm.deck_text(spec), precision=prec)
        p.step(STEPS)
        U, _ = p.state()
        ms, launches = p.kernel_timing()
        err = rel(U, U64)
        print(f"[full] {name} {label}: {STEPS} steps, stage {ms:.4f} ms x {launches}, "
              f"max rel displacement error vs fp64 {err:.3e} (launch {p.stage_launch()}, "
              f"tuning {p.tuning()})", flush=True)
        assert err <= 1e-6
        p.close()


def test_plain_bar_fp64_against_oracle(plain):
    spec, orc = plain
    prob = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64)
    assert prob.tuning()["tuned_stage_ms"] > 0.0  # the autotuner ran
    X = check_bond_set(prob, orc, "plain")
    pyoracle.set_square_mode(True)
    try:
        check_forces(prob, orc, X, 1e-12, "plain fp64")
    finally:
        pyoracle.set_square_mode(False)
    check_rk4_oracle(prob, orc, "plain")
    run_precisions(spec, prob, "plain")


def test_slit_bar_fp64_against_oracle(slit):
    spec, orc = slit
    prob = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64)
    X = check_bond_set(prob, orc, "slit")
    assert prob.csr()[2].sum() > 0
    pyoracle.set_square_mode(True)
    try:
        check_forces(prob, orc, X, 1e-12, "slit fp64")
    finally:
        pyoracle.set_square_mode(False)
    check_rk4_oracle(prob, orc, "slit")
    run_precisions(spec, prob, "slit")


@pytest.mark.parametrize("precision", [pm.PDM_FP32C, pm.PDM_FP32C2])
def test_slit_bar_fp32_forces_autotuned(slit, precision):
    spec, orc = slit
    prob = pm.Problem(pm.deck_text(spec), precision=precision)
    X = check_bond_set(prob, orc, f"slit precision={precision}")
    print(f"[full] slit precision={precision} stage launch {prob.stage_launch()}", flush=True)
    check_forces(prob, orc, X, 1e-6, f"slit precision={precision}")


def test_slit_bar_fp32c_three_tiles_per_cu(slit):
    """Tile edge 4, 8 wavefronts, 4 families per wavefront: a tile's LDS leaves room for three
    per CU, so the launches take the 80-VGPR stage kernel; its forces and its 2,000-step
    trajectory keep the FP32C bars."""
    spec, orc = slit
    prob = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP32C, tile_edge=4, waves_per_tile=8,
                      families_per_wave=4)
    launch = prob.stage_launch()
    print(f"[full] slit fp32c layout 4/8/4: {launch}", flush=True)
    assert launch["max_vgprs"] == 80 and 3 * launch["lds_bytes"] <= 160 * 1024
    X, _, _ = prob.nodes()
    check_forces(prob, orc, X, 1e-6, "slit fp32c 4/8/4")
    p64 = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64)
    p64.step(STEPS)
    prob.step(STEPS)
    U64, _ = p64.state()
    U32, _ = prob.state()
    err = rel(U32, U64)
    print(f"[full] slit fp32c 4/8/4 after {STEPS} steps: err vs fp64 {err:.3e}", flush=True)
    assert err <= 1e-6


def test_graph_remainder_chunks(slit):
    """Step counts that are not a multiple of the graph chunk run as captured graphs of the
    remainder too, bit-identical to eager launches."""
    spec, _ = slit
    a = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP32C, graph_chunk=50)
    b = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP32C, use_graph=0,
                   **{k: v for k, v in zip(("tile_edge", "waves_per_tile", "families_per_wave"),
                                           (a.tuning()["tile_edge"], a.tuning()["waves_per_tile"],
                                            a.tuning()["families_per_wave"]))})
    for n in (20, 7, 50, 20):
        a.step(n)
        b.step(n)
    Ua, Va = a.state()
    Ub, Vb = b.state()
    print(f"[full] graph remainder chunks: step {a.clock()[1]}", flush=True)
    assert np.array_equal(Ua, Ub) and np.array_equal(Va, Vb)
    assert a.energy() == b.energy()
