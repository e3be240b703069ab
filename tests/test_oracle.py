"""Pins the CPU oracle against facts that do not come from its own code.

Each check is chosen so that a plausible slip in the restatement (dropped term,
wrong sign or index, transposed operand) fails: closed forms, brute-force
geometry, conservation/derivative identities, and counts fixed by the
reference's bond criterion.  Runs on the CPU only.
"""

import math

import numpy as np
import pytest

import pdmtel_amd as pm
import pyoracle

DX = 1e-3


def small_bar(**kw):
    return pm.bar_spec(kw.pop("nx", 10), kw.pop("ny", 6), kw.pop("nz", 6), **kw)


# ---------------------------------------------------------------- lattice / regions
def test_lattice_closed_form():
    spec = small_bar(nx=7, ny=5, nz=4)
    o = pyoracle.Oracle(spec)
    X, mat, bc = o.nodes()
    assert o.n == 7 * 5 * 4
    ids = np.arange(o.n)
    l = ids % 7 + 1
    m = (ids // 7) % 5 + 1
    n = ids // 35 + 1
    cx, cy, cz = spec["grid_center"]
    # pdDataManager.cpp:852-854: (l - (nx+1)/2) dx + centre, x fastest
    assert np.array_equal(X[:, 0], (l - 8 * 0.5) * DX + cx)
    assert np.array_equal(X[:, 1], (m - 6 * 0.5) * DX + cy)
    assert np.array_equal(X[:, 2], (n - 5 * 0.5) * DX + cz)
    # x = 0 layer carries bc 0 (fixed), the x = L layer bc 1 (traction)
    assert np.all(bc[l == 1] == 0) and np.all(bc[l == 7] == 1) and np.all(bc[(l > 1) & (l < 7)] == -1)


def test_deletion_and_last_space_wins():
    spec = small_bar(nx=6, ny=6, nz=6)
    spec["deletion_spaces"] = [dict(geo=1, box=(0.0, 0.0021, 0.0, 0.0021, 0.0, 0.0021)),
                               dict(geo=2, cyl=(0.0012, 0.006, 0.006, 0.0, 0.0021))]
    spec["materials"].append(dict(type=1, E=90e9, sigy=1e30, ecrit=1e10, rho=2000.0))
    # second material space covers everything: it decides every node (pdDataManager.cpp:918-968)
    spec["material_spaces"] = [dict(geo=1, ref=1, box=(-1, 1, -1, 1, -1, 1)),
                               dict(geo=1, ref=2, box=(-1, 1, -1, 1, -1, 1))]
    o = pyoracle.Oracle(spec)
    X, mat, bc = o.nodes()
    # cube deletes nodes with all coordinates <= 2.1 mm (2 x 2 x 2 = 8 nodes); cylinder of radius
    # 1.2 mm around (6, 6) mm deletes the corner nodes at (5.5, 5.5) mm for z <= 2.1 mm (2 nodes)
    assert o.n == 216 - 8 - 2
    assert np.all(mat == 1)
    # a node outside the LAST material space is fatal even if an earlier space holds it
    spec["material_spaces"][1]["box"] = (-1, 0.003, -1, 1, -1, 1)
    with pytest.raises(RuntimeError):
        pyoracle.Oracle(spec)


# ---------------------------------------------------------------- bond predicate
def true_cube_distance(xi, xj, dx):
    d = np.maximum(np.abs(np.asarray(xj) - np.asarray(xi)) - 0.5 * dx, 0.0)
    return float(np.sqrt((d * d).sum()))


def test_pair_dmin_is_point_to_cube_distance():
    """CalcDistanceToCube's dmin equals the geometric point-to-cell distance."""
    spec = small_bar()
    prob = pyoracle.problem_from_spec(spec)
    L = pyoracle.lib()
    rng = np.random.default_rng(1)
    C = pyoracle.C
    worst = 0.0
    near_tol = 0
    for _ in range(4000):
        xi = rng.uniform(-0.01, 0.01, 3)
        xj = xi + rng.uniform(-4.6, 4.6, 3) * DX
        dmin, dmax = C.c_double(), C.c_double()
        a = np.ascontiguousarray(xi)
        b = np.ascontiguousarray(xj)
        ok = L.orc_pair_bonded(C.byref(prob), pyoracle._dp(a), pyoracle._dp(b), C.byref(dmin), C.byref(dmax))
        d = np.linalg.norm(xj - xi)
        if prob.horizon < d - 1.5 * DX:
            assert ok == 0
            continue
        if np.all(np.abs(xj - xi) < 0.5 * DX + 1e-6 * DX):
            continue  # node i inside node j's cell: the reference measures to a face; no lattice pair does this
        t = true_cube_distance(xi, xj, DX)
        err = abs(dmin.value - t)
        # face tolerance 1e-4 dx can only move dmin by a few tolerances
        assert err <= 3e-4 * DX
        if err > 1e-12 * DX:
            near_tol += 1
        worst = max(worst, err)
        if abs(t - (prob.horizon + 1e-4 * DX)) > 5e-4 * DX:
            assert bool(ok) == (t < prob.horizon + 1e-4 * DX)
    assert near_tol < 40


@pytest.mark.parametrize("delta,expected", [(3.015, 250), (4.015, 460)])
def test_interior_family_and_max_family(delta, expected):
    spec = small_bar(nx=13, ny=13, nz=13, delta_over_dx=delta)
    o = pyoracle.Oracle(spec)
    rp, col, cut = o.csr()
    fam = np.diff(rp)
    assert o.max_family == expected
    centre = 6 + 13 * 6 + 169 * 6
    assert fam[centre] == expected
    # the interior family is exactly the lattice offsets whose cell lies within delta
    offs = np.array([(a, b, c) for a in range(-5, 6) for b in range(-5, 6) for c in range(-5, 6) if (a, b, c) != (0, 0, 0)])
    d = np.sqrt((np.maximum(np.abs(offs) - 0.5, 0) ** 2).sum(axis=1))
    assert (d < delta + 1e-4).sum() == expected


def test_reference_bond_total_2000_node_bar():
    o = pyoracle.Oracle(pm.bar_spec(20, 10, 10))
    assert o.E // 2 == 168392


def test_brute_force_equals_cells_and_symmetry():
    spec = small_bar(nx=9, ny=7, nz=6, jitter=0.01, seed=5)
    spec["deletion_spaces"] = [dict(geo=2, cyl=(0.0015, 0.0045, 0.0035, 0.0, 0.01))]
    a = pyoracle.Oracle(spec, bond_method=0)
    b = pyoracle.Oracle(spec, bond_method=1)
    rpa, cola, _ = a.csr()
    rpb, colb, _ = b.csr()
    assert np.array_equal(rpa, rpb) and np.array_equal(cola, colb)
    rows = np.repeat(np.arange(a.n), np.diff(rpa))
    pairs = set(zip(rows.tolist(), cola.tolist()))
    assert all((j, i) in pairs for i, j in pairs)
    for i in range(a.n):
        r = cola[rpa[i]:rpa[i + 1]]
        assert np.all(np.diff(r) > 0)


def test_sampled_rows_equal_whole_csr():
    """The row sampler (families of selected nodes only, for problems too large to build whole)
    returns exactly the rows, cut flags and bdryEff factors of the whole-problem oracle, which
    the O(N^2) reference loop pins above; the whole-problem search also agrees for any thread
    count."""
    spec = small_bar(nx=14, ny=9, nz=7, jitter=0.01, seed=2, delta_over_dx=4.015)
    spec["slits"] = [pm.mid_slit(spec, 30.0), dict(type=1, box=(0.0, 1.0, 0.004, 0.004, 0.0, 1.0))]
    whole = pyoracle.Oracle(spec, bond_method=1)
    serial = pyoracle.Oracle(spec, bond_method=1, nthreads=1)
    rp, col, cut = whole.csr()
    rps, cols, cuts = serial.csr()
    assert np.array_equal(rp, rps) and np.array_equal(col, cols) and np.array_equal(cut, cuts)
    assert np.array_equal(whole.bdryeff(), serial.bdryeff()) and whole.min_bond == serial.min_bond
    be = whole.bdryeff()
    rows = np.random.default_rng(1).permutation(whole.n)[:150]
    rows = np.concatenate([rows, [0, whole.n - 1]])
    srp, scol, scut, sbe = pyoracle.sample_rows(spec, rows)
    assert cut.sum() > 0 and (be != 1.0).any()
    for k, i in enumerate(rows):
        assert np.array_equal(scol[srp[k]:srp[k + 1]], col[rp[i]:rp[i + 1]])
        assert np.array_equal(scut[srp[k]:srp[k + 1]], cut[rp[i]:rp[i + 1]])
    assert np.array_equal(sbe, be[rows])
    u = np.random.default_rng(3).uniform(-1e-6, 1e-6, size=(whole.n, 3))
    F, el = whole.forces(u, with_elast=True)
    sF, sel = pyoracle.sample_forces(spec, rows, u)
    assert np.array_equal(sF, F[rows]) and np.array_equal(sel, el[rows])


# ---------------------------------------------------------------- material / beta / time
def test_spring_constant_and_dt_closed_form():
    spec = small_bar()
    o = pyoracle.Oracle(spec)
    E, rho, delta = 72e9, 2440.0, 3.015 * DX
    assert math.isclose(o.spring(0), 12 * E / (math.pi * delta ** 4), rel_tol=1e-14)
    c = math.sqrt(E / 1.5 / rho)
    assert math.isclose(o.min_bond, DX, rel_tol=1e-12)
    assert math.isclose(o.dt, 0.8 * DX / c, rel_tol=1e-12)


def test_beta_closed_form():
    L = pyoracle.lib()
    d = 3.015
    for mdf, outer in ((1, d + 0.5), (0, d)):
        for r in np.linspace(2.0, 4.0, 81):
            b = L.orc_beta(r, d, 1.0, 1, mdf)
            if r <= d - 0.5:
                assert b == 1.0
            elif r <= outer:
                assert math.isclose(b, d + 0.5 - r, rel_tol=1e-13)
            else:
                assert b == 0.0


def test_time_factor_closed_form():
    L = pyoracle.lib()
    bc = pyoracle._BC()
    bc.t_end = 2.0
    for typ, load, t, want in ((1, 1.0, 0.5, 0.25), (1, 1.0, 3.0, 1.0), (1, 2.0, 0.5, 1.0), (3, 1.0, 1.0, 0.5),
                               (2, 1.0, 1.5, 1.0), (2, 1.0, 2.5, 0.0)):
        bc.type, bc.load = typ, load
        assert L.orc_time_factor(pyoracle.C.byref(bc), t) == want


# ---------------------------------------------------------------- boundary-effect factor
def test_bdryeff_closed_forms():
    spec = small_bar(nx=14, ny=14, nz=14)
    o = pyoracle.Oracle(spec)
    be = o.bdryeff()
    X, mat, _ = o.nodes()
    rp, col, _ = o.csr()
    fam = np.diff(rp)
    E, delta, vol = 72e9, 3.015 * DX, DX ** 3
    c = 12 * E / (math.pi * delta ** 4)
    assert np.all(be[fam == 250] == 1.0)
    lo, hi = 0.0, 14 * DX
    ns = 3 * DX
    seen_one = seen_energy = 0
    for i in np.nonzero(fam != 250)[0]:
        x = X[i]
        inner = [(x[k] >= lo + ns - 1e-12 and x[k] <= hi - ns + 1e-12) for k in range(3)]
        if (x[0] <= lo + ns and inner[1] and inner[2]):
            h = x[0] - lo
            want = 3 * E / (3 * E - 0.5 * c * math.pi * (delta ** 4 / 4 - h * delta ** 3 / 3 + h ** 4 / 12))
            assert math.isclose(be[i], want, rel_tol=1e-13)
            seen_one += 1
        elif sum(inner) <= 1:
            # energy branch: uniform strain e gives extension e r, so p_grid = sum c e^2 r beta V / 4
            # and the factor is 12 E / (c V sum r beta)
            nb = col[rp[i]:rp[i + 1]]
            r = np.linalg.norm(X[nb] - x, axis=1)
            beta = np.where(r <= delta - 0.5 * DX, 1.0, np.where(r <= delta + 0.5 * DX, (delta + 0.5 * DX - r) / DX, 0.0))
            want = 12 * E / (c * vol * np.sum(r * beta))
            assert math.isclose(be[i], want, rel_tol=1e-8)
            seen_energy += 1
    assert seen_one > 0 and seen_energy > 0


# ---------------------------------------------------------------- forces
def test_force_two_node_closed_form():
    prob = pyoracle.problem_from_spec(small_bar())
    L = pyoracle.lib()
    C = pyoracle.C
    c, beff = 3.3e21, 1.25
    for r0, du in ((1.0, 1e-7), (2.3, -3e-8), (3.4, 5e-8)):
        xi = np.array([r0 * DX, 0.0, 0.0])
        eta = np.array([du, 0.0, 0.0])
        f = np.zeros(3)
        el = C.c_double()
        L.orc_bond_force_cci(C.byref(prob), pyoracle._dp(xi), pyoracle._dp(eta), 1e10, c, beff, 2.0, pyoracle._dp(f),
                             C.byref(el))
        r = r0 * DX
        beta = 1.0 if r <= 3.015 * DX - 0.5 * DX else (3.015 * DX + 0.5 * DX - r) / DX
        want = beff * DX ** 3 * beta * c * (du / r)
        assert math.isclose(f[0], want, rel_tol=1e-9) and f[1] == 0.0 and f[2] == 0.0
        assert math.isclose(el.value, beff * 0.25 * beta * c * du * du / r, rel_tol=1e-9)


def test_force_invariants():
    spec = small_bar(nx=8, ny=6, nz=5, jitter=0.01, seed=2)
    spec["bdryeff_on"] = 0
    o = pyoracle.Oracle(spec)
    rng = np.random.default_rng(3)
    assert np.all(o.forces(np.zeros((o.n, 3))) == 0.0)
    assert np.all(o.forces(np.tile(rng.normal(size=3) * 1e-6, (o.n, 1))) == 0.0)
    u = rng.uniform(-1e-6, 1e-6, size=(o.n, 3))
    F = o.forces(u)
    # pairwise antisymmetry with symmetric coefficients: total force vanishes
    assert np.abs(F.sum(axis=0)).max() <= 1e-10 * np.abs(F).sum()


def test_force_is_minus_energy_gradient():
    spec = small_bar(nx=6, ny=5, nz=5, jitter=0.01, seed=4)
    o = pyoracle.Oracle(spec)
    rng = np.random.default_rng(5)
    u = rng.uniform(-2e-6, 2e-6, size=(o.n, 3))
    F = o.forces(u)
    vol = DX ** 3
    for _ in range(12):
        k, c = int(rng.integers(o.n)), int(rng.integers(3))
        h = 1e-9
        up, um = u.copy(), u.copy()
        up[k, c] += h
        um[k, c] -= h
        _, ep = o.forces(up, with_elast=True)
        _, em = o.forces(um, with_elast=True)
        grad = (ep.sum() - em.sum()) * vol / (2 * h)
        assert math.isclose(-grad, F[k, c], rel_tol=2e-5, abs_tol=1e-6 * np.abs(F).max())


def test_tangent_is_force_linearisation():
    spec = small_bar(nx=7, ny=6, nz=5, jitter=0.01, seed=6)
    o = pyoracle.Oracle(spec)
    rng = np.random.default_rng(7)
    u = rng.uniform(-1.0, 1.0, size=(o.n, 3))
    Ku = o.tangent(u)
    for eps in (1e-9, 1e-10):
        F = o.forces(eps * u)
        err = np.abs(F / eps + Ku).max() / np.abs(Ku).max()
        assert err < 5e-4
    # symmetric operator: v.Ku == u.Kv
    v = rng.uniform(-1.0, 1.0, size=(o.n, 3))
    assert math.isclose(np.vdot(v, Ku), np.vdot(u, o.tangent(v)), rel_tol=1e-10)


# ---------------------------------------------------------------- time integration
def test_rk4_rigid_motion_advances_three_quarters():
    """No bonds strain under a uniform velocity: each step U += 0.75 V dt (reference quirk)."""
    spec = small_bar(nx=6, ny=5, nz=5)
    spec["boundary_spaces"] = []
    spec["bcs"] = []
    spec["traction_body"] = 0
    o = pyoracle.Oracle(spec)
    V0 = np.tile([3.0, -1.0, 2.0], (o.n, 1))
    o.set_state(np.zeros((o.n, 3)), V0)
    o.run(3)
    U, V = o.state()
    assert np.allclose(V, V0, rtol=0, atol=1e-12)
    assert np.allclose(U, 3 * 0.75 * o.dt * V0, rtol=1e-12, atol=0)


def test_rk4_matches_amplification_matrix():
    """A linearised node under small loads follows the scheme's amplification matrix.

    For a(u) = -w2 u the reference's stage algebra (velocity buffer accumulated over
    stages with stage steps dt/2, dt/2, dt, dt; U += (k1 + 2k2 + 2k3 + k4)/6) is the
    linear map built below from its definition; a 2-node bar with one node fixed
    and a tiny displacement obeys it to linearisation accuracy.
    """
    spec = small_bar(nx=2, ny=1, nz=1, delta_over_dx=1.2)
    spec["bcs"] = [dict(type=1, flags=(1, 1, 1), value=(0.0, 0.0, 0.0), t_end=1.0, load=2)]
    spec["boundary_spaces"] = [dict(geo=1, ref=1, box=(-1, 0.001, -1, 1, -1, 1))]
    spec["traction_body"] = 0
    spec["bdryeff_on"] = 0
    o = pyoracle.Oracle(spec)
    assert o.n == 2 and o.E == 2
    u0 = 1e-9
    U = np.zeros((2, 3))
    U[1, 0] = u0
    o.set_state(U, np.zeros((2, 3)))
    # stiffness of the free node along x: F = -w2 * rho * u
    F = o.forces(U)
    w2 = -F[1, 0] / u0 / 2440.0
    dt = o.dt

    def stage_map(h):
        # returns M with [u_new, v_new] = M [u, v] for one RK4 step of a(u) = -w2 u
        cols = []
        for uu, vv in ((1.0, 0.0), (0.0, 1.0)):
            vb = vv
            vb += dt / 2 * (-w2 * uu); k1 = dt / 2 * vb
            vb += dt / 2 * (-w2 * (uu + k1 / 2)); k2 = dt / 2 * vb
            vb += dt * (-w2 * (uu + k2 / 2)); k3 = dt * vb
            vb += dt * (-w2 * (uu + k3)); k4 = dt * vb
            cols.append([uu + (k1 + 2 * k2 + 2 * k3 + k4) / 6, vb])
        return np.array(cols).T

    M = stage_map(dt)
    x = np.array([u0, 0.0])
    for _ in range(5):
        o.run(1)
        x = M @ x
        Ux, Vx = o.state()
        assert math.isclose(Ux[1, 0], x[0], rel_tol=1e-6)
        assert math.isclose(Vx[1, 0], x[1], rel_tol=1e-6)


def test_driver_loop_shortened_last_step():
    spec = small_bar(nx=4, ny=3, nz=3, steps=10)
    o0 = pyoracle.Oracle(spec)
    dt = o0.dt
    spec["total_time"] = 2.5 * dt
    o = pyoracle.Oracle(spec)
    o.run(10)
    # steps at t = dt, 2 dt, then a shortened step to the end time; then the loop stops
    assert o.steps_done == 3
    assert math.isclose(o.t, 2.5 * dt, rel_tol=1e-15)


# ---------------------------------------------------------------- slits
def seg_rect(xi, xj, c, n, t, ht, hw):
    """Independent segment / oriented-rectangle test: solve xi + s (xj - xi) = c + a t + b w."""
    n = np.asarray(n) / np.linalg.norm(n)
    t = np.asarray(t) - np.dot(t, n) * n
    t = t / np.linalg.norm(t)
    w = np.cross(n, t)
    A = np.column_stack([np.asarray(xj) - np.asarray(xi), -t, -w])
    try:
        s, a, b = np.linalg.solve(A, np.asarray(c) - np.asarray(xi))
    except np.linalg.LinAlgError:
        return None
    return (0 < s < 1) and abs(a) <= ht and abs(b) <= hw, s, a, b


def test_slit_cut_bruteforce():
    L = pyoracle.lib()
    C = pyoracle.C
    rng = np.random.default_rng(11)
    spec = small_bar()
    checked = 0
    for trial in range(300):
        c = rng.uniform(0, 0.01, 3)
        n = rng.normal(size=3)
        t = rng.normal(size=3)
        ht, hw = rng.uniform(0.001, 0.004, 2)
        spec["slits"] = [dict(type=2, c=tuple(c), n=tuple(n), t=tuple(t), half_t=ht, half_s=hw)]
        prob = pyoracle.problem_from_spec(spec)
        xi = c + rng.uniform(-0.004, 0.004, 3)
        xj = xi + rng.uniform(-0.004, 0.004, 3)
        res = seg_rect(xi, xj, c, n, t, ht, hw)
        if res is None:
            continue
        inside, s, a, b = res
        # skip pairs within rounding of an edge or an endpoint
        if min(abs(s), abs(1 - s), abs(abs(a) - ht), abs(abs(b) - hw)) < 1e-9:
            continue
        got = L.orc_pair_cut(C.byref(prob), pyoracle._dp(np.ascontiguousarray(xi)), pyoracle._dp(np.ascontiguousarray(xj)))
        assert bool(got) == inside
        checked += 1
    assert checked > 250


def test_axis_slit_matches_reference_plane_test():
    """Type 1 (axis-normal box) is the reference's commented crossing test (pdDataManager.cpp:1246-1290)."""
    L = pyoracle.lib()
    C = pyoracle.C
    rng = np.random.default_rng(12)
    spec = small_bar()
    box = (0.002, 0.008, 0.005, 0.005, 0.001, 0.004)
    spec["slits"] = [dict(type=1, box=box)]
    prob = pyoracle.problem_from_spec(spec)
    for _ in range(500):
        xi = rng.uniform(0, 0.01, 3)
        xj = xi + rng.uniform(-0.004, 0.004, 3)
        yp = 0.005
        crosses = (xi[1] - yp) * (xj[1] - yp) < 0
        if crosses:
            f = (yp - xi[1]) / (xj[1] - xi[1])
            p1, p3 = xi[0] + f * (xj[0] - xi[0]), xi[2] + f * (xj[2] - xi[2])
            crosses = (p1 - box[0]) * (p1 - box[1]) <= 0 and (p3 - box[4]) * (p3 - box[5]) <= 0
        got = L.orc_pair_cut(C.byref(prob), pyoracle._dp(xi), pyoracle._dp(np.ascontiguousarray(xj)))
        assert bool(got) == crosses


# ---------------------------------------------------------------- fixed Gaussian integration
def fgi_spec(spec, n):
    s = dict(spec)
    s["cci"] = None
    s["fgi"] = (1, n)
    return s


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8])
def test_gauss_rules_match_legendre(n):
    a, w = pyoracle.gauss_rule(n)
    x, v = np.polynomial.legendre.leggauss(n)
    order = np.argsort(a)
    assert np.allclose(a[order], x, rtol=0, atol=2e-15)
    assert np.allclose(w[order], v, rtol=0, atol=2e-15)
    # the reference's point order: zero first for odd n, +/- pairs from the centre out
    # (n = 8: the four positive points, then the four negative ones)
    if n == 8:
        assert np.all(np.diff(a[:4]) > 0) and np.array_equal(a[4:], -a[:4])
    elif n > 1:
        pos = a[n % 2::2]
        assert np.all(pos > 0) and np.all(np.diff(pos) > 0) and np.array_equal(a[n % 2 + 1::2], -pos)
    # exact for x^(2n-1) and x^(2n-2) on [-1, 1]
    assert abs(np.dot(w, a ** (2 * n - 1))) < 1e-14
    assert abs(np.dot(w, a ** (2 * n - 2)) - 2.0 / (2 * n - 1)) < 1e-14


def test_gauss_rule_orders_without_reference_table():
    for n in (0, 7, 9):
        with pytest.raises(ValueError):
            pyoracle.gauss_rule(n)


def test_gauss_point_positions():
    spec = fgi_spec(small_bar(nx=5, ny=4, nz=4, jitter=0.01, seed=1), 3)
    o = pyoracle.Oracle(spec)
    X, _, _ = o.nodes()
    Xg, Ug = o.gauss()
    x, _ = np.polynomial.legendre.leggauss(3)
    want = X[:, None, :] + 0.5 * DX * np.stack(np.meshgrid(x, x, x, indexing="ij"), -1).reshape(-1, 3)[None]
    for i in range(o.n):
        got = np.array(sorted(map(tuple, Xg[i])))
        ref = np.array(sorted(map(tuple, want[i])))
        assert np.allclose(got, ref, rtol=0, atol=1e-18)
    assert np.all(Ug == 0.0)


def test_fgi_one_point_is_modified_cci():
    """One Gauss point (weight 2 per axis) is the modified CCI rule: same forces, same steps."""
    pyoracle.set_square_mode(True)
    try:
        base = small_bar(nx=8, ny=6, nz=5, jitter=0.01, seed=2, steps=20)
        oc, o1 = pyoracle.Oracle(base), pyoracle.Oracle(fgi_spec(base, 1))
        u = np.random.default_rng(0).uniform(-1e-6, 1e-6, size=(oc.n, 3))
        Fc, ec = oc.forces(u, with_elast=True)
        F1, e1 = o1.forces(u, with_elast=True)
        assert np.abs(F1 - Fc).max() <= 1e-14 * np.abs(Fc).max()
        assert np.abs(e1 - ec).max() <= 1e-14 * np.abs(ec).max()
        oc.run(5)
        o1.run(5)
        Uc, Vc = oc.state()
        U1, V1 = o1.state()
        # the per-bond association (dx^3 / 8 * 8 outside the force) differs in the last bits; this
        # bar's step load amplifies a 1e-16 perturbation of the state to ~1e-13 within five steps
        assert np.abs(U1 - Uc).max() <= 1e-11 * np.abs(Uc).max()
        assert np.abs(V1 - Vc).max() <= 1e-11 * np.abs(Vc).max()
    finally:
        pyoracle.set_square_mode(False)


def quartic(d):
    return np.where(d > 1.0, 0.0, 1 - 6 * d ** 2 + 8 * d ** 3 - 3 * d ** 4)


@pytest.mark.parametrize("n", [2, 3])
def test_mls_reproduces_quadratics_and_matches_lstsq(n):
    spec = fgi_spec(small_bar(nx=6, ny=5, nz=5, jitter=0.02, seed=3), n)
    rw = 3.0 * DX
    o = pyoracle.Oracle(spec)
    X, _, _ = o.nodes()
    Xg, _ = o.gauss()
    rp, cb, phi = o.mls()
    ng = n ** 3
    rng = np.random.default_rng(4)
    worst = 0.0
    for i in range(o.n):
        k = cb[rp[i]:rp[i + 1]]
        assert k[0] == i and len(k) >= 10
        rows = phi[ng * rp[i]:ng * rp[i + 1]].reshape(ng, len(k))
        d = (X[k][None, :, :] - Xg[i][:, None, :]) / DX            # (ng, m, 3)
        worst = max(worst, np.abs(rows.sum(1) - 1).max())
        for a in range(3):
            worst = max(worst, np.abs((rows * d[..., a]).sum(1)).max())
            for b in range(a, 3):
                worst = max(worst, np.abs((rows * d[..., a] * d[..., b]).sum(1)).max())
        if i in rng.choice(o.n, 6, replace=False):
            # weighted least squares through numpy: phi = W P (P^T W P)^-1 p(x_g)
            for g in range(ng):
                w = quartic(np.linalg.norm(X[k] - Xg[i, g], axis=1) / rw)
                q = (X[k] - Xg[i, g]) / DX
                P = np.column_stack([np.ones(len(k)), q, q[:, 0] * q[:, 1], q[:, 0] * q[:, 2], q[:, 1] * q[:, 2],
                                     q ** 2])
                sw = np.sqrt(w)
                # min || sqrt(W) (P c - e_k) || per unit vector: phi_k = (P (P^T W P)^-1 P^T W)_0k at q = 0
                coef = np.linalg.lstsq(sw[:, None] * P, np.diag(sw), rcond=None)[0]
                assert np.allclose(coef[0], rows[g], rtol=0, atol=1e-12)
    assert worst < 1e-10


@pytest.mark.parametrize("n", [2, 3])
def test_fgi_patch_test_and_translation(n):
    """Regular lattice: an affine field leaves a full-family node unloaded (lattice symmetry,
    exact MLS for affine fields); a rigid translation loads nobody."""
    spec = fgi_spec(small_bar(nx=9, ny=9, nz=9), n)
    spec["bdryeff_on"] = 0
    o = pyoracle.Oracle(spec)
    X, _, _ = o.nodes()
    rng = np.random.default_rng(5)
    u = (X - X.mean(0)) @ (rng.normal(size=(3, 3)) * 1e-4).T
    F = o.forces(u)
    scale = np.abs(F).max()
    rp, _, _ = o.csr()
    full = np.flatnonzero(np.diff(rp) == o.max_family)
    assert len(full) > 0 and np.abs(F[full]).max() <= 1e-11 * scale
    Ft = o.forces(np.tile(rng.normal(size=3) * 1e-6, (o.n, 1)))
    assert np.abs(Ft).max() <= 1e-11 * scale


def test_fgi_converges_in_number_of_points():
    """Uniform expansion: the family sums of boundary nodes converge as the rule grows, and keep
    the lattice's mirror symmetry."""
    base = small_bar(nx=10, ny=9, nz=9)
    base["bdryeff_on"] = 0
    X = None
    res = {}
    for n in (2, 3, 4, 5, 6, 8):
        o = pyoracle.Oracle(fgi_spec(base, n))
        if X is None:
            X, _, _ = o.nodes()
        res[n] = o.forces(1e-4 * (X - X.mean(0)))
    corner = 0
    face = int(np.argmin(np.abs(X[:, 0] - X[:, 0].min()) + np.linalg.norm(X[:, 1:] - X[:, 1:].mean(0), axis=1)))
    ref = res[8]
    for n, F in res.items():
        tol = 1.5e-2 if n == 2 else 6e-3
        for i in (corner, face):
            assert np.linalg.norm(F[i] - ref[i]) <= tol * np.linalg.norm(ref[i]), (n, i)
        assert np.abs(F[face, 1:]).max() <= 1e-12 * abs(F[face, 0])
        assert np.ptp(F[corner]) <= 1e-12 * np.abs(F[corner]).max()


def test_fgi_gauss_points_follow_state_and_bcs():
    """After a step: BC nodes' points hold the prescribed components (the rest keep 0); the other
    points interpolate U with the MLS coefficients."""
    spec = fgi_spec(small_bar(nx=8, ny=5, nz=5, jitter=0.01, seed=6, steps=10), 2)
    spec["bcs"][0] = dict(type=1, flags=(1, 0, 1), value=(2e-7, 0.0, -1e-7), t_end=1.0, load=1)
    o = pyoracle.Oracle(spec)
    o.run(2)
    _, _, bc = o.nodes()
    U, _ = o.state()
    _, Ug = o.gauss()
    rp, cb, phi = o.mls()
    fac = o.t / 1.0  # ramp load
    bcn = bc == 0
    assert bcn.sum() > 0
    assert np.allclose(Ug[bcn][..., 0], 2e-7 * fac, rtol=1e-15, atol=0)
    assert np.all(Ug[bcn][..., 1] == 0.0)
    assert np.allclose(Ug[bcn][..., 2], -1e-7 * fac, rtol=1e-15, atol=0)
    for i in np.flatnonzero(bc < 0)[:20]:
        k = cb[rp[i]:rp[i + 1]]
        rows = phi[8 * rp[i]:8 * rp[i + 1]].reshape(8, len(k))
        assert np.allclose(Ug[i], rows @ U[k], rtol=1e-14, atol=1e-24)
