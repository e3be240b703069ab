"""Pinning of the oracle's adaptive integration (ai_on 1) by mathematics.

The reference's adaptive integration (pdIntgManager.cpp:49-451, 523-3482; CalcBondForceIntg,
pdDataManager2.cpp:1222-1291) does not build and has no recorded output, so its restatement
is checked against what geometry and the integrand fix:

- a cell wholly inside the horizon is the tensor trapezoidal rule on the cell (closed form);
- a clipped cell's weights converge to the volume of cell and ball, computed independently by
  adaptive quadrature in the opposite integration order (z outer, x inner);
- the integrated bond force of a cell under an affine field converges to an independent
  adaptive cubature of the same integral;
- an interior node's family under an affine field is in equilibrium (mirror symmetry);
- the MLS coefficients at the points reproduce quadratic fields, so an affine field through
  MLS gives the same forces as the prescribed-gradient mode; zero and rigid fields give 0.
"""

import math
import os
import sys

import numpy as np
import pytest
from scipy import integrate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

import pyoracle  # noqa: E402
from pdmtel_amd import deck  # noqa: E402

DX = 1e-3
DELTA = 3.015 * DX


def spec_ai(nx=6, ny=5, nz=5, mls=1, n=2, jitter=0.0, gradient=None):
    spec = deck.bar_spec(nx, ny, nz, steps=10, jitter=jitter, seed=0)
    spec["cci"] = None
    spec["ai"] = (1, mls, n)
    if gradient is not None:
        # every node carries the same DISPLGRAD condition (ai_on 1 0 requires one on each node)
        spec["bcs"] = [dict(type=3, flags=(1, 1, 1), value=tuple(np.asarray(gradient).reshape(-1)), t_end=1.0,
                            load=2)]
        spec["boundary_spaces"] = [dict(geo=1, ref=1, box=(-1e3, 1e3, -1e3, 1e3, -1e3, 1e3))]
        spec["traction_body"] = 0
    return spec


def cell_ball_volume(c, h, delta):
    """vol([c - h, c + h]^3 intersected with the ball |x| <= delta), z outer, x inner."""
    lo = np.asarray(c) - h
    hi = np.asarray(c) + h

    def length_x(y, z):
        s2 = delta * delta - y * y - z * z
        if s2 <= 0.0:
            return 0.0
        s = math.sqrt(s2)
        return max(0.0, min(hi[0], s) - max(lo[0], -s))

    v, _ = integrate.dblquad(lambda y, z: length_x(y, z), lo[2], hi[2], lo[1], hi[1], epsabs=1e-16, epsrel=1e-11)
    return v


def cell_force(c, h, delta, G):
    """Integral over cell and ball of s (xi + G xi) / p, s = (p - r) / r (c = bdryEff = 1)."""
    lo = np.asarray(c) - h
    hi = np.asarray(c) + h
    G = np.asarray(G).reshape(3, 3)

    def f(x, y, z, k):
        xi = np.array([x, y, z])
        a = xi + G @ xi
        r = math.sqrt(xi @ xi)
        p = math.sqrt(a @ a)
        return (p - r) / r * a[k] / p

    def xlim(y, z):
        s2 = delta * delta - y * y - z * z
        s = math.sqrt(s2) if s2 > 0 else 0.0
        a, b = max(lo[0], -s), min(hi[0], s)
        return (a, b) if b > a else (a, a)

    out = []
    for k in range(3):
        v, _ = integrate.nquad(lambda x, y, z: f(x, y, z, k), [lambda y, z: xlim(y, z), [lo[1], hi[1]], [lo[2], hi[2]]],
                               opts={"epsabs": 1e-22, "epsrel": 1e-7, "limit": 50})
        out.append(v)
    return np.array(out)


def test_full_cell_is_tensor_trapezoid():
    """dmax <= delta: the cell taken whole (CalcConfigurationType3), 2^(n-1)+1 equally spaced
    points per direction, end points weighted 1/2."""
    spec = spec_ai()
    c = np.array([1.0, 1.0, 0.0]) * DX
    ok, dmin, dmax = pyoracle.pair_dists(spec, np.zeros(3), c)
    assert ok and dmax < DELTA
    for n in (1, 2, 3, 4):
        spec["ai"] = (1, 1, n)
        xw = pyoracle.ai_cell_points(spec, c, dmin, dmax)
        m = 2 ** (n - 1) + 1
        t = np.linspace(-0.5, 0.5, m) * DX
        w1 = np.full(m, DX / (m - 1))
        w1[[0, -1]] *= 0.5
        grid = np.stack(np.meshgrid(c[0] + t, c[1] + t, c[2] + t, indexing="ij"), -1).reshape(-1, 3)
        wgt = np.einsum("i,j,k->ijk", w1, w1, w1).reshape(-1)
        assert xw.shape == (m ** 3, 4)
        np.testing.assert_allclose(xw[:, :3], grid, rtol=0, atol=1e-18)
        np.testing.assert_allclose(xw[:, 3], wgt, rtol=1e-14)
        assert abs(xw[:, 3].sum() - DX ** 3) < 1e-14 * DX ** 3


def test_no_points_beyond_the_horizon():
    spec = spec_ai()
    c = np.array([4.0, 0.0, 0.0]) * DX
    _, dmin, dmax = pyoracle.pair_dists(spec, np.zeros(3), c)
    assert dmin >= DELTA
    assert pyoracle.ai_cell_points(spec, c, dmin, dmax).shape[0] == 0


# clipped cells of several configurations: on an axis, in a coordinate plane, off-plane, a
# corner barely inside, and a jittered centre
CLIPPED = [(3, 0, 0), (2, 2, 0), (3, 1, 0), (2, 2, 1), (1, 1, 3), (2, 2, 2), (3.004, 0.996, -0.003)]


@pytest.mark.parametrize("cell", CLIPPED)
def test_clipped_cell_volume_converges(cell):
    """Weights of a clipped cell sum to vol(cell and ball) in the limit; the error shrinks at
    every refinement and is below 1 % of a cell at n = 5."""
    spec = spec_ai()
    c = np.array(cell) * DX
    ok, dmin, dmax = pyoracle.pair_dists(spec, np.zeros(3), c)
    assert ok and dmin < DELTA < dmax
    exact = cell_ball_volume(c, 0.5 * DX, DELTA)
    assert exact > 0
    errs = []
    for n in (1, 2, 3, 4, 5):
        spec["ai"] = (1, 1, n)
        xw = pyoracle.ai_cell_points(spec, c, dmin, dmax)
        # every point lies in the cell and in the ball
        assert np.all(np.abs(xw[:, :3] - c) <= 0.5 * DX * (1 + 1e-12))
        assert np.all(np.linalg.norm(xw[:, :3], axis=1) <= DELTA * (1 + 1e-12))
        errs.append(abs(xw[:, 3].sum() - exact))
    assert errs[-1] < 1e-2 * DX ** 3, errs
    assert errs[-1] < errs[1], errs


@pytest.mark.parametrize("cell", [(3, 0, 0), (2, 2, 1), (1, 1, 1)])
def test_cell_force_converges_to_cubature(cell):
    """The integrated bond force of a cell under an affine field (eta = G xi) approaches an
    independent adaptive cubature of the integral as the trapezoidal index grows."""
    spec = spec_ai()
    G = np.array([[2e-4, 3e-5, 0.0], [3e-5, -5e-5, 1e-5], [0.0, 1e-5, -5e-5]])
    c = np.array(cell) * DX
    ok, dmin, dmax = pyoracle.pair_dists(spec, np.zeros(3), c)
    full = dmax <= DELTA
    ref = cell_force(c, 0.5 * DX, DELTA if not full else 10 * DELTA, G)
    scale = np.abs(ref).max()
    errs = []
    for n in (2, 3, 4, 5):
        spec["ai"] = (1, 1, n)
        F, el = pyoracle.ai_cell_force(spec, c, dmin, dmax, G)
        errs.append(np.abs(F - ref).max() / scale)
        assert el > 0
    # trapezoidal error on the sphere's edge falls about 3x per refinement
    assert errs[-1] < (1e-2 if not full else 2e-3), errs
    assert all(b < a for a, b in zip(errs, errs[1:])), errs


def test_interior_node_equilibrium_under_affine_field():
    """An interior node of a regular lattice under an affine field: its family's integrated
    forces cancel (mirror-symmetric point sets), while a single cell's force does not."""
    G = np.array([[1e-4, 2e-5, 0.0], [2e-5, -3e-5, 0.0], [0.0, 0.0, -3e-5]])
    spec = spec_ai(9, 9, 9, mls=0, n=2, gradient=G)
    o = pyoracle.Oracle(spec)
    F = o.forces(np.zeros((o.n, 3)))
    X, _, _ = o.nodes()
    i = 4 + 9 * 4 + 81 * 4
    assert np.allclose(X[i], [4.5e-3, 4.5e-3, 4.5e-3])
    Fc, _ = pyoracle.ai_cell_force(spec, np.array([3.0, 0.0, 0.0]) * DX, *pyoracle.pair_dists(spec, np.zeros(3),
                                   np.array([3.0, 0.0, 0.0]) * DX)[1:], G, spring=o.spring(0))
    assert np.abs(F[i]).max() < 1e-9 * np.abs(Fc).max()


def test_points_match_family_and_cells():
    """The state's point table: entry by entry, the points of the cell primitive (same
    distances), nothing for cut entries; point counts per entry at most (2^(n-1)+1)^3."""
    spec = spec_ai(7, 5, 5, n=2, jitter=0.01)
    spec["slits"] = [deck.mid_slit(spec, 30.0)]
    o = pyoracle.Oracle(spec)
    ep, xw = o.ai_points()
    rp, col, cut = o.csr()
    X, _, _ = o.nodes()
    assert np.all(np.diff(ep)[cut.astype(bool)] == 0)
    assert np.diff(ep).max() <= 27
    rng = np.random.default_rng(1)
    for e in rng.choice(len(col), 40, replace=False):
        i = int(np.searchsorted(rp, e, side="right") - 1)
        j = int(col[e])
        if cut[e]:
            continue
        a, b = (i, j) if i < j else (j, i)
        _, dmin, dmax = pyoracle.pair_dists(spec, X[a], X[b])
        ref = pyoracle.ai_cell_points(spec, X[j] - X[i], dmin, dmax)
        np.testing.assert_array_equal(xw[ep[e]:ep[e + 1]], ref)


def test_mls_coefficients_reproduce_quadratics():
    """At every sampled point, sum_k phi_k p(x_k) = p(x_point) for the quadratic basis."""
    spec = spec_ai(7, 6, 6, n=2, jitter=0.01)
    o = pyoracle.Oracle(spec)
    ep, xw = o.ai_points()
    po, phi = o.ai_mls()
    cb_rp, cb, _ = o.mls()
    rp, col, _ = o.csr()
    X, _, _ = o.nodes()
    rng = np.random.default_rng(2)
    ent = np.repeat(np.arange(len(col)), np.diff(ep))
    for q in rng.choice(len(xw), 200, replace=False):
        e = ent[q]
        i = int(np.searchsorted(rp, e, side="right") - 1)
        j = int(col[e])
        xg = xw[q, :3] + X[i]
        xk = X[cb[cb_rp[j]:cb_rp[j + 1]]]
        ph = phi[po[q]:po[q + 1]]
        assert len(ph) == len(xk)
        for f in (lambda x: np.ones(len(x)), lambda x: x[:, 0], lambda x: x[:, 1] * x[:, 2],
                  lambda x: x[:, 0] ** 2, lambda x: x[:, 2] ** 2 + x[:, 0] * x[:, 1]):
            val = f(xg[None, :])[0]
            got = ph @ f(xk)
            assert abs(got - val) <= 1e-9 * max(1e-6, abs(val)), (q, got, val)


def test_mls_affine_equals_prescribed_gradient_and_null_fields():
    """An affine u = G x through the MLS points gives the prescribed-gradient mode's forces;
    zero and rigid fields give none."""
    G = np.array([[1e-4, 2e-5, 0.0], [2e-5, -3e-5, 1e-5], [0.0, 1e-5, -3e-5]])
    sm = spec_ai(7, 6, 6, mls=1, n=2, jitter=0.01)
    sg = spec_ai(7, 6, 6, mls=0, n=2, jitter=0.01, gradient=G)
    om, og = pyoracle.Oracle(sm), pyoracle.Oracle(sg)
    X, _, _ = om.nodes()
    u = X @ G.T
    Fm = om.forces(u)
    Fg = og.forces(np.zeros_like(u))
    # the bdryEff factors do not depend on the boundary conditions
    np.testing.assert_array_equal(om.bdryeff(), og.bdryeff())
    assert np.abs(Fm - Fg).max() <= 1e-8 * np.abs(Fg).max()
    assert np.abs(om.forces(np.zeros_like(u))).max() == 0.0
    Fr = om.forces(np.tile([3e-6, -1e-6, 2e-6], (om.n, 1)))
    assert np.abs(Fr).max() <= 1e-8 * np.abs(Fg).max()
