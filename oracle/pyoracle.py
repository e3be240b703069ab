"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  It takes the same plain-dict
problem spec that pdmtel_amd.deck renders into a pds_input.txt deck, so a test
feeds one spec to both sides: the deck text to the HIP library's parser, the
values directly to the oracle.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Any, Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MAX_MAT, MAX_BC, MAX_SPACE, MAX_SLIT = 8, 16, 16, 8


class _Mat(C.Structure):
    _fields_ = [("type", C.c_int), ("E", C.c_double), ("sigy", C.c_double), ("ecrit", C.c_double),
                ("rho", C.c_double), ("nu", C.c_double), ("g", C.c_double * 3)]


class _BC(C.Structure):
    _fields_ = [("type", C.c_int), ("f", C.c_int * 3), ("v", C.c_double * 9), ("t_end", C.c_double),
                ("load", C.c_double)]


class _Space(C.Structure):
    _fields_ = [("geo", C.c_int), ("ref", C.c_int), ("b", C.c_double * 6)]


class _Slit(C.Structure):
    _fields_ = [("type", C.c_int), ("p", C.c_double * 11)]


class Problem(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("horizon", C.c_double), ("cx", C.c_double), ("cy", C.c_double), ("cz", C.c_double),
        ("total_time", C.c_double), ("total_step", C.c_int), ("dump", C.c_int),
        ("safety", C.c_double), ("fixed_dt_on", C.c_int), ("fixed_dt", C.c_double),
        ("nmat", C.c_int), ("mat", _Mat * MAX_MAT),
        ("nbc", C.c_int), ("bc", _BC * MAX_BC),
        ("nms", C.c_int), ("ms", _Space * MAX_SPACE),
        ("nds", C.c_int), ("ds", _Space * MAX_SPACE),
        ("nbs", C.c_int), ("bs", _Space * MAX_SPACE),
        ("damping", C.c_double), ("bdryeff_on", C.c_int),
        ("cci_on", C.c_int), ("cci_mdf", C.c_int), ("fgi_on", C.c_int), ("intg_point", C.c_int),
        ("mls_on", C.c_int), ("rw", C.c_double),
        ("jitter", C.c_double), ("seed", C.c_uint64),
        ("nslit", C.c_int), ("slit", _Slit * MAX_SLIT),
        ("traction_body", C.c_int),
        ("ai_on", C.c_int),
    ]


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(Problem)]
        L.orc_destroy.argtypes = [vp]
        L.orc_init.argtypes = [vp, C.c_int, C.c_int]
        for name in ("orc_num_nodes", "orc_num_entries"):
            getattr(L, name).restype = C.c_int64
            getattr(L, name).argtypes = [vp]
        L.orc_num_max_family.argtypes = [vp]
        for name in ("orc_min_bond_length", "orc_stable_dt", "orc_time", "orc_dt_current", "orc_ke", "orc_ee"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [vp]
        L.orc_spring.restype = C.c_double
        L.orc_spring.argtypes = [vp, C.c_int]
        L.orc_get_nodes.argtypes = [vp, dp, ip, ip]
        L.orc_get_csr.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)]
        L.orc_get_bdryeff.argtypes = [vp, dp]
        L.orc_set_bdryeff.argtypes = [vp, dp]
        L.orc_forces.argtypes = [vp, dp, dp, dp, C.c_int]
        L.orc_apply_tangent.argtypes = [vp, dp, dp, C.c_int]
        L.orc_run_dynamic.argtypes = [vp, C.c_int, C.c_int]
        L.orc_get_state.argtypes = [vp, dp, dp]
        L.orc_set_state.argtypes = [vp, dp, dp]
        L.orc_steps_done.argtypes = [vp]
        L.orc_pair_bonded.argtypes = [C.POINTER(Problem), dp, dp, dp, dp]
        L.orc_pair_cut.argtypes = [C.POINTER(Problem), dp, dp]
        L.orc_beta.restype = C.c_double
        L.orc_beta.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
        L.orc_bond_force_cci.argtypes = [C.POINTER(Problem), dp, dp, C.c_double, C.c_double, C.c_double,
                                         C.c_double, dp, dp]
        L.orc_set_square_mode.argtypes = [C.c_int]
        L.orc_static_system.argtypes = [vp, dp, dp, dp]
        L.orc_time_factor.restype = C.c_double
        L.orc_time_factor.argtypes = [C.POINTER(_BC), C.c_double]
        L.orc_gauss_rule.argtypes = [C.c_int, dp, dp]
        L.orc_num_gauss.argtypes = [vp]
        L.orc_num_contrib.restype = C.c_int64
        L.orc_num_contrib.argtypes = [vp]
        L.orc_get_gauss.argtypes = [vp, dp, dp]
        L.orc_get_mls.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32), dp]
        L.orc_ai_cell_points.argtypes = [C.POINTER(Problem), dp, C.c_double, C.c_double, dp, C.c_int]
        L.orc_ai_cell_force.argtypes = [C.POINTER(Problem), dp, C.c_double, C.c_double, dp, C.c_double, C.c_double,
                                        C.c_double, dp, dp]
        L.orc_num_ai_points.restype = C.c_int64
        L.orc_num_ai_points.argtypes = [vp]
        L.orc_get_ai_points.argtypes = [vp, C.POINTER(C.c_int64), dp]
        L.orc_get_ai_mls.argtypes = [vp, C.POINTER(C.c_int64), dp]
        L.orc_areal_force_density.argtypes = [vp, C.POINTER(C.c_int64), C.c_int64, dp, dp]
        L.orc_sample_forces.argtypes = [C.POINTER(Problem), C.POINTER(C.c_int64), C.c_int64, dp, dp, dp, C.c_int]
        L.orc_sample_rows.restype = C.c_int64
        L.orc_sample_rows.argtypes = [C.POINTER(Problem), C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int32), C.POINTER(C.c_uint8), dp, C.c_int64, C.c_int]
        _lib = L
    return _lib


def _space(dst: _Space, s: Dict[str, Any], with_ref: bool) -> None:
    geo = int(s.get("geo", 1))
    dst.geo = geo
    dst.ref = int(s["ref"]) - 1 if with_ref else -1
    body = s["box"] if geo == 1 else s["cyl"]
    for k, v in enumerate(body):
        dst.b[k] = float(v)


def problem_from_spec(spec: Dict[str, Any]) -> Problem:
    """Fill the oracle's problem struct from the same spec the deck writer renders."""
    p = Problem()
    p.nx, p.ny, p.nz = (int(v) for v in spec["grid_dimension"])
    p.dx = float(spec["grid_spacing"])
    p.horizon = float(spec["horizon"]) if spec.get("horizon") is not None else 3.015 * p.dx
    p.cx, p.cy, p.cz = (float(v) for v in spec.get("grid_center", (0.0, 0.0, 0.0)))
    p.total_time = float(spec["total_time"])
    p.total_step = int(spec["total_step"])
    p.dump = int(spec.get("step_dump_frequency", 1000000))
    p.safety = float(spec["safety_factor"]) if spec.get("safety_factor") is not None else 0.8
    if spec.get("fixed_time_step") is not None:
        p.fixed_dt_on = 1
        p.fixed_dt = float(spec["fixed_time_step"])
    mats = spec["materials"]
    p.nmat = len(mats)
    for i, m in enumerate(mats):
        p.mat[i].type = int(m.get("type", 1))
        p.mat[i].E = float(m["E"])
        p.mat[i].sigy = float(m.get("sigy", 1e30))
        p.mat[i].ecrit = float(m["ecrit"])
        p.mat[i].rho = float(m["rho"])
        p.mat[i].nu = 0.25
        for k, g in enumerate(m.get("gravity", (0.0, 0.0, 0.0))):
            p.mat[i].g[k] = float(g)
    bcs = spec.get("bcs", [])
    p.nbc = len(bcs)
    for i, b in enumerate(bcs):
        p.bc[i].type = int(b["type"])
        for k in range(3):
            p.bc[i].f[k] = int(b["flags"][k])
        for k, v in enumerate(b["value"]):
            p.bc[i].v[k] = float(v)
        p.bc[i].t_end = float(b["t_end"])
        p.bc[i].load = float(b["load"])
    for name, cnt, arr, with_ref in (("material_spaces", "nms", "ms", True),
                                     ("deletion_spaces", "nds", "ds", False),
                                     ("boundary_spaces", "nbs", "bs", True)):
        lst = spec.get(name, [])
        setattr(p, cnt, len(lst))
        for i, s in enumerate(lst):
            _space(getattr(p, arr)[i], s, with_ref)
    p.damping = float(spec.get("damping", 0.0))
    p.bdryeff_on = int(spec.get("bdryeff_on", 1))
    cci = spec.get("cci", (1, 1))
    if cci is not None and cci[0]:
        p.cci_on, p.cci_mdf = 1, int(cci[1])
    elif spec.get("ai") is not None:
        ai = spec["ai"]
        p.ai_on, p.mls_on = 1, int(ai[1])
        p.intg_point = int(ai[2]) if len(ai) > 2 and ai[2] is not None else 3
    else:
        fgi = spec["fgi"]
        p.fgi_on, p.intg_point = 1, int(fgi[1])
        p.mls_on = 1 if p.intg_point > 1 else 0
    rw = spec.get("support_domain_size")
    p.rw = float(rw) if rw is not None else 3.0 * p.dx
    p.jitter = float(spec.get("jitter", 0.0) or 0.0)
    p.seed = int(spec.get("seed", 0))
    slits = spec.get("slits", [])
    p.nslit = len(slits)
    for i, s in enumerate(slits):
        p.slit[i].type = int(s["type"])
        vals = list(s["box"]) if int(s["type"]) == 1 else [*s["c"], *s["n"], *s["t"], s["half_t"], s["half_s"]]
        for k, v in enumerate(vals):
            p.slit[i].p[k] = float(v)
    p.traction_body = int(spec.get("traction_body", 0))
    return p


def sample_rows(spec: Dict[str, Any], rows, nthreads: int = 0):
    """Families, cut flags and bdryEff of the given rows only (for problems too large to build
    whole): (rp, col, cut, be), rows in the given order."""
    p = problem_from_spec(spec)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    nr = len(rows)
    cap = 512 * max(nr, 1)
    rp = np.zeros(nr + 1, dtype=np.int64)
    col = np.zeros(cap, dtype=np.int32)
    cut = np.zeros(cap, dtype=np.uint8)
    be = np.zeros(max(nr, 1))
    tot = lib().orc_sample_rows(C.byref(p), rows.ctypes.data_as(C.POINTER(C.c_int64)), nr,
                                rp.ctypes.data_as(C.POINTER(C.c_int64)), col.ctypes.data_as(C.POINTER(C.c_int32)),
                                cut.ctypes.data_as(C.POINTER(C.c_uint8)), _dp(be), cap, nthreads)
    if tot < 0:
        raise RuntimeError(f"orc_sample_rows failed ({tot})")
    return rp, col[:tot].copy(), cut[:tot].copy(), be[:nr].copy()


def sample_forces(spec: Dict[str, Any], rows, u, nthreads: int = 0):
    """Family force sums and energy densities of the given rows only, for the displacement field
    u of every node: (F (nrows, 3), el (nrows))."""
    p = problem_from_spec(spec)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    F = np.zeros(3 * max(len(rows), 1))
    el = np.zeros(max(len(rows), 1))
    rc = lib().orc_sample_forces(C.byref(p), rows.ctypes.data_as(C.POINTER(C.c_int64)), len(rows), _dp(u), _dp(F),
                                 _dp(el), nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_sample_forces failed ({rc})")
    return F[: 3 * len(rows)].reshape(-1, 3), el[: len(rows)]


def set_square_mode(exact: bool) -> None:
    """Evaluate the reference's pow(x, 2.0) as glibc pow (False) or as x*x (True)."""
    lib().orc_set_square_mode(1 if exact else 0)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def gauss_rule(n: int):
    """The oracle's order-n Gauss-Legendre rule (abscissae, weights), reference point order."""
    a, w = np.zeros(8), np.zeros(8)
    if lib().orc_gauss_rule(int(n), _dp(a), _dp(w)) != 0:
        raise ValueError(f"no Gauss rule of order {n}")
    return a[:n].copy(), w[:n].copy()


def ai_cell_points(spec: Dict[str, Any], c, dmin: float, dmax: float) -> np.ndarray:
    """The adaptive-integration points (x, y, z, weight) of one cell with centre c relative to
    node i, for the bond distances dmin / dmax; shape (P, 4)."""
    p = problem_from_spec(spec)
    cc = np.ascontiguousarray(c, dtype=np.float64)
    cnt = lib().orc_ai_cell_points(C.byref(p), _dp(cc), float(dmin), float(dmax), None, 0)
    xw = np.zeros(4 * max(cnt, 1))
    lib().orc_ai_cell_points(C.byref(p), _dp(cc), float(dmin), float(dmax), _dp(xw), cnt)
    return xw[: 4 * cnt].reshape(-1, 4)


def ai_cell_force(spec: Dict[str, Any], c, dmin: float, dmax: float, G, bij: float = 1.0, spring: float = 1.0,
                  ecrit: float = 1e10):
    """Integrated bond force (3) and energy density of one AI cell for eta = G xi."""
    p = problem_from_spec(spec)
    cc = np.ascontiguousarray(c, dtype=np.float64)
    g = np.ascontiguousarray(G, dtype=np.float64).reshape(-1)
    F = np.zeros(3)
    el = C.c_double()
    lib().orc_ai_cell_force(C.byref(p), _dp(cc), float(dmin), float(dmax), _dp(g), float(bij), float(spring),
                            float(ecrit), _dp(F), C.byref(el))
    return F, el.value


def pair_dists(spec: Dict[str, Any], xi, xj):
    """(bonded, dmin, dmax) of the pair as InitBond computes it (xi the lower id)."""
    p = problem_from_spec(spec)
    a = np.ascontiguousarray(xi, dtype=np.float64)
    b = np.ascontiguousarray(xj, dtype=np.float64)
    dmin, dmax = C.c_double(), C.c_double()
    ok = lib().orc_pair_bonded(C.byref(p), _dp(a), _dp(b), C.byref(dmin), C.byref(dmax))
    return bool(ok), dmin.value, dmax.value


class Oracle:
    """One oracle problem instance (init = InitNode/InitBond/InitMaterial/InitBdryEffectFactor)."""

    def __init__(self, spec: Dict[str, Any], bond_method: int = 1, nthreads: int = 0):
        self.L = lib()
        self.spec = spec
        self.prob = problem_from_spec(spec)
        self.nthreads = nthreads
        self.h = self.L.orc_create(C.byref(self.prob))
        rc = self.L.orc_init(self.h, bond_method, nthreads)
        if rc != 0:
            self.L.orc_destroy(self.h)
            self.h = None
            raise RuntimeError(f"oracle init failed rc={rc}")
        self.n = int(self.L.orc_num_nodes(self.h))
        self.E = int(self.L.orc_num_entries(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def nodes(self):
        X = np.zeros(3 * self.n)
        mat = np.zeros(self.n, dtype=np.int32)
        bc = np.zeros(self.n, dtype=np.int32)
        self.L.orc_get_nodes(self.h, _dp(X), mat.ctypes.data_as(C.POINTER(C.c_int)),
                             bc.ctypes.data_as(C.POINTER(C.c_int)))
        return X.reshape(-1, 3), mat, bc

    def csr(self):
        rp = np.zeros(self.n + 1, dtype=np.int64)
        col = np.zeros(max(self.E, 1), dtype=np.int32)
        cut = np.zeros(max(self.E, 1), dtype=np.uint8)
        self.L.orc_get_csr(self.h, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                           col.ctypes.data_as(C.POINTER(C.c_int32)), cut.ctypes.data_as(C.POINTER(C.c_uint8)))
        return rp, col[: self.E], cut[: self.E]

    def bdryeff(self):
        be = np.zeros(self.n)
        self.L.orc_get_bdryeff(self.h, _dp(be))
        return be

    def forces(self, u: np.ndarray, with_elast: bool = False):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        F = np.zeros(3 * self.n)
        el = np.zeros(self.n) if with_elast else None
        self.L.orc_forces(self.h, _dp(u), _dp(F), _dp(el) if el is not None else None, self.nthreads)
        return (F.reshape(-1, 3), el) if with_elast else F.reshape(-1, 3)

    def tangent(self, u: np.ndarray):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        y = np.zeros(3 * self.n)
        self.L.orc_apply_tangent(self.h, _dp(u), _dp(y), self.nthreads)
        return y.reshape(-1, 3)

    def run(self, nsteps: int) -> int:
        return int(self.L.orc_run_dynamic(self.h, nsteps, self.nthreads))

    def state(self):
        U = np.zeros(3 * self.n)
        V = np.zeros(3 * self.n)
        self.L.orc_get_state(self.h, _dp(U), _dp(V))
        return U.reshape(-1, 3), V.reshape(-1, 3)

    def set_state(self, U, V) -> None:
        """U, V: 3N arrays (reference node order), or None to keep that field."""
        U = None if U is None else np.ascontiguousarray(U, dtype=np.float64).reshape(-1)
        V = None if V is None else np.ascontiguousarray(V, dtype=np.float64).reshape(-1)
        self.L.orc_set_state(self.h, _dp(U) if U is not None else None, _dp(V) if V is not None else None)

    def areal_force_density(self, nodes, be=None):
        """Areal force density (9 components per node, row n = plane normal) from the current U."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        sig = np.zeros(9 * max(len(nodes), 1))
        bep = None if be is None else np.ascontiguousarray(be, dtype=np.float64)
        self.L.orc_areal_force_density(self.h, nodes.ctypes.data_as(C.POINTER(C.c_int64)), len(nodes),
                                       _dp(bep) if bep is not None else None, _dp(sig))
        return sig[: 9 * len(nodes)].reshape(-1, 3, 3)

    def gauss(self):
        """Gauss-point positions and displacements, each (N, n^3, 3); None for CCI."""
        ng = int(self.L.orc_num_gauss(self.h))
        if ng == 0:
            return None
        Xg, Ug = np.zeros(3 * ng * self.n), np.zeros(3 * ng * self.n)
        self.L.orc_get_gauss(self.h, _dp(Xg), _dp(Ug))
        return Xg.reshape(self.n, ng, 3), Ug.reshape(self.n, ng, 3)

    def mls(self):
        """MLS contributing nodes and coefficients: (cb_rp, cb, phi), node i's coefficients
        phi[ng*cb_rp[i] : ng*cb_rp[i+1]] as [point][contributing node]; None for n = 1."""
        m = int(self.L.orc_num_contrib(self.h))
        if m == 0:
            return None
        ng = int(self.L.orc_num_gauss(self.h))
        rp = np.zeros(self.n + 1, dtype=np.int64)
        cb = np.zeros(m, dtype=np.int32)
        phi = np.zeros(ng * m)
        self.L.orc_get_mls(self.h, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                           cb.ctypes.data_as(C.POINTER(C.c_int32)), _dp(phi))
        return rp, cb, phi

    def ai_points(self):
        """Adaptive-integration points: (ent_ptr (E + 1), xw (P, 4)) in reference entry order;
        None without ai_on."""
        P = int(self.L.orc_num_ai_points(self.h))
        if P == 0:
            return None
        ep = np.zeros(self.E + 1, dtype=np.int64)
        xw = np.zeros(4 * P)
        self.L.orc_get_ai_points(self.h, ep.ctypes.data_as(C.POINTER(C.c_int64)), _dp(xw))
        return ep, xw.reshape(-1, 4)

    def ai_mls(self):
        """MLS coefficients of the AI points: (po (P + 1), phi); point q's over node j's
        contributing nodes (mls()'s cb_rp / cb) at phi[po[q]:po[q + 1]]."""
        P = int(self.L.orc_num_ai_points(self.h))
        po = np.zeros(P + 1, dtype=np.int64)
        self.L.orc_get_ai_mls(self.h, po.ctypes.data_as(C.POINTER(C.c_int64)), None)
        phi = np.zeros(max(int(po[-1]), 1))
        self.L.orc_get_ai_mls(self.h, po.ctypes.data_as(C.POINTER(C.c_int64)), _dp(phi))
        return po, phi[: int(po[-1])]

    def static_system(self):
        m, uc, b = np.zeros(3 * self.n), np.zeros(3 * self.n), np.zeros(3 * self.n)
        self.L.orc_static_system(self.h, _dp(m), _dp(uc), _dp(b))
        return m.reshape(-1, 3), uc.reshape(-1, 3), b.reshape(-1, 3)

    def dense_tangent(self) -> np.ndarray:
        """K assembled column by column from the matrix-free tangent (small problems only)."""
        n3 = 3 * self.n
        K = np.zeros((n3, n3))
        e = np.zeros(n3)
        for k in range(n3):
            e[k] = 1.0
            K[:, k] = self.tangent(e).reshape(-1)
            e[k] = 0.0
        return K

    def solve_static_dense(self):
        """Static solution by a dense direct solve of the masked system (small problems only)."""
        m, uc, b = self.static_system()
        K = self.dense_tangent()
        free = m.reshape(-1) > 0
        ucf = uc.reshape(-1)
        rhs = b.reshape(-1) - K @ ucf
        u = ucf.copy()
        u[free] = np.linalg.solve(K[np.ix_(free, free)], rhs[free])
        return u.reshape(-1, 3)

    @property
    def max_family(self) -> int:
        return int(self.L.orc_num_max_family(self.h))

    @property
    def min_bond(self) -> float:
        return float(self.L.orc_min_bond_length(self.h))

    @property
    def dt(self) -> float:
        return float(self.L.orc_stable_dt(self.h))

    @property
    def t(self) -> float:
        return float(self.L.orc_time(self.h))

    @property
    def ke(self) -> float:
        return float(self.L.orc_ke(self.h))

    @property
    def ee(self) -> float:
        return float(self.L.orc_ee(self.h))

    @property
    def steps_done(self) -> int:
        return int(self.L.orc_steps_done(self.h))

    def spring(self, m: int = 0) -> float:
        return float(self.L.orc_spring(self.h, m))
