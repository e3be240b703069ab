"""pdmtel_amd -- MI355X-native linear bond-based peridynamics (Python host over the C ABI).

The product is the HIP library ``libpdm.so`` (C ABI in include/pdm.h).  This
module is a thin ctypes binding with the reference's driver vocabulary
(pdSolve::InitialProblem / SolveDynamic / SolveStatic, pdDataManager getters).
There is no CPU fallback: importing the binding without the built library
raises, and every call goes to the GPU library.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Tuple

import numpy as np

from .deck import bar_spec, deck_text, mid_slit  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpdm.so")

PDM_FP64 = 0
PDM_FP32C = 1
PDM_FP32C2 = 2

ERRORS = {
    -1: "PDM_E_ARG", -2: "PDM_E_IO", -3: "PDM_E_PARSE", -4: "PDM_E_MODEL", -5: "PDM_E_HIP",
    -6: "PDM_E_STATE", -7: "PDM_E_LIMIT", -8: "PDM_E_CONVERGE",
}
globals().update({name: code for code, name in ERRORS.items()})


class PdmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [("precision", C.c_int), ("device", C.c_int), ("use_graph", C.c_int), ("graph_chunk", C.c_int),
                ("energies", C.c_int), ("tile_edge", C.c_int), ("waves_per_tile", C.c_int),
                ("families_per_wave", C.c_int), ("verbose", C.c_int)]


class StaticStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("refinements", C.c_int), ("rel_residual", C.c_double),
                ("seconds", C.c_double)]


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load libpdm.so (built in-tree by __graft_entry__.build() / `make -C pdmtel_amd`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `make -C pdmtel_amd` or __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, dp, i32p, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    L.pdm_options_default.argtypes = [C.POINTER(Options)]
    L.pdm_create.argtypes = [C.c_char_p, C.POINTER(Options), C.POINTER(vp)]
    L.pdm_create_from_text.argtypes = [C.c_char_p, C.POINTER(Options), C.POINTER(vp)]
    L.pdm_destroy.argtypes = [vp]
    L.pdm_last_error.restype = C.c_char_p
    L.pdm_last_error.argtypes = [vp]
    L.pdm_init.argtypes = [vp]
    for name in ("pdm_num_nodes", "pdm_num_bonds", "pdm_num_entries", "pdm_num_live_entries"):
        getattr(L, name).restype = C.c_int64
        getattr(L, name).argtypes = [vp]
    L.pdm_max_family.argtypes = [vp]
    L.pdm_get_nodes.argtypes = [vp, dp, i32p, i32p]
    L.pdm_get_csr.argtypes = [vp, i64p, i32p, C.POINTER(C.c_uint8)]
    L.pdm_get_bdryeff.argtypes = [vp, dp]
    L.pdm_get_timestep.argtypes = [vp, dp, dp]
    L.pdm_compute_forces.argtypes = [vp, dp, dp, dp]
    L.pdm_apply_tangent.argtypes = [vp, dp, dp]
    L.pdm_compute_forces_noalloc.argtypes = [vp, vp, vp, vp, C.c_int]
    L.pdm_areal_force_density.argtypes = [vp, i64p, C.c_int64, dp, dp]
    L.pdm_set_state.argtypes = [vp, dp, dp]
    L.pdm_get_state.argtypes = [vp, dp, dp]
    L.pdm_step_rk4.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.pdm_prepare_steps.argtypes = [vp, C.c_int]
    L.pdm_get_clock.argtypes = [vp, dp, C.POINTER(C.c_int), dp]
    L.pdm_get_energy.argtypes = [vp, dp, dp]
    L.pdm_solve_dynamic.argtypes = [vp, C.c_char_p]
    L.pdm_solve_static.argtypes = [vp, C.c_double, C.c_int, C.POINTER(StaticStats)]
    L.pdm_write_tecplot.argtypes = [vp, C.c_char_p, C.c_int]
    L.pdm_last_kernel_timing.argtypes = [vp, dp, i64p]
    ip = C.POINTER(C.c_int)
    L.pdm_get_tuning.argtypes = [vp, ip, ip, ip, dp]
    L.pdm_get_stage_launch.argtypes = [vp, ip, i64p]
    L.pdm_num_gauss_points.argtypes = [vp]
    L.pdm_num_live_gauss_pairs.restype = C.c_int64
    L.pdm_num_live_gauss_pairs.argtypes = [vp]
    L.pdm_get_gauss.argtypes = [vp, dp, dp]
    L.pdm_num_ai_points.restype = C.c_int64
    L.pdm_num_ai_points.argtypes = [vp]
    L.pdm_get_ai_points.argtypes = [vp, i32p, i32p, dp]
    L.pdm_num_ai_coefficients.restype = C.c_int64
    L.pdm_num_ai_coefficients.argtypes = [vp]
    cpp = C.POINTER(C.c_char_p)
    L.pdm_batch_create.argtypes = [cpp, C.c_int, C.POINTER(Options), C.POINTER(vp)]
    L.pdm_batch_create_from_text.argtypes = [cpp, C.c_int, C.POINTER(Options), C.POINTER(vp)]
    L.pdm_batch_destroy.argtypes = [vp]
    L.pdm_batch_last_error.restype = C.c_char_p
    L.pdm_batch_last_error.argtypes = [vp]
    L.pdm_batch_init.argtypes = [vp]
    L.pdm_batch_num_specimens.argtypes = [vp]
    for name in ("pdm_batch_num_nodes", "pdm_batch_num_entries", "pdm_batch_num_live_entries"):
        getattr(L, name).restype = C.c_int64
        getattr(L, name).argtypes = [vp, C.c_int]
    L.pdm_batch_step_rk4.argtypes = [vp, C.c_int, ip]
    L.pdm_batch_prepare_steps.argtypes = [vp, C.c_int]
    L.pdm_batch_get_state.argtypes = [vp, C.c_int, dp, dp]
    L.pdm_batch_set_state.argtypes = [vp, C.c_int, dp, dp]
    L.pdm_batch_get_clock.argtypes = [vp, C.c_int, dp, ip, dp]
    L.pdm_batch_get_energy.argtypes = [vp, C.c_int, dp, dp]
    L.pdm_batch_get_tuning.argtypes = [vp, ip, ip, ip, dp]
    L.pdm_batch_get_spec_tuning.argtypes = [vp, C.c_int, ip, ip, ip]
    L.pdm_batch_last_kernel_timing.argtypes = [vp, dp, i64p]
    _lib = L
    return L


def options(precision: int = PDM_FP64, **kw) -> Options:
    o = Options()
    lib().pdm_options_default(C.byref(o))
    o.precision = precision
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Problem:
    """One specimen: pdm_create + pdm_init (the InitialProblem sequence, pdSolve.cpp:10-22)."""

    def __init__(self, deck: str, precision: int = PDM_FP64, from_text: bool = True, init: bool = True, **opts):
        self.L = lib()
        self.h = C.c_void_p()
        o = options(precision, **opts)
        if from_text:
            rc = self.L.pdm_create_from_text(deck.encode(), C.byref(o), C.byref(self.h))
        else:
            rc = self.L.pdm_create(deck.encode(), C.byref(o), C.byref(self.h))
        if rc != 0:
            raise PdmError(rc, self.L.pdm_last_error(None).decode())
        self.precision = precision
        if init:
            self.init()

    def _chk(self, rc: int) -> None:
        if rc != 0:
            raise PdmError(rc, self.L.pdm_last_error(self.h).decode())

    def init(self) -> None:
        self._chk(self.L.pdm_init(self.h))
        self.n = int(self.L.pdm_num_nodes(self.h))
        self.E = int(self.L.pdm_num_entries(self.h))
        self.E_live = int(self.L.pdm_num_live_entries(self.h))

    def close(self) -> None:
        if getattr(self, "h", None) and self.h.value:
            self.L.pdm_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def max_family(self) -> int:
        return int(self.L.pdm_max_family(self.h))

    def nodes(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        X = np.zeros(3 * self.n)
        mat = np.zeros(self.n, dtype=np.int32)
        bc = np.zeros(self.n, dtype=np.int32)
        self._chk(self.L.pdm_get_nodes(self.h, _dp(X), mat.ctypes.data_as(C.POINTER(C.c_int32)),
                                       bc.ctypes.data_as(C.POINTER(C.c_int32))))
        return X.reshape(-1, 3), mat, bc

    def csr(self):
        rp = np.zeros(self.n + 1, dtype=np.int64)
        col = np.zeros(max(self.E, 1), dtype=np.int32)
        cut = np.zeros(max(self.E, 1), dtype=np.uint8)
        self._chk(self.L.pdm_get_csr(self.h, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                     col.ctypes.data_as(C.POINTER(C.c_int32)),
                                     cut.ctypes.data_as(C.POINTER(C.c_uint8))))
        return rp, col[: self.E], cut[: self.E]

    def bdryeff(self) -> np.ndarray:
        be = np.zeros(self.n)
        self._chk(self.L.pdm_get_bdryeff(self.h, _dp(be)))
        return be

    def timestep(self) -> Tuple[float, float]:
        dt, mb = C.c_double(), C.c_double()
        self._chk(self.L.pdm_get_timestep(self.h, C.byref(dt), C.byref(mb)))
        return dt.value, mb.value

    def forces(self, u: np.ndarray, with_elast: bool = False):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        F = np.zeros(3 * self.n)
        el = np.zeros(self.n)
        self._chk(self.L.pdm_compute_forces(self.h, _dp(u), _dp(F), _dp(el)))
        return (F.reshape(-1, 3), el) if with_elast else F.reshape(-1, 3)

    def areal_force_density(self, nodes, bdryeff: Optional[np.ndarray] = None) -> np.ndarray:
        """PostAFD's stress at the given nodes from the current U: (count, 3, 3), row = plane normal."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        sig = np.zeros(9 * max(len(nodes), 1))
        be = None if bdryeff is None else np.ascontiguousarray(bdryeff, dtype=np.float64)
        self._chk(self.L.pdm_areal_force_density(self.h, nodes.ctypes.data_as(C.POINTER(C.c_int64)), len(nodes),
                                                  _dp(be) if be is not None else None, _dp(sig)))
        return sig[: 9 * len(nodes)].reshape(-1, 3, 3)

    def forces_noalloc(self, u_dev: int, F_dev: int, elast_dev: int = 0, sync: bool = True) -> None:
        """pdm_compute_forces_noalloc on device buffers given by address (reference-order 3N u
        and F, N energy densities or 0): no device allocation in the call."""
        self._chk(self.L.pdm_compute_forces_noalloc(self.h, C.c_void_p(u_dev), C.c_void_p(F_dev),
                                                     C.c_void_p(elast_dev or None), 1 if sync else 0))

    @property
    def live_gauss_pairs(self) -> int:
        """(entry, Gauss point) pairs the fixed-Gaussian gathers evaluate (0 for CCI)."""
        return int(self.L.pdm_num_live_gauss_pairs(self.h))

    def gauss(self):
        """Gauss-point positions and displacements, each (N, n^3, 3); None for a CCI deck."""
        ng = int(self.L.pdm_num_gauss_points(self.h))
        if ng <= 0:
            return None
        Xg = np.zeros(3 * ng * self.n)
        Ug = np.zeros(3 * ng * self.n)
        self._chk(self.L.pdm_get_gauss(self.h, _dp(Xg), _dp(Ug)))
        return Xg.reshape(self.n, ng, 3), Ug.reshape(self.n, ng, 3)

    @property
    def ai_counts(self) -> Tuple[int, int]:
        """(points, MLS coefficients) of adaptive integration (0, 0 for other modes)."""
        return int(self.L.pdm_num_ai_points(self.h)), int(self.L.pdm_num_ai_coefficients(self.h))

    def ai_points(self):
        """Adaptive-integration points: (node i, node j (reference ids), xw (P, 4): x, y, z relative
        to node i and the weight); the points of an entry are contiguous.  None for other modes."""
        P = int(self.L.pdm_num_ai_points(self.h))
        if P <= 0:
            return None
        ni = np.zeros(P, dtype=np.int32)
        nj = np.zeros(P, dtype=np.int32)
        xw = np.zeros(4 * P)
        self._chk(self.L.pdm_get_ai_points(self.h, ni.ctypes.data_as(C.POINTER(C.c_int32)),
                                           nj.ctypes.data_as(C.POINTER(C.c_int32)), _dp(xw)))
        return ni, nj, xw.reshape(-1, 4)

    def tangent(self, u: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        y = np.zeros(3 * self.n)
        self._chk(self.L.pdm_apply_tangent(self.h, _dp(u), _dp(y)))
        return y.reshape(-1, 3)

    def set_state(self, U: Optional[np.ndarray], V: Optional[np.ndarray]) -> None:
        Ua = np.ascontiguousarray(U, dtype=np.float64).reshape(-1) if U is not None else None
        Va = np.ascontiguousarray(V, dtype=np.float64).reshape(-1) if V is not None else None
        self._chk(self.L.pdm_set_state(self.h, _dp(Ua) if Ua is not None else None,
                                       _dp(Va) if Va is not None else None))

    def state(self) -> Tuple[np.ndarray, np.ndarray]:
        U = np.zeros(3 * self.n)
        V = np.zeros(3 * self.n)
        self._chk(self.L.pdm_get_state(self.h, _dp(U), _dp(V)))
        return U.reshape(-1, 3), V.reshape(-1, 3)

    def step(self, nsteps: int) -> int:
        done = C.c_int()
        self._chk(self.L.pdm_step_rk4(self.h, int(nsteps), C.byref(done)))
        return done.value

    def prepare(self, nsteps: int) -> None:
        """Capture the graph of an nsteps-step chunk now, so stepping that many replays it."""
        self._chk(self.L.pdm_prepare_steps(self.h, int(nsteps)))

    def clock(self) -> Tuple[float, int, float]:
        t, s, dt = C.c_double(), C.c_int(), C.c_double()
        self._chk(self.L.pdm_get_clock(self.h, C.byref(t), C.byref(s), C.byref(dt)))
        return t.value, s.value, dt.value

    def energy(self) -> Tuple[float, float]:
        ke, ee = C.c_double(), C.c_double()
        self._chk(self.L.pdm_get_energy(self.h, C.byref(ke), C.byref(ee)))
        return ke.value, ee.value

    def kernel_timing(self) -> Tuple[float, int]:
        ms, n = C.c_double(), C.c_int64()
        self._chk(self.L.pdm_last_kernel_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def tuning(self) -> dict:
        """Tile layout in use (autotuned at init for large specimens when no layout option is set)."""
        t, w, f, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        self._chk(self.L.pdm_get_tuning(self.h, C.byref(t), C.byref(w), C.byref(f), C.byref(ms)))
        return dict(tile_edge=t.value, waves_per_tile=w.value, families_per_wave=f.value, tuned_stage_ms=ms.value)

    def stage_launch(self) -> dict:
        """Register budget (128, or 80) and LDS bytes per workgroup of the RK4 stage kernel."""
        v, b = C.c_int(), C.c_int64()
        self._chk(self.L.pdm_get_stage_launch(self.h, C.byref(v), C.byref(b)))
        return dict(max_vgprs=v.value, lds_bytes=b.value)

    def solve_dynamic(self, out_dir: str) -> None:
        self._chk(self.L.pdm_solve_dynamic(self.h, out_dir.encode()))

    def solve_static(self, rel_tol: float = 1e-10, max_iterations: int = 100000) -> StaticStats:
        st = StaticStats()
        self._chk(self.L.pdm_solve_static(self.h, rel_tol, max_iterations, C.byref(st)))
        return st

    def write_tecplot(self, path: str, step: int = 0) -> None:
        self._chk(self.L.pdm_write_tecplot(self.h, path.encode(), int(step)))


class Batch:
    """Many specimens stepped together (pdm_batch_*): one launch per RK4 stage over all tiles."""

    def __init__(self, decks, precision: int = PDM_FP32C, init: bool = True, **opts):
        self.L = lib()
        self.h = C.c_void_p()
        o = options(precision, **opts)
        arr = (C.c_char_p * len(decks))(*[d.encode() for d in decks])
        rc = self.L.pdm_batch_create_from_text(arr, len(decks), C.byref(o), C.byref(self.h))
        if rc != 0:
            raise PdmError(rc, self.L.pdm_batch_last_error(None).decode())
        self.nspec = len(decks)
        if init:
            self.init()

    def _chk(self, rc: int) -> None:
        if rc != 0:
            raise PdmError(rc, self.L.pdm_batch_last_error(self.h).decode())

    def init(self) -> None:
        self._chk(self.L.pdm_batch_init(self.h))
        self.n = [int(self.L.pdm_batch_num_nodes(self.h, k)) for k in range(self.nspec)]
        self.E = [int(self.L.pdm_batch_num_entries(self.h, k)) for k in range(self.nspec)]
        self.E_live = [int(self.L.pdm_batch_num_live_entries(self.h, k)) for k in range(self.nspec)]

    def close(self) -> None:
        if getattr(self, "h", None) and self.h.value:
            self.L.pdm_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps: int) -> int:
        done = C.c_int()
        self._chk(self.L.pdm_batch_step_rk4(self.h, int(nsteps), C.byref(done)))
        return done.value

    def prepare(self, nsteps: int) -> None:
        """Capture the graph of an nsteps-step chunk now, so stepping that many replays it."""
        self._chk(self.L.pdm_batch_prepare_steps(self.h, int(nsteps)))

    def state(self, spec: int) -> Tuple[np.ndarray, np.ndarray]:
        U = np.zeros(3 * self.n[spec])
        V = np.zeros(3 * self.n[spec])
        self._chk(self.L.pdm_batch_get_state(self.h, spec, _dp(U), _dp(V)))
        return U.reshape(-1, 3), V.reshape(-1, 3)

    def set_state(self, spec: int, U: Optional[np.ndarray], V: Optional[np.ndarray]) -> None:
        Ua = np.ascontiguousarray(U, dtype=np.float64).reshape(-1) if U is not None else None
        Va = np.ascontiguousarray(V, dtype=np.float64).reshape(-1) if V is not None else None
        self._chk(self.L.pdm_batch_set_state(self.h, spec, _dp(Ua) if Ua is not None else None,
                                             _dp(Va) if Va is not None else None))

    def clock(self, spec: int) -> Tuple[float, int, float]:
        t, s, dt = C.c_double(), C.c_int(), C.c_double()
        self._chk(self.L.pdm_batch_get_clock(self.h, spec, C.byref(t), C.byref(s), C.byref(dt)))
        return t.value, s.value, dt.value

    def energy(self, spec: int) -> Tuple[float, float]:
        ke, ee = C.c_double(), C.c_double()
        self._chk(self.L.pdm_batch_get_energy(self.h, spec, C.byref(ke), C.byref(ee)))
        return ke.value, ee.value

    def tuning(self, spec: int = -1) -> dict:
        """The batch's layout (that of its largest tile class), or with spec >= 0 the layout
        that specimen's tiles run with."""
        t, w, f, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        if spec >= 0:
            self._chk(self.L.pdm_batch_get_spec_tuning(self.h, spec, C.byref(t), C.byref(w), C.byref(f)))
            return dict(tile_edge=t.value, waves_per_tile=w.value, families_per_wave=f.value)
        self._chk(self.L.pdm_batch_get_tuning(self.h, C.byref(t), C.byref(w), C.byref(f), C.byref(ms)))
        return dict(tile_edge=t.value, waves_per_tile=w.value, families_per_wave=f.value, tuned_stage_ms=ms.value)

    def kernel_timing(self) -> Tuple[float, int]:
        ms, n = C.c_double(), C.c_int64()
        self._chk(self.L.pdm_batch_last_kernel_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value
