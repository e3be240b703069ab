"""CPU tests of the host side: the C ABI library loads and exports every entry
point of include/pdm.h, the deck parser (through pdm_create_* and
pdm_get_problem_info, which need no GPU) matches the reference's keyword rules
on its own sample deck, and the error paths return the documented codes."""

import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import pdmtel_amd as pm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


class Info(C.Structure):
    _fields_ = [
        ("grid_dimension", C.c_int * 3), ("grid_spacing", C.c_double), ("horizon", C.c_double),
        ("grid_center", C.c_double * 3), ("total_time", C.c_double), ("total_step", C.c_int),
        ("step_dump_frequency", C.c_int), ("safety_factor", C.c_double), ("fixed_time_step", C.c_double),
        ("num_materials", C.c_int), ("num_bcs", C.c_int), ("num_material_spaces", C.c_int),
        ("num_deletion_spaces", C.c_int), ("num_boundary_spaces", C.c_int), ("num_slits", C.c_int),
        ("damping_coeff", C.c_double), ("bdryeff_on", C.c_int), ("cci_on", C.c_int), ("cci_modified", C.c_int),
        ("fgi_on", C.c_int), ("ai_on", C.c_int), ("mls_on", C.c_int), ("intg_point", C.c_int),
        ("support_domain_size", C.c_double), ("tolerance", C.c_double), ("jitter", C.c_double),
        ("jitter_seed", C.c_longlong), ("traction_body", C.c_int),
    ]


def create(text):
    L = pm.lib()
    h = C.c_void_p()
    o = pm.options()
    rc = L.pdm_create_from_text(text.encode(), C.byref(o), C.byref(h))
    return rc, h


def info(text):
    rc, h = create(text)
    assert rc == 0, pm.lib().pdm_last_error(None)
    L = pm.lib()
    L.pdm_get_problem_info.argtypes = [C.c_void_p, C.POINTER(Info)]
    inf = Info()
    assert L.pdm_get_problem_info(h, C.byref(inf)) == 0
    L.pdm_destroy(h)
    return inf


def test_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "pdm.h")).read()
    names = sorted(set(re.findall(r"\b(pdm_[a-z_0-9]+)\s*\(", hdr)))
    assert len(names) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", pm.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (pdm_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_reference_sample_deck():
    """The reference's own sample deck (CRLF): values read by the reference's keyword rules."""
    raw = open(os.path.join(GOLDEN, "pds_input_sample.txt"), "rb").read().decode()
    assert "\r\n" in raw
    inf = info(raw)
    assert list(inf.grid_dimension) == [20, 40, 20]
    assert inf.grid_spacing == 2.0 and inf.horizon == 6.0
    assert list(inf.grid_center) == [0.0, 0.0, 0.0]
    assert inf.total_time == 2.0 and inf.total_step == 50 and inf.step_dump_frequency == 50
    assert inf.safety_factor == 1.1 and inf.fixed_time_step < 0
    assert inf.num_materials == 1 and inf.num_bcs == 3 and inf.num_material_spaces == 1
    # the deck says 2 boundary spaces and lists 6; the reader stops at the count (pdDataManager.cpp:433)
    assert inf.num_boundary_spaces == 2 and inf.num_deletion_spaces == 0
    assert inf.damping_coeff == 0.06 and inf.bdryeff_on == 1
    assert inf.cci_on == 1 and inf.cci_modified == 0 and inf.fgi_on == 0 and inf.ai_on == 0
    assert inf.tolerance == 1e-4 * 2.0
    # LF and CRLF parse identically
    lf = info(raw.replace("\r\n", "\n"))
    assert bytes(inf) == bytes(lf)


def test_defaults_and_extensions():
    spec = pm.bar_spec(4, 3, 3, jitter=0.01, seed=7, slits=[dict(type=1, box=(0, 1, 0.0015, 0.0015, 0, 1))])
    spec["horizon"] = None
    spec["safety_factor"] = None
    inf = info(pm.deck_text(spec))
    assert math_close(inf.horizon, 3.015 * 1e-3) and inf.safety_factor == 0.8
    assert inf.jitter == 0.01 and inf.jitter_seed == 7 and inf.num_slits == 1 and inf.traction_body == 1


def test_fgi_deck_values():
    spec = pm.bar_spec(4, 3, 3)
    spec["cci"] = None
    spec["fgi"] = (1, 2)
    inf = info(pm.deck_text(spec))
    assert inf.cci_on == 0 and inf.fgi_on == 1 and inf.intg_point == 2 and inf.mls_on == 1
    assert math_close(inf.support_domain_size, 3.0 * 1e-3)
    spec["fgi"] = (1, 1)
    inf = info(pm.deck_text(spec))
    assert inf.intg_point == 1 and inf.mls_on == 0


def math_close(a, b):
    return abs(a - b) <= 1e-15 * abs(b)


def test_error_codes():
    L = pm.lib()
    spec = pm.bar_spec(4, 3, 3)
    text = pm.deck_text(spec)
    rc, _ = create(text.replace("grid_spacing", "grid_spacingX"))
    assert rc == -3 and b"grid_spacing" in L.pdm_last_error(None)
    rc, _ = create(text.replace("cci_on\n 1 1", "cci_on\n 0 0"))
    assert rc == -3  # fgi_on missing: REQ keyword not found
    rc, _ = create(text.replace("cci_on\n 1 1", "cci_on\n 0 0\nfgi_on\n 0\nai_on\n 0 0"))
    assert rc == -3 and b"no integration method" in L.pdm_last_error(None)
    rc, _ = create(text.replace("material_1\n 1", "material_1\n 7"))
    assert rc == -3
    fgi = text.replace("cci_on\n 1 1", "cci_on\n 0 0\nfgi_on\n 1\nintg_point\n 7")
    rc, _ = create(fgi)
    assert rc == -3 and b"intg_point" in L.pdm_last_error(None)
    ai = text.replace("cci_on\n 1 1", "cci_on\n 0 0\nfgi_on\n 0\nai_on\n 1 1")
    rc, h = create(ai)
    assert rc == 0  # adaptive integration, default trapezoidal index 3
    L.pdm_destroy(h)
    rc, _ = create(ai + "error_control_on\n 1\n")
    assert rc == -4 and b"error_control_on" in L.pdm_last_error(None)
    rc, _ = create(ai + "intg_point\n 6\n")
    assert rc == -3 and b"intg_point" in L.pdm_last_error(None)
    h = C.c_void_p()
    rc = L.pdm_create_from_text(ai.encode(), C.byref(pm.options(pm.PDM_FP32C)), C.byref(h))
    assert rc == -4 and b"FP64" in L.pdm_last_error(None)
    o32 = pm.options(pm.PDM_FP32C)
    h = C.c_void_p()
    rc = L.pdm_create_from_text(fgi.replace("intg_point\n 7", "intg_point\n 2").encode(), C.byref(o32), C.byref(h))
    assert rc == 0  # fixed Gaussian integration with FP32 bond arithmetic
    L.pdm_destroy(h)
    h = C.c_void_p()
    o = pm.options()
    assert L.pdm_create(b"/nonexistent/pds_input.txt", C.byref(o), C.byref(h)) == -2
    rc, h = create(text)
    assert rc == 0
    # calls that need pdm_init fail cleanly before it
    assert L.pdm_step_rk4(h, 1, None) == -6
    L.pdm_destroy(h)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_deck_writer_roundtrip_values():
    spec = pm.bar_spec(5, 4, 3, E=90e9, rho=2000.0, steps=17, safety=0.5)
    inf = info(pm.deck_text(spec, crlf=True))
    assert list(inf.grid_dimension) == [5, 4, 3]
    assert inf.total_step == 17 and inf.safety_factor == 0.5
    assert inf.num_bcs == 2 and inf.num_boundary_spaces == 2


def _run_cli(wkdir, *args):
    exe = os.path.join(ROOT, "pdmtel_amd", "pdm_run")
    return subprocess.run([exe, str(wkdir), *args], capture_output=True, text=True, timeout=60)


def test_cli_quit_option_and_errors(tmp_path):
    """pdm_run parses the deck through the C++ facade; option 3 quits like the reference menu."""
    import shutil

    shutil.copy(os.path.join(GOLDEN, "pds_input_sample.txt"), tmp_path / "pds_input.txt")
    r = _run_cli(tmp_path, "3")
    assert r.returncode == 0, r.stderr
    assert "user quits without doing any calculation!" in (tmp_path / "pds_output.dat").read_text()
    r = _run_cli(tmp_path / "missing", "1")
    assert r.returncode == 1 and "code -2" in r.stderr
    (tmp_path / "pds_input.txt").write_text(raw_without(open(os.path.join(GOLDEN, "pds_input_sample.txt")).read(), "grid_spacing"))
    r = _run_cli(tmp_path, "1")
    assert r.returncode == 1 and "code -3" in r.stderr


def raw_without(text, keyword):
    lines = text.splitlines()
    i = next(k for k, l in enumerate(lines) if l.strip() == keyword)
    return "\n".join(lines[:i] + lines[i + 2:])
