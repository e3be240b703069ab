"""GPU: the whole-driver entry points (pdm_solve_dynamic through the pdm_run CLI
and the C++ facade, tecplot output) against the oracle's trajectory."""

import os
import subprocess

import numpy as np
import pytest

import pdmtel_amd as pm
import pyoracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_tecplot(path):
    rows = []
    with open(path) as f:
        lines = f.read().splitlines()
    assert lines[0] == "Tecplot formatting output mode"
    assert lines[1] == "VARIABLES = "
    assert lines[3].startswith("ZONE I=")
    for l in lines[5:]:
        if l.strip():
            rows.append([float(v) for v in l.split()])
    return np.array(rows)


def test_cli_dynamic_run_matches_oracle(tmp_path):
    spec = pm.bar_spec(12, 6, 6, steps=5, jitter=0.01, seed=1)
    spec["step_dump_frequency"] = 2
    (tmp_path / "pds_input.txt").write_text(pm.deck_text(spec, crlf=True))
    exe = os.path.join(ROOT, "pdmtel_amd", "pdm_run")
    print(f"[cli] {exe} {tmp_path} 1", flush=True)
    r = subprocess.run([exe, str(tmp_path), "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    files = sorted(p.name for p in tmp_path.iterdir() if p.name.startswith("pds_tecplot_"))
    assert files == ["pds_tecplot_1.dat", "pds_tecplot_2.dat", "pds_tecplot_4.dat", "pds_tecplot_5.dat"]
    out = (tmp_path / "pds_output.dat").read_text()
    assert "** Dynamic solution: starting from cycle 1 **" in out and "\n5: dt= " in out
    assert "The total CPU tick = " in out
    orc = pyoracle.Oracle(spec, bond_method=1)
    orc.run(5)
    Uo, Vo = orc.state()
    T = read_tecplot(tmp_path / "pds_tecplot_5.dat")
    assert T.shape == (orc.n, 14)
    assert np.array_equal(T[:, 0], np.arange(orc.n))  # 0-based ids, pdDataManager.cpp:901
    # 5 significant digits in the file
    assert np.allclose(T[:, 4:7], Uo, rtol=2e-5, atol=2e-5 * np.abs(Uo).max())
    assert np.allclose(T[:, 7:10], Vo, rtol=2e-5, atol=2e-5 * np.abs(Vo).max())


def test_cli_static_run(tmp_path):
    spec = pm.bar_spec(12, 6, 6, jitter=0.01, seed=1)
    (tmp_path / "pds_input.txt").write_text(pm.deck_text(spec))
    exe = os.path.join(ROOT, "pdmtel_amd", "pdm_run")
    r = subprocess.run([exe, str(tmp_path), "2", "--fp32c"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "** Static solution: matrix-free PCG" in (tmp_path / "pds_output.dat").read_text()
    orc = pyoracle.Oracle(spec, bond_method=1)
    Ud = orc.solve_static_dense()
    T = read_tecplot(tmp_path / "pds_tecplot_1.dat")
    assert np.allclose(T[:, 4:7], Ud, rtol=2e-5, atol=2e-5 * np.abs(Ud).max())
