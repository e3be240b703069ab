"""GPU: the post-processing of a tecplot dump, against the oracle.

  - the areal force density (CalcArealForceDensity, pdDataManager3.cpp:442-620) at every node
    of a coordinate plane and at single nodes (corners, faces, next to a slit), FP64, with the
    state's bdryEff factors and with the factors a dump carries, within 1e-12 of the largest
    component (the same per-bond expressions; only the order of the sums differs);
  - the pdm_run menu options 4 (PostAFD, pdSolve.cpp:810-834) and 5 (PostFormatData,
    pdSolve.cpp:782-808, pdDataManager3.cpp:940-1099) on a dump written by pdm_write_tecplot:
    the prompts are answered on stdin, pds_AFD_<step>.dat and pds_format_<step>.dat are read
    back and checked against the oracle's AFD on the dump's (rounded) displacements and
    factors, and against the dump's own columns, to the 6 significant digits the files hold.
"""

import os
import subprocess

import numpy as np
import pytest

import pdmtel_amd as pm
import pyoracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pdmtel_amd", "pdm_run")
DX = 1e-3


def afd_spec():
    """A 20x10x10 lattice bar (no jitter, so planes hold whole node layers) with an edge crack."""
    spec = pm.bar_spec(20, 10, 10, steps=20)
    spec["slits"] = [pm.mid_slit(spec, 30.0, edge=True)]
    spec["traction_body"] = 0
    return spec


def field(X, rng):
    """A smooth strain field plus noise: stretches of ~1e-4, some bonds compressed."""
    G = rng.uniform(-1e-4, 1e-4, size=(3, 3))
    return X @ G.T + rng.uniform(-2e-8, 2e-8, size=X.shape)


def plane_nodes(X, axis, value):
    return np.nonzero(X[:, axis] == value)[0]


@pytest.mark.parametrize("dump_factors", [False, True])
def test_areal_force_density_plane_and_nodes(dump_factors):
    spec = afd_spec()
    prob = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64)
    orc = pyoracle.Oracle(spec)
    X, _, _ = prob.nodes()
    rp, _, cut = prob.csr()
    assert cut.sum() > 0
    rng = np.random.default_rng(31)
    u = field(X, rng)
    prob.set_state(u, None)
    orc.set_state(u, None)
    be = None
    if dump_factors:  # the 6 significant digits of a tecplot dump, not the state's factors
        be = np.array([float(f"{v:.5e}") for v in prob.bdryeff()])
        assert np.any(be != prob.bdryeff())
    xs = np.unique(X[:, 0])
    sets = {
        "plane x": plane_nodes(X, 0, xs[len(xs) // 2]),
        "plane y": plane_nodes(X, 1, np.unique(X[:, 1])[2]),
        "plane z": plane_nodes(X, 2, np.unique(X[:, 2])[0]),
        "single": np.array([0, prob.n - 1, prob.n // 2, 7, 211, 1203]),
    }
    for name, nodes in sets.items():
        assert len(nodes) > 0
        sig = prob.areal_force_density(nodes, be)
        want = orc.areal_force_density(nodes, be)
        scale = np.abs(want).max()
        err = np.abs(sig - want).max() / scale
        print(f"[afd] {name}: {len(nodes)} nodes, max |sigma| {scale:.3e}, rel err {err:.2e}", flush=True)
        assert scale > 0 and err <= 1e-12
    # the empty request, and an id out of range is refused
    assert prob.areal_force_density(np.zeros(0, dtype=np.int64)).shape == (0, 3, 3)
    with pytest.raises(pm.PdmError):
        prob.areal_force_density([prob.n])


def read_dump(path):
    rows = np.loadtxt(path, skiprows=4)
    ids = rows[:, 0].astype(np.int64)
    out = np.empty_like(rows)
    out[ids] = rows
    return out


def run_menu(wkdir, option, answers):
    r = subprocess.run([EXE, str(wkdir), str(option)], input=answers, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_post_afd_and_format_menu(tmp_path):
    spec = afd_spec()
    (tmp_path / "pds_input.txt").write_text(pm.deck_text(spec))
    prob = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64)
    X, mat, bc = prob.nodes()
    rng = np.random.default_rng(5)
    u = field(X, rng)
    v = rng.uniform(-1e-3, 1e-3, size=X.shape)
    prob.set_state(u, v)
    step = 7
    prob.write_tecplot(str(tmp_path / f"pds_tecplot_{step}.dat"), step)
    d = read_dump(tmp_path / f"pds_tecplot_{step}.dat")
    Ud, Vd, bed = d[:, 4:7], d[:, 7:10], d[:, 12]
    orc = pyoracle.Oracle(spec)
    orc.set_state(Ud, None)

    xs = np.unique(X[:, 0])
    xc = float(xs[len(xs) // 2])  # a Python float: its repr is the shortest round-trip text
    single = 1203
    stdout = run_menu(tmp_path, 4, f"{step}\n1\n1\n{xc!r}\ny\n2\n{single}\nn\n")
    assert "Give the location of the plane" in stdout and "Another run? (y/n)" in stdout
    lines = (tmp_path / f"pds_AFD_{step}.dat").read_text().splitlines()
    assert lines[0] == f" norm = 1, location = {xc:g}"
    plane = plane_nodes(X, 0, xc)
    body = np.array([[float(t) for t in ln.split("\t")] for ln in lines[1:1 + len(plane)]])
    assert np.array_equal(body[:, 0].astype(np.int64), plane)
    want = orc.areal_force_density(plane, bed).reshape(-1, 9)
    scale = np.abs(want).max()
    assert np.all(np.abs(body[:, 4:] - want) <= 1e-5 * np.abs(want) + 1e-12 * scale)
    assert np.allclose(body[:, 1:4], X[plane], rtol=1e-5, atol=1e-12)
    # the single-node block: a header line, then sigma11..sigma33 over three lines
    k = lines.index(next(ln for ln in lines if ln.startswith(" AFD at node")))
    # " sigma11 = a sigma12 = b sigma13 = c": the first word after each '='
    vals = [float(part.split()[0]) for ln in lines[k + 1:k + 4] for part in ln.split("=")[1:]]
    want1 = orc.areal_force_density([single], bed).reshape(9)
    assert np.all(np.abs(np.array(vals) - want1) <= 1e-5 * np.abs(want1) + 1e-12 * scale)
    print(f"[post afd] plane of {len(plane)} nodes and node {single} match the oracle", flush=True)

    ys = np.unique(X[:, 1])
    yc = float(ys[3])
    run_menu(tmp_path, 5, f"{step}\n1\n2\n{yc!r}\ny\n2\n{single}\ny\n3\n1\n{xc!r}\nn\n")
    lines = (tmp_path / f"pds_format_{step}.dat").read_text().splitlines()
    py = plane_nodes(X, 1, yc)
    node_rows = np.array([[float(t) for t in ln.split("\t")] for ln in lines[1:1 + len(py)]])
    assert np.array_equal(node_rows[:, 0].astype(np.int64), py)
    ref = np.column_stack([X[py], Ud[py], Vd[py], mat[py], bc[py], bed[py], np.zeros(len(py))])
    assert np.allclose(node_rows[:, 1:], ref, rtol=1e-5, atol=1e-12)
    k = 1 + len(py) + 1  # the blank line after the plane
    one = np.array([float(t) for t in lines[k].split("\t")])
    assert one[0] == single and np.allclose(one[4:10], np.r_[Ud[single], Vd[single]], rtol=1e-5, atol=1e-12)
    k += 1
    assert lines[k].startswith(" norm = 1, location = ")
    strain = [ln.split("\t") for ln in lines[k + 1:] if ln.strip()]
    ids = np.array([int(s[0]) for s in strain])
    assert np.array_equal(ids, plane)
    for s in strain:  # e11 from the member one spacing further along x, where there is one
        i = int(s[0])
        j = np.nonzero((X[:, 0] == X[i, 0] + DX) & (X[:, 1] == X[i, 1]) & (X[:, 2] == X[i, 2]))[0]
        if len(s) == 5:
            e = (Ud[j[0], 0] - Ud[i, 0]) / (X[j[0], 0] - X[i, 0])
            assert abs(float(s[4]) - e) <= 1e-5 * abs(e) + 1e-12
        else:
            assert len(s) == 4
    print(f"[post format] {len(py)} plane rows, node {single}, {len(strain)} strain rows", flush=True)
