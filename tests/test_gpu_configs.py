"""GPU: the batched-specimen and the large-plate workloads at their full sizes.

Batch (pdmtel_amd.deck.specimen_specs(512, seed=0)): 512 bars with an edge crack each, node
count log-uniform in [1e3, 1e5] (every side at least 5 nodes), aspect ratios uniform in
[1, 5], delta / dx from {3.015, 4.015}, crack normal uniform in [0, 90] degrees, E uniform
in [50, 200] GPa; one launch per RK4 stage over all of them, graph-captured chunks, tile
layout from the autotuner; 1,000 steps in FP64, FP32C and FP32C2.
  - every specimen's FP32C and FP32C2 state against its FP64 state, max |U32 - U64| /
    max |U64| at the last step, within 1e-6 for every specimen;
  - sampled specimens (the thinnest delta = 4.015 dx ones among them) bit-identical to the
    same deck run alone with the batch's layout;
  - the three smallest specimens' FP64 trajectories against the oracle over all 1,000 steps.

Plate (pdmtel_amd.deck.plate_spec): 400 x 400 x 50 nodes, jitter 0.01 dx from seed 0, two
crossing part-through slits at the centre, 1.93e9 family entries; 500 steps.
  - the whole bond set: every family ascending, the relation symmetric (j in F(i) iff i in
    F(j), by two independent random-weight checksums over all entries);
  - complete rows of sampled blocks of nodes (a crack tip, the slit crossing, a corner, an
    edge, the interior) and of random nodes, with their cut flags and bdryEff factors,
    bit-identical to the oracle's search of those rows;
  - FP64 forces of sampled rows within 1e-12 of the oracle's for a random field;
  - FP32C against FP64 after 500 steps, max relative displacement error <= 1e-6.
"""

import numpy as np
import pytest

import pdmtel_amd as pm
import pyoracle
from pdmtel_amd.deck import plate_spec, specimen_specs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


# ---------------------------------------------------------------- batch of 512 specimens
BATCH_STEPS = 1000


def test_batch_512_specimens():
    specs = specimen_specs(512, seed=0, steps=BATCH_STEPS)
    decks = [pm.deck_text(s) for s in specs]
    runs = {}
    for prec, label in ((pm.PDM_FP64, "fp64"), (pm.PDM_FP32C, "fp32c")):
        print(f"[batch512] {label}: init", flush=True)
        b = pm.Batch(decks, precision=prec)
        print(f"[batch512] {label}: nodes={sum(b.n)} entries={sum(b.E)} live={sum(b.E_live)} "
              f"tuning={b.tuning()}", flush=True)
        for chunk in range(0, BATCH_STEPS, 250):
            assert b.step(250) == 250
            ms, launches = b.kernel_timing()
            print(f"[batch512] {label}: step {chunk + 250}, stage {ms:.3f} ms", flush=True)
        runs[label] = b
    b64, b32 = runs["fp64"], runs["fp32c"]
    errs = np.zeros(len(specs))
    for k in range(len(specs)):
        U64, _ = b64.state(k)
        U32, _ = b32.state(k)
        assert np.isfinite(U32).all() and np.isfinite(U64).all()
        assert b64.clock(k)[1] == b32.clock(k)[1] == BATCH_STEPS
        errs[k] = rel(U32, U64)
    over = np.flatnonzero(errs > 1e-6)
    print(f"[batch512] fp32c vs fp64 after {BATCH_STEPS} steps: median {np.median(errs):.3e} "
          f"p99 {np.quantile(errs, 0.99):.3e} max {errs.max():.3e}; over 1e-6: {len(over)} of {len(specs)} "
          f"{[(int(k), specs[k]['grid_dimension'], f'{errs[k]:.2e}') for k in over]}", flush=True)
    assert np.median(errs) <= 1e-7
    assert len(over) == 0 and errs.max() <= 1e-6

    # the two-float (compensated) family sums on the same batch
    print("[batch512] fp32c2: init", flush=True)
    b2 = pm.Batch(decks, precision=pm.PDM_FP32C2)
    assert b2.step(BATCH_STEPS) == BATCH_STEPS
    errs2 = np.array([rel(b2.state(k)[0], b64.state(k)[0]) for k in range(len(specs))])
    b2.close()
    print(f"[batch512] fp32c2 vs fp64 after {BATCH_STEPS} steps: median {np.median(errs2):.3e} "
          f"max {errs2.max():.3e}; over 1e-6: {int((errs2 > 1e-6).sum())}", flush=True)
    assert np.median(errs2) <= 1e-7 and errs2.max() <= 1e-6

    # sampled specimens bit-identical to their decks run alone with the layout of their tile
    # class in the batch (both horizons, so both classes)
    thin = sorted(range(len(specs)), key=lambda k: min(specs[k]["grid_dimension"]))
    thin = [k for k in thin if specs[k]["horizon"] > 4e-3][:3]
    sample = sorted(set(thin) | {0, 255, 511} | {int(k) for k in over[:2]})
    assert len({specs[k]["horizon"] for k in sample}) == 2
    for k in sample:
        p = pm.Problem(decks[k], precision=pm.PDM_FP32C, **b32.tuning(k))
        assert p.step(BATCH_STEPS) == BATCH_STEPS
        U, V = p.state()
        Ub, Vb = b32.state(k)
        print(f"[batch512] specimen {k} {specs[k]['grid_dimension']} alone == in batch", flush=True)
        assert np.array_equal(U, Ub) and np.array_equal(V, Vb), k
        assert p.clock() == b32.clock(k) and p.energy() == b32.energy(k), k
        p.close()
    b32.close()

    # FP64 batch against the oracle on the smallest specimens, whole run
    small = sorted(range(len(specs)), key=lambda k: b64.n[k])[:3]
    pyoracle.set_square_mode(True)
    try:
        for k in small:
            o = pyoracle.Oracle(specs[k], bond_method=1)
            o.run(BATCH_STEPS)
            U, V = b64.state(k)
            Uo, Vo = o.state()
            eu, ev = rel(U, Uo), rel(V, Vo)
            print(f"[batch512] specimen {k} fp64 vs oracle after {BATCH_STEPS} steps: eU={eu:.2e} eV={ev:.2e}",
                  flush=True)
            assert eu <= 1e-9 and ev <= 1e-9
    finally:
        pyoracle.set_square_mode(False)
    b64.close()


# ---------------------------------------------------------------- 8M-node plate
PLATE_STEPS = 500


def bond_set_checksums(rp, col, n, chunk=250000):
    """(symmetric, ascending) over the whole CSR: symmetry from two random-weight checksums,
    sum_e f(i) g(j) == sum_e f(j) g(i) (mod 2^64), which a missing or extra directed entry
    breaks with probability ~1 - 2^-64."""
    rng = np.random.default_rng(12345)
    f = rng.integers(1, 2 ** 63, size=n, dtype=np.uint64)
    g = rng.integers(1, 2 ** 63, size=n, dtype=np.uint64)
    s1 = np.uint64(0)
    s2 = np.uint64(0)
    asc = True
    with np.errstate(over="ignore"):
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            e0, e1 = int(rp[a]), int(rp[b])
            c = col[e0:e1]
            off = (rp[a:b] - e0).astype(np.int64)
            lens = np.diff(rp[a:b + 1])
            assert (lens > 0).all()
            s1 += np.sum(f[a:b] * np.add.reduceat(g[c], off))
            s2 += np.sum(g[a:b] * np.add.reduceat(f[c], off))
            d = np.diff(c)
            keep = np.ones(len(d), dtype=bool)
            starts = off[1:] - 1
            keep[starts] = False
            asc = asc and bool(np.all(d[keep] > 0))
    return s1 == s2, asc


def plate_rows(spec, X, rng):
    """Node ids of whole blocks (8 x 8 x 6 lattice cells around a crack tip, the slit
    crossing, a corner, an edge midpoint, the interior) plus random nodes."""
    nx, ny, nz = spec["grid_dimension"]
    dx = spec["grid_spacing"]
    ids = np.arange(nx * ny * nz)
    lx, my, kz = ids % nx, (ids // nx) % ny, ids // (nx * ny)
    cx, cy = nx // 2, ny // 2
    tip = int(round(0.25 * nx))  # crack_fraction 0.5: each slit spans half the width
    blocks = [(cx + tip, cy, nz // 2), (cx, cy, nz // 2), (0, 0, 0), (nx // 2, 0, nz - 1),
              (nx // 4, ny // 4, nz // 2), (cx, cy - tip, 2)]
    sel = np.zeros(len(ids), dtype=bool)
    for (bx, by, bz) in blocks:
        sel |= (np.abs(lx - bx) <= 4) & (np.abs(my - by) <= 4) & (np.abs(kz - bz) <= 3)
    rows = np.concatenate([np.flatnonzero(sel), rng.choice(len(ids), 3000, replace=False)])
    assert len(X) == len(ids)
    return np.unique(rows)


def test_plate_8m_nodes():
    spec = plate_spec(400, 400, 50, jitter=0.01, seed=0, steps=PLATE_STEPS)
    print("[plate] fp32c init", flush=True)
    p32 = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP32C)
    print(f"[plate] n={p32.n} E={p32.E} live={p32.E_live} tuning={p32.tuning()} launch={p32.stage_launch()}",
          flush=True)
    assert p32.n == 8_000_000 and p32.E > 1.9e9
    rp, col, cut = p32.csr()
    X, _, _ = p32.nodes()
    sym, asc = bond_set_checksums(rp, col, p32.n)
    print(f"[plate] whole bond set: symmetric={sym} ascending={asc} cut={int(cut.sum())}", flush=True)
    assert sym and asc and cut.sum() > 0

    rows = plate_rows(spec, X, np.random.default_rng(7))
    orp, ocol, ocut, obe = pyoracle.sample_rows(spec, rows)
    be = p32.bdryeff()
    for k, i in enumerate(rows):
        assert np.array_equal(col[rp[i]:rp[i + 1]], ocol[orp[k]:orp[k + 1]]), i
        assert np.array_equal(cut[rp[i]:rp[i + 1]], ocut[orp[k]:orp[k + 1]]), i
    assert np.array_equal(be[rows], obe)
    ncut = int(sum(ocut[orp[k]:orp[k + 1]].sum() for k in range(len(rows))))
    print(f"[plate] {len(rows)} sampled rows bit-identical to the oracle ({int(orp[-1])} entries, "
          f"{ncut} cut, {int((obe != 1.0).sum())} rows with bdryEff != 1)", flush=True)
    assert ncut > 0
    del col, cut

    print("[plate] fp64 init", flush=True)
    p64 = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP64)
    u = np.random.default_rng(8).uniform(-1e-6, 1e-6, size=(p64.n, 3))
    F = p64.forces(u)
    pyoracle.set_square_mode(True)
    try:
        Fo, _ = pyoracle.sample_forces(spec, rows, u)
    finally:
        pyoracle.set_square_mode(False)
    ef = rel(F[rows], Fo)
    print(f"[plate] fp64 forces of sampled rows vs oracle: {ef:.3e}", flush=True)
    assert ef <= 1e-12
    del u, F

    for p, label in ((p64, "fp64"), (p32, "fp32c")):
        for chunk in range(0, PLATE_STEPS, 100):
            assert p.step(100) == 100
        ms, launches = p.kernel_timing()
        print(f"[plate] {label}: {PLATE_STEPS} steps, stage {ms:.3f} ms", flush=True)
    U64, _ = p64.state()
    U32, _ = p32.state()
    err = rel(U32, U64)
    print(f"[plate] fp32c vs fp64 after {PLATE_STEPS} steps: {err:.3e}", flush=True)
    assert np.isfinite(U32).all() and err <= 1e-6
