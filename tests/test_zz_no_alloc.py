"""GPU: pdm_compute_forces_noalloc (device buffers, no allocation in the call) through the C ABI
against pdm_compute_forces (host buffers) on the same fields: bit-identical forces and energy
densities, for two calls in a row on one GPU, the second on a smaller problem (the shared
scratch keeps the first call's contents beyond what the second uses), then the first problem
again; a problem beyond the scratch's node limit is refused with PDM_E_LIMIT before any work."""

import numpy as np
import pytest
import torch

import pdmtel_amd as pm

pytestmark = pytest.mark.gpu


def noalloc_forces(prob, u):
    ud = torch.from_numpy(np.ascontiguousarray(u, dtype=np.float64).reshape(-1)).to("cuda")
    Fd = torch.full((3 * prob.n,), float("nan"), dtype=torch.float64, device="cuda")
    ed = torch.full((prob.n,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()  # the library runs on its own stream
    prob.forces_noalloc(ud.data_ptr(), Fd.data_ptr(), ed.data_ptr(), sync=True)
    return Fd.cpu().numpy().reshape(-1, 3), ed.cpu().numpy()


def test_noalloc_matches_compute_forces():
    big = pm.bar_spec(40, 20, 20, jitter=0.01, seed=1)
    big["slits"] = [pm.mid_slit(big, 30.0)]
    small = pm.bar_spec(12, 6, 6, jitter=0.01, seed=2)
    pa = pm.Problem(pm.deck_text(big), precision=pm.PDM_FP32C)
    pb = pm.Problem(pm.deck_text(small), precision=pm.PDM_FP64)
    assert pb.n < pa.n
    rng = np.random.default_rng(5)
    ua = rng.uniform(-1e-6, 1e-6, size=(pa.n, 3))
    ub = rng.uniform(-1e-6, 1e-6, size=(pb.n, 3))
    for prob, u, name in ((pa, ua, "big fp32c"), (pb, ub, "small fp64"), (pa, 2.0 * ua, "big fp32c again")):
        print(f"[noalloc] {name}: n={prob.n}", flush=True)
        F, el = noalloc_forces(prob, u)
        Fh, elh = prob.forces(u, with_elast=True)
        assert np.array_equal(F, Fh) and np.array_equal(el, elh), name
        assert np.isfinite(F).all() and np.abs(F).max() > 0


def test_noalloc_refuses_beyond_its_scratch():
    spec = pm.bar_spec(100, 80, 70)  # 560,000 nodes > 524,288
    prob = pm.Problem(pm.deck_text(spec), precision=pm.PDM_FP32C)
    with pytest.raises(pm.PdmError) as ei:
        prob.forces_noalloc(0, 0, 0)  # null device pointers are refused first
    assert ei.value.code == pm.PDM_E_ARG
    d = torch.zeros(3 * prob.n, dtype=torch.float64, device="cuda")
    with pytest.raises(pm.PdmError) as ei:
        prob.forces_noalloc(d.data_ptr(), d.data_ptr(), 0)
    assert ei.value.code == pm.PDM_E_LIMIT
