"""Full-size configurations against the oracle's outputs stored in
tests/golden/golden.json (written by tests/golden/make_golden.py on the same
seeded inputs):

  - the 256x256x160 VecPixel<f32,3> field (spacing 0.7/0.7/2.5, sigma 3);
  - the 64 batched volumes (extents 64..320, per-volume sigma and spacing),
    through a batched plan and through plain per-volume calls;
  - 256^3 sigma 1.8 forward and adjoint;
  - the batched adjoint over 16 uneven gradients (extents 48..192).

EXACT outputs must hash to the oracle's output (bit-identical); FAST outputs
and f32 adjoints must lie within 1e-5 of the oracle at the sampled voxels,
relative to the max |value| of the whole output; the f64 adjoint within 1e-12.
"""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL_F32 = 1e-5
TOL_F64 = 1e-12


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as fh:
        return json.load(fh)


def _sha(t: torch.Tensor) -> str:
    return hashlib.sha256(t.contiguous().cpu().numpy().tobytes()).hexdigest()


def _f32_at(entry, scale_of: np.ndarray):
    idx = np.array(entry["sample_flat_index"], dtype=np.int64)
    ref = np.array(entry["sample_value_bits"], dtype=np.uint32).view(np.float32)
    return idx, ref.astype(np.float64), float(np.abs(scale_of).max())


def _check_f32_samples(got: torch.Tensor, entry, exact_out: np.ndarray, tol=TOL_F32):
    idx, ref, scale = _f32_at(entry, exact_out)
    g = got.contiguous().cpu().numpy().reshape(-1)[idx].astype(np.float64)
    err = float(np.abs(g - ref).max()) / scale
    assert err <= tol, err
    return err


def _check_f64_samples(got: torch.Tensor, entry, tol):
    idx = np.array(entry["sample_flat_index"], dtype=np.int64)
    ref = np.array(entry["sample_value"], dtype=np.float64)
    g = got.contiguous().cpu().numpy().reshape(-1)[idx].astype(np.float64)
    err = float(np.abs(g - ref).max()) / entry["max_abs"]
    assert err <= tol, err
    return err


def test_c3_vecpixel_full_size(itk, dev, gold):
    from itk_rs_amd import workloads as wl
    g = gold["c3"]
    v, sigma, sp = wl.c3("cpu")
    assert hashlib.sha256(v.numpy().tobytes()).hexdigest() == g["input_sha256"]
    vd = v.to(dev)
    exact = itk.gaussian(vd, sigma, sp, ndim=3, arith=itk.EXACT)
    torch.cuda.synchronize()
    assert _sha(exact) == g["output_sha256"]
    ex = exact.cpu().numpy()
    idx = np.array(g["sample_flat_index"], dtype=np.int64)
    assert np.array_equal(ex.reshape(-1)[idx].view(np.uint32),
                          np.array(g["sample_value_bits"], dtype=np.uint32))
    fast = itk.gaussian(vd, sigma, sp, ndim=3, arith=itk.FAST)
    torch.cuda.synchronize()
    _check_f32_samples(fast, g, ex)
    # the whole FAST field against the (bit-identical) EXACT field
    assert float((fast.double() - exact.double()).abs().max()) / float(exact.abs().max()) <= TOL_F32


@pytest.fixture(scope="module")
def c4_inputs(dev, gold):
    from itk_rs_amd import workloads as wl
    vols, params = wl.c4(dev)
    gv = gold["c4"]["volumes"]
    assert len(vols) == len(gv) == 64
    for v, p, g in zip(vols, params, gv):
        assert list(v.shape) == g["shape_zyx"] and p[1] == g["sigma"] and p[2][0] == g["spacing"]
        assert _sha(v) == g["input_sha256"]
    return vols, params


def test_c4_batched_exact_full_size(itk, dev, gold, c4_inputs):
    """Every volume of the 64-volume batched EXACT plan hashes to the
    oracle's output, on the captured first run and on a replay."""
    vols, params = c4_inputs
    gv = gold["c4"]["volumes"]
    radii = {g["radius"] for g in gv}
    assert max(radii) >= 16  # the configuration reaches past radius 15
    outs = [torch.full_like(v, float("nan")) for v in vols]
    plan = itk.BatchPlan(vols, outs, [p[1] for p in params], [p[2] for p in params],
                         direction=itk.BATCH_FORWARD, arith=itk.EXACT)
    s = torch.cuda.Stream()
    for rep in range(2):
        for o in outs:
            o.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.run(s)
        s.synchronize()
        for k, (o, g) in enumerate(zip(outs, gv)):
            assert _sha(o) == g["output_sha256"], (rep, k, g["shape_zyx"], g["radius"])
    plan.close()


def test_c4_batched_fast_full_size(itk, dev, gold, c4_inputs):
    """The autotuned 64-volume FAST plan, and the plain FAST loop, within
    1e-5 of the oracle at every volume's sampled voxels; batched and plain
    FAST outputs are bit-identical (same tap chains)."""
    vols, params = c4_inputs
    gv = gold["c4"]["volumes"]
    outs = [torch.empty_like(v) for v in vols]
    plan = itk.BatchPlan(vols, outs, [p[1] for p in params], [p[2] for p in params],
                         direction=itk.BATCH_FORWARD, arith=itk.FAST)
    s = torch.cuda.Stream()
    plan.autotune(s)
    plan.run(s)
    s.synchronize()
    worst = 0.0
    for k, (v, o, p, g) in enumerate(zip(vols, outs, params, gv)):
        exact = itk.gaussian(v, p[1], p[2], arith=itk.EXACT)
        plain = itk.gaussian(v, p[1], p[2], arith=itk.FAST)
        torch.cuda.synchronize()
        assert _sha(exact) == g["output_sha256"], (k, g["radius"])
        ex = exact.cpu().numpy()
        worst = max(worst, _check_f32_samples(o, g, ex))
        _check_f32_samples(plain, g, ex)
        assert torch.equal(o, plain), (k, g["shape_zyx"], g["radius"])
    plan.close()
    print(f"c4 FAST worst sampled error {worst:.3g}")


def test_c5_forward_and_adjoint_full_size(itk, dev, gold):
    from itk_rs_amd import workloads as wl
    g = gold["c5"]
    x, gr, sigma, sp = wl.c5("cpu")
    assert hashlib.sha256(x.numpy().tobytes()).hexdigest() == g["input_sha256"]
    assert hashlib.sha256(gr.numpy().tobytes()).hexdigest() == g["grad_sha256"]
    xd, gd = x.to(dev), gr.to(dev)
    exact = itk.gaussian(xd, sigma, sp, arith=itk.EXACT)
    fast = itk.gaussian(xd, sigma, sp, arith=itk.FAST)
    adj = itk.gaussian_backward(gd, sigma, sp)
    adj64 = itk.gaussian_backward(gd.double(), sigma, sp)
    torch.cuda.synchronize()
    assert _sha(exact) == g["forward"]["output_sha256"]
    _check_f32_samples(fast, g["forward"], exact.cpu().numpy())
    _check_f64_samples(adj, g["adjoint_f64"], TOL_F32)
    _check_f64_samples(adj64, g["adjoint_f64"], TOL_F64)


def test_c5_batched_adjoint_full_size(itk, dev, gold):
    """The 16-gradient batched adjoint plan within 1e-5 of the oracle's f64
    transpose at every volume's sampled voxels, before and after autotuning."""
    from itk_rs_amd import workloads as wl
    gv = gold["c5b"]["volumes"]
    grads, params = wl.c5_batch("cpu")
    assert len(grads) == len(gv) == 16
    for q, g in zip(grads, gv):
        assert list(q.shape) == g["shape_zyx"]
        assert hashlib.sha256(q.numpy().tobytes()).hexdigest() == g["grad_sha256"]
    gd = [q.to(dev) for q in grads]
    outs = [torch.full_like(q, float("nan")) for q in gd]
    plan = itk.BatchPlan(gd, outs, [p[1] for p in params], [p[2] for p in params],
                         direction=itk.BATCH_BACKWARD)
    s = torch.cuda.Stream()
    for tune in (False, True):
        if tune:
            plan.autotune(s)
        plan.run(s)
        s.synchronize()
        for o, g in zip(outs, gv):
            _check_f64_samples(o, g, TOL_F32)
    plan.close()
