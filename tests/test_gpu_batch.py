"""Batched many-volume launch: every volume's output equals its plain
single-volume output bit for bit, on the first (captured) and later
(replayed) runs."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _plain(itk, x, sigma, sp, direction, arith):
    if direction == 0:
        return itk.gaussian(x, sigma, sp, arith=arith)
    return itk.gaussian_backward(x, sigma, sp)


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("arith", [0, 1])
def test_batch_equals_plain(itk, dev, direction, arith):
    from itk_rs_amd import workloads as wl
    vols, params = wl.c4(dev, nvol=12, lo=16, hi=70, seed=5)
    if direction == 1 and arith == 1:
        pytest.skip("the adjoint has one arithmetic")
    outs = [torch.full_like(v, float("nan")) for v in vols]
    plan = itk.BatchPlan(vols, outs, [p[1] for p in params], [p[2] for p in params],
                         direction=direction, arith=arith)
    assert plan.voxels == sum(v.numel() for v in vols)
    s = torch.cuda.Stream()
    for rep in range(3):
        for o in outs:
            o.fill_(float("nan"))
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            plan.run(s)
        s.synchronize()
        for v, o, (shape, sigma, sp) in zip(vols, outs, params):
            ref = _plain(itk, v, sigma, sp, direction, arith)
            torch.cuda.synchronize()
            assert torch.equal(o, ref), (rep, shape, sigma)
    plan.close()


def test_batch_validates(itk, dev):
    x = torch.zeros((4, 4, 4), device=dev)
    with pytest.raises(itk.ItkgError):
        itk.BatchPlan([x], [x.clone()], [200.0], [(1.0, 1.0, 1.0)])
