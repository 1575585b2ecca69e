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


def test_batch_autotune_keeps_results(itk, dev, tmp_path, monkeypatch):
    from itk_rs_amd import workloads as wl
    monkeypatch.setenv("ITKG_TUNE_CACHE", str(tmp_path / "tune.txt"))
    vols, params = wl.c4(dev, nvol=6, lo=20, hi=90, seed=8)
    outs = [torch.empty_like(v) for v in vols]
    plan = itk.BatchPlan(vols, outs, [p[1] for p in params], [p[2] for p in params])
    s = torch.cuda.Stream()
    plan.run(s)
    s.synchronize()
    first = [o.clone() for o in outs]
    plan.autotune(s)
    s.synchronize()
    for o in outs:
        o.zero_()
    torch.cuda.synchronize()
    plan.run(s)
    s.synchronize()
    for a, b in zip(first, outs):
        assert torch.equal(a, b)
    lines = (tmp_path / "tune.txt").read_text().split("\n")
    assert any(len(ln.split()) == 5 for ln in lines)   # direction arith radius tz tile_rows
    # a new plan picks the cached choice up and gives the same results
    outs2 = [torch.empty_like(v) for v in vols]
    plan2 = itk.BatchPlan(vols, outs2, [p[1] for p in params], [p[2] for p in params])
    plan2.run(s)
    s.synchronize()
    for a, b in zip(first, outs2):
        assert torch.equal(a, b)
    plan.close()
    plan2.close()


@pytest.mark.parametrize("direction", [0, 1])
def test_batch_per_volume_z_run(itk, dev, tmp_path, monkeypatch, direction):
    """A cached z run of 0 (each volume takes its own plain launch's z run)
    gives every volume's plain result, for both tile shapes."""
    from itk_rs_amd import workloads as wl
    vols, params = wl.c4(dev, nvol=8, lo=20, hi=120, seed=11)
    arith = itk.FAST
    for rows in (64, 32):
        cache = tmp_path / ("tune%d_%d.txt" % (direction, rows))
        cache.write_text("".join("%d %d %d 0 %d\n" % (direction, arith, r, rows)
                                 for r in range(1, 16)))
        monkeypatch.setenv("ITKG_TUNE_CACHE", str(cache))
        outs = [torch.full_like(v, float("nan")) for v in vols]
        plan = itk.BatchPlan(vols, outs, [p[1] for p in params], [p[2] for p in params],
                             direction=direction, arith=arith)
        s = torch.cuda.Stream()
        for rep in range(2):
            plan.run(s)
            s.synchronize()
            for v, o, (shape, sigma, sp) in zip(vols, outs, params):
                ref = _plain(itk, v, sigma, sp, direction, arith)
                torch.cuda.synchronize()
                assert torch.equal(o, ref), (rows, rep, shape, sigma)
        plan.close()


def test_batched_adjoint_uneven_gradients(itk, dev):
    """The batched adjoint over uneven gradient volumes (sigma 1.8) equals the
    plain adjoint of each, before and after autotuning."""
    from itk_rs_amd import workloads as wl
    grads, params = wl.c5_batch(dev, nvol=6)
    outs = [torch.full_like(g, float("nan")) for g in grads]
    plan = itk.BatchPlan(grads, outs, [p[1] for p in params], [p[2] for p in params],
                         direction=itk.BATCH_BACKWARD)
    s = torch.cuda.Stream()
    for tune in (False, True):
        if tune:
            plan.autotune(s)
        plan.run(s)
        s.synchronize()
        for g, o, (shape, sigma, sp) in zip(grads, outs, params):
            ref = itk.gaussian_backward(g, sigma, sp)
            torch.cuda.synchronize()
            assert torch.equal(o, ref), (shape, tune)
    plan.close()
