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


def _radius_volumes(dev, radii, seed, lo=8, hi=90):
    """One volume per radius (spacing 1, sigma (r - 0.5) / 3 gives radius r),
    extents drawn in [lo, hi): tiles past the volume edge, extents below the
    tile and below 2r + 1."""
    rng = np.random.default_rng(seed)
    vols, sig = [], []
    for r in radii:
        shape = tuple(int(x) for x in rng.integers(lo, hi, size=3))
        vols.append(torch.rand(shape, device=dev))
        sig.append((r - 0.5) / 3.0)
    return vols, sig


@pytest.mark.parametrize("direction", [0, 1])
def test_single_launch_every_radius(itk, dev, direction):
    """FAST forward (radii 1..24, all in the one persistent launch) and the
    adjoint (radii 1..15 in it, 16..20 on the per-axis path beside it): every
    volume equals its plain result bit for bit, on the captured first run,
    on replays, and after autotuning (each run starts from NaN-filled
    outputs, so queue heads not reset between launches would show)."""
    radii = list(range(1, 25)) if direction == 0 else list(range(1, 21))
    vols, sig = _radius_volumes(dev, radii, 21 + direction)
    for v, r, s in zip(vols, radii, sig):
        assert itk.kernel_radius(s, 1.0) == r
    sps = [(1.0, 1.0, 1.0)] * len(vols)
    outs = [torch.full_like(v, float("nan")) for v in vols]
    plan = itk.BatchPlan(vols, outs, sig, sps, direction=direction, arith=itk.FAST)
    refs = [_plain(itk, v, s, sp, direction, itk.FAST) for v, s, sp in zip(vols, sig, sps)]
    torch.cuda.synchronize()
    st = torch.cuda.Stream()
    for rep in range(5):
        if rep == 3:
            plan.autotune(st)
        for o in outs:
            o.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.run(st)
        st.synchronize()
        for o, ref, r in zip(outs, refs, radii):
            assert torch.equal(o, ref), (direction, rep, r, tuple(o.shape))
    plan.close()


def test_single_launch_large_volumes_and_tiny_ones(itk, dev):
    """Volumes large enough to be split along z next to volumes smaller than
    one tile, in one plan."""
    shapes = [(200, 150, 130), (3, 5, 7), (1, 1, 1), (64, 64, 64), (65, 33, 129), (9, 300, 70)]
    sig = [2.0, 1.2, 0.9, 3.7, 6.1, 4.4]
    vols = [torch.rand(sh, device=dev) for sh in shapes]
    sps = [(1.0, 1.0, 1.0)] * len(vols)
    outs = [torch.full_like(v, float("nan")) for v in vols]
    plan = itk.BatchPlan(vols, outs, sig, sps, direction=0, arith=itk.FAST)
    st = torch.cuda.Stream()
    for _ in range(2):
        plan.run(st)
        st.synchronize()
        for v, o, s in zip(vols, outs, sig):
            assert torch.equal(o, itk.gaussian(v, s, (1.0,) * 3, arith=itk.FAST)), (tuple(v.shape), s)
    plan.close()
