"""The committed golden fixtures are the oracle's outputs: a few entries of
every configuration are recomputed here on the host (the smallest volumes,
to keep the CPU suite short) and must match tests/golden/golden.json."""
import hashlib
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as fh:
        return json.load(fh)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_fixture_keys(gold):
    for k in ("c1", "c2", "c3", "c4", "c5", "c5b"):
        assert k in gold
    assert len(gold["c4"]["volumes"]) == 64
    assert len(gold["c5b"]["volumes"]) == 16


def test_c4_smallest_volumes_are_oracle_outputs(itk, orc, gold):
    from itk_rs_amd import workloads as wl
    params = wl.c4_params()
    gv = gold["c4"]["volumes"]
    order = sorted(range(64), key=lambda v: int(np.prod(params[v][0])))
    # the two smallest volumes, and the smallest with radius >= 16
    big_r = [v for v in order if gv[v]["radius"] >= 16][:1]
    for v in order[:2] + big_r:
        shape, sigma, sp = params[v]
        g = gv[v]
        assert list(shape) == g["shape_zyx"] and sigma == g["sigma"]
        assert orc.kernel_radius(sigma, sp[0]) == g["radius"]
        x = wl.uniform(3, 1000 + v, shape, "cpu").numpy()
        assert _sha(x) == g["input_sha256"]
        size = list(reversed(shape))
        y = orc.generate_region(x, [0, 0, 0], sp, sigma, [0, 0, 0], size)
        assert _sha(y) == g["output_sha256"], v
        idx = np.array(g["sample_flat_index"], dtype=np.int64)
        assert np.array_equal(y.reshape(-1)[idx].view(np.uint32),
                              np.array(g["sample_value_bits"], dtype=np.uint32))


def test_c5b_smallest_volume_is_oracle_adjoint(itk, orc, gold):
    from itk_rs_amd import workloads as wl
    grads, params = wl.c5_batch("cpu")
    gv = gold["c5b"]["volumes"]
    v = min(range(16), key=lambda k: grads[k].numel())
    ga = grads[v].numpy()
    assert _sha(ga) == gv[v]["grad_sha256"]
    size = list(reversed(ga.shape))
    gin, _, _ = orc.adjoint_f64(ga, [0, 0, 0], size, params[v][2], params[v][1], [0, 0, 0], size)
    idx = np.array(gv[v]["sample_flat_index"], dtype=np.int64)
    assert np.array_equal(gin.reshape(-1)[idx], np.array(gv[v]["sample_value"]))
    assert float(np.abs(gin).max()) == gv[v]["max_abs"]
