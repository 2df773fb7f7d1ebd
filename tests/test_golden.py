"""Committed oracle results on small seeded problems (tests/golden/make_golden.py).

On the CPU the oracle must reproduce them bit for bit; on the GPU the HIP path must.
"""
import os
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    return np.load(os.path.join(HERE, "golden", "oracle_small.npz"))


def _cases():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.CASES


@pytest.mark.parametrize("name", list(_cases()))
def test_oracle_reproduces_golden(name, oracle):
    from lakcluster_c_amd import gen
    g = _golden()
    with tempfile.TemporaryDirectory() as root:
        ds = gen.make_dataset(root, name, **_cases()[name])
        r = oracle.run_text(open(ds.in_path).read())
    assert r.trace.tobytes() == g[f"{name}/trace"].tobytes()
    assert np.array_equal(r.xdrop_a, g[f"{name}/xdrop_a"])
    assert np.array_equal(r.xdrop_b, g[f"{name}/xdrop_b"])
    assert r.ops == int(g[f"{name}/ops"][0])


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(_cases()))
def test_gpu_reproduces_golden(name):
    from lakcluster_c_amd import gen, lkc
    g = _golden()
    with tempfile.TemporaryDirectory() as root:
        ds = gen.make_dataset(root, name, **_cases()[name])
        with lkc.Bcc(open(ds.in_path).read()) as run:
            run.run()
            tr = run.trace()
            xa, xb = run.xdrop()
            ops = run.ops
    assert tr.tobytes() == g[f"{name}/trace"].tobytes()
    assert np.array_equal(xa, g[f"{name}/xdrop_a"])
    assert np.array_equal(xb, g[f"{name}/xdrop_b"])
    assert ops == int(g[f"{name}/ops"][0])
