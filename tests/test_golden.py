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


# ---------------------------------------------------------------------------
# The 4096 x 16384 case-only problem at full size (tests/golden/make_caseonly_golden.py)
# ---------------------------------------------------------------------------

def _caseonly_4096():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_caseonly_golden", os.path.join(HERE, "golden", "make_caseonly_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.CASE_ONLY_4096, np.load(os.path.join(HERE, "golden", "oracle_caseonly_4096x16384.npz"))


def test_caseonly_4096_fixture_follows_reference_run(oracle):
    """The oracle's full-size run agrees with the reference's own run of this problem
    (recorded in xdrop_schedules.json: 115 trace lines, last 114 2 2, 20480 xdrop
    lines), every count follows get_xdrop, and every row and column leaves once."""
    import json
    kw, g = _caseonly_4096()
    ref = json.load(open(os.path.join(HERE, "golden", "xdrop_schedules.json")))["c1_4096x16384"]
    tr = g["trace"]
    assert tr.shape[0] == ref["n_trace"]
    assert [int(x) for x in tr[-1, :3]] == ref["last"]
    assert g["xdrop_a"].shape[0] == ref["xdrop_lines"]
    assert np.isfinite(tr).all()
    for i in range(tr.shape[0] - 1):
        rd, cd = oracle.get_xdrop(tr[i, 1], tr[i, 2], kw["gamma"])
        assert tr[i + 1, 1] == tr[i, 1] - min(rd, tr[i, 1]) and tr[i + 1, 2] == tr[i, 2] - min(cd, tr[i, 2])
    xa = g["xdrop_a"].astype(np.int64)
    rows, cols = xa[xa[:, 0] >= 0, 0], xa[xa[:, 1] >= 0, 1]
    assert np.array_equal(np.sort(rows), np.arange(kw["nA"]))
    assert np.array_equal(np.sort(cols), np.arange(kw["ncols"]))


@pytest.mark.gpu
def test_gpu_caseonly_4096_full_size_matches_oracle():
    """The HIP path on the 4096 x 16384 case-only problem: trace (bit patterns), xdrop
    lists and bit-op count equal the oracle's committed full-size run."""
    from lakcluster_c_amd import gen, lkc
    kw, g = _caseonly_4096()
    with tempfile.TemporaryDirectory() as root:
        ds = gen.make_dataset(root, "caseonly_4096x16384", **kw)
        with lkc.Bcc(open(ds.in_path).read()) as run:
            run.run()
            tr = run.trace()
            xa, xb = run.xdrop()
            ops = run.ops
    assert tr.tobytes() == g["trace"].tobytes()
    assert np.array_equal(xa, g["xdrop_a"].astype(xa.dtype))
    assert np.array_equal(xb, g["xdrop_b"].astype(xb.dtype))
    assert ops == int(g["ops"][0])


# ---------------------------------------------------------------------------
# The 16384 x 65536 case-control problem's family at 2048 x 8192
# (tests/golden/make_casecontrol_golden.py): the benchmark's own parameters (T/S 3
# covariates, gamma 0.02, B_MLT 32, planted bicluster in A) at a size the oracle finishes
# ---------------------------------------------------------------------------

def _casecontrol_2048():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_casecontrol_golden",
                                                  os.path.join(HERE, "golden", "make_casecontrol_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.CASE_CONTROL_2048, np.load(os.path.join(HERE, "golden", "oracle_casecontrol_2048x8192.npz"))


def test_casecontrol_2048_fixture_invariants(oracle):
    """The committed oracle run: counts follow get_xdrop at every line down to 2 x 2, and
    every row and column leaves once."""
    kw, g = _casecontrol_2048()
    tr = g["trace"]
    assert np.isfinite(tr).all() and [int(x) for x in tr[0, :3]] == [0, kw["nA"], kw["ncols"]] and tr[-1, 1] <= 2
    for i in range(tr.shape[0] - 1):
        rd, cd = oracle.get_xdrop(tr[i, 1], tr[i, 2], kw["gamma"])
        assert tr[i + 1, 1] == tr[i, 1] - min(rd, tr[i, 1]) and tr[i + 1, 2] == tr[i, 2] - min(cd, tr[i, 2])
    xa = g["xdrop_a"].astype(np.int64)
    rows, cols = xa[xa[:, 0] >= 0, 0], xa[xa[:, 1] >= 0, 1]
    assert np.array_equal(np.sort(rows), np.arange(kw["nA"]))
    assert np.array_equal(np.sort(cols), np.arange(kw["ncols"]))


@pytest.mark.gpu
def test_gpu_casecontrol_2048x8192_matches_oracle():
    """The HIP path on the 2048 x 8192 case-control problem: trace (bit patterns), xdrop
    lists and bit-op count equal the oracle's committed run."""
    from lakcluster_c_amd import gen, lkc
    kw, g = _casecontrol_2048()
    with tempfile.TemporaryDirectory() as root:
        ds = gen.make_dataset(root, "casecontrol_2048x8192", **kw)
        with lkc.Bcc(open(ds.in_path).read()) as run:
            run.run()
            tr = run.trace()
            xa, xb = run.xdrop()
            ops = run.ops
    assert tr.tobytes() == g["trace"].tobytes()
    assert np.array_equal(xa, g["xdrop_a"].astype(xa.dtype))
    assert np.array_equal(xb, g["xdrop_b"].astype(xb.dtype))
    assert ops == int(g["ops"][0])
