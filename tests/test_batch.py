"""Many eliminations on one GPU: jobs drawn from a device-resident pool, run alone, in a
concurrent batch and as a served stream.  A job's trace and drop order must equal the
oracle's on the same problem written as files, and be the same alone and in a mix."""
import numpy as np
import pytest

from lakcluster_c_amd import gen


def _pool(seed=11):
    return gen.make_pool(140, 120, 2600, T=3, seed=seed)


def test_null_jobs_shape():
    pool = _pool()
    jobs = gen.null_jobs(pool, 12, seed=3, rows=(60, 200), cols=(1000, 2600))
    for j in jobs:
        n = j.rowsA.size + j.rowsZ.size
        assert 59 <= n <= 201
        assert np.all(np.diff(j.rowsA) > 0) and np.all(np.diff(j.rowsZ) > 0)
        assert not set(j.rowsA.tolist()) & set(j.rowsZ.tolist())
        cm = np.unpackbits(j.colmask, bitorder="big")[: pool.ncols]
        assert int(cm.sum()) == j.ncols_active and 999 <= j.ncols_active <= 2601
        assert j.colmask.size == gen.bsize(pool.ncols)


def test_written_job_runs_on_the_oracle(tmp_path, oracle):
    pool = _pool()
    job = gen.null_jobs(pool, 1, seed=4, rows=(60, 120), cols=(1200, 2000))[0]
    ds = gen.write_job(str(tmp_path), "j0", pool, job, gamma=0.05, B=32)
    r = oracle.run_text(open(ds.in_path).read())
    assert r.n_iter > 3
    assert r.trace[0, 1] == job.rowsA.size and r.trace[0, 2] == job.ncols_active


def _solo(P, j, gamma, B):
    with P.job(j.rowsA, j.rowsZ, j.colmask, gamma, B) as run:
        run.run()
        tr = run.trace()
        xa, xb = run.xdrop()
    return tr, xa, xb


@pytest.mark.gpu
def test_job_from_pool_matches_oracle(tmp_path, oracle):
    from lakcluster_c_amd import lkc
    pool = _pool()
    jobs = gen.null_jobs(pool, 3, seed=5, rows=(60, 260), cols=(1000, 2600))
    with lkc.Pool(pool.rows, pool.nrows, pool.ncols, pool.cov, pool.ncov) as P:
        for i, j in enumerate(jobs):
            print(f"job {i}: A {j.rowsA.size} Z {j.rowsZ.size} cols {j.ncols_active}", flush=True)
            ds = gen.write_job(str(tmp_path), f"j{i}", pool, j, gamma=0.05, B=32)
            ref = oracle.run_text(open(ds.in_path).read())
            tr, xa, xb = _solo(P, j, 0.05, 32)
            assert tr.tobytes() == ref.trace.tobytes()
            np.testing.assert_array_equal(xa, ref.xdrop_a)
            np.testing.assert_array_equal(xb, ref.xdrop_b)


@pytest.mark.gpu
def test_batch_and_serving_match_solo_runs():
    from lakcluster_c_amd import lkc
    pool = _pool(seed=12)
    jobs = gen.null_jobs(pool, 8, seed=6, rows=(60, 260), cols=(1000, 2600))
    spec = [(j.rowsA, j.rowsZ, j.colmask, 0.05, 32) for j in jobs]
    with lkc.Pool(pool.rows, pool.nrows, pool.ncols, pool.cov, pool.ncov) as P:
        solo = [_solo(P, j, 0.05, 32) for j in jobs]
        print("solo runs done", flush=True)
        batched = P.batch(spec, max_concurrent=4)
        print("batch done", flush=True)
        # a budget that admits one job at a time, and a served stream with arrivals
        one_at_a_time = P.batch(spec, max_concurrent=4, budget_bytes=1.0)
        served = P.batch(spec, max_concurrent=3, arrivals=[0.05 * k for k in range(len(jobs))])
    for out in (batched, one_at_a_time, served):
        for (tr, xa, xb), r in zip(solo, out):
            assert r["status"] == 0
            assert r["trace"].tobytes() == tr.tobytes()
            np.testing.assert_array_equal(r["xdrop_a"], xa)
            np.testing.assert_array_equal(r["xdrop_b"], xb)
    for r in served:
        assert r["end_s"] >= r["start_s"] >= r["arrival_s"] - 1e-9
    # one at a time: no two jobs overlap
    iv = sorted((r["start_s"], r["end_s"]) for r in one_at_a_time)
    assert all(b[0] >= a[1] - 1e-9 for a, b in zip(iv, iv[1:]))


def _mixed_jobs(pool, n, seed):
    """Jobs as a served stream draws them: half case-only (every picked row a case, no
    controls), half case-control, gamma ~ U[0.01, 0.1] per job, one with gamma_type 1 and one
    with kappa_squared 0.5."""
    rng = np.random.default_rng(seed)
    jobs = gen.null_jobs(pool, n, seed=seed, rows=(60, 260), cols=(1000, 2600))
    out = []
    for i, j in enumerate(jobs):
        if i % 2:
            j = gen.NullJob(np.sort(np.r_[j.rowsA, j.rowsZ]).astype(np.int32), np.zeros(0, np.int32), j.colmask,
                            j.ncols_active)
        kw = {"gamma_type": 1 if i == 2 else 0, "kappa_squared": 0.5 if i == 3 else 0.0}
        out.append((j, float(rng.uniform(0.01, 0.1)), kw))
    return out


@pytest.mark.gpu
def test_batch_and_served_mixed_jobs_match_oracle(tmp_path, oracle):
    """Every job of a concurrent batch and of a served stream (Poisson-like arrivals, a
    budget that holds two jobs) against its own oracle run of the job written as files:
    case-only and case-control jobs, per-job gamma, gamma_type and kappa_squared."""
    from lakcluster_c_amd import lkc
    pool = _pool(seed=13)
    mixed = _mixed_jobs(pool, 8, seed=21)
    refs = []
    for i, (j, g, kw) in enumerate(mixed):
        ds = gen.write_job(str(tmp_path), f"m{i}", pool, j, gamma=g, B=32, **kw)
        refs.append(oracle.run_text(open(ds.in_path).read()))
        assert refs[-1].trace[0, 1] == j.rowsA.size and refs[-1].trace[0, 2] == j.ncols_active
    spec = [(j.rowsA, j.rowsZ, j.colmask, g, 32, kw) for j, g, kw in mixed]
    arr = np.cumsum(np.random.default_rng(5).exponential(0.03, size=len(spec)))
    with lkc.Pool(pool.rows, pool.nrows, pool.ncols, pool.cov, pool.ncov) as P:
        print("mixed batch", flush=True)
        batched = P.batch(spec, max_concurrent=4)
        two = 2.05 * max(lkc.job_bytes(j.rowsA.size, j.rowsZ.size, pool.ncols, 32) for j, _, _ in mixed)
        print("mixed served stream", flush=True)
        served = P.batch(spec, max_concurrent=4, budget_bytes=two, arrivals=arr - arr[0])
    for out in (batched, served):
        for i, (r, ref) in enumerate(zip(out, refs)):
            assert r["status"] == 0, i
            assert r["trace"].tobytes() == ref.trace.tobytes(), i
            np.testing.assert_array_equal(r["xdrop_a"], ref.xdrop_a)
            np.testing.assert_array_equal(r["xdrop_b"], ref.xdrop_b)
    for r in served:
        assert r["end_s"] >= r["start_s"] >= r["arrival_s"] - 1e-9
