"""The seeded .b16 generator and the .in files it writes."""
import numpy as np
import pytest

from lakcluster_c_amd import gen


def test_b16_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    for shape in [(1, 1), (17, 5), (40, 33), (3, 100)]:
        m = rng.random(shape) < 0.5
        p = str(tmp_path / "x.b16")
        gen.write_b16(p, m)
        assert np.array_equal(gen.read_b16(p), m)
        hdr = np.fromfile(p, dtype=np.int32, count=3)
        assert hdr.tolist() == [16, shape[0], shape[1]]


def test_dataset_files_and_keys(tmp_path):
    ds = gen.make_dataset(str(tmp_path), "t", nA=30, ncols=50, nZ=20, T=3, seed=1, row_mask_frac=0.2)
    A_n = gen.read_b16(str(tmp_path / "dir_t_in" / "A_n.b16"))
    A_t = gen.read_b16(str(tmp_path / "dir_t_in" / "A_t.b16"))
    assert A_n.shape == (30, 50) and np.array_equal(A_t, A_n.T)
    T_n = gen.read_b16(str(tmp_path / "dir_t_in" / "T_n.b16"))
    assert T_n.shape == (30, 3) and T_n[:, 0].all()
    text = open(ds.in_path).read()
    for key in ("GLOBAL_TEST_TYPE= lakcluster_driver;", "GLOBAL_NBINS= 1;", "GLOBAL_T_n_cols= 3;",
                "GLOBAL_A_n_rind_= mr_A.b16;", "GLOBAL_Z_n_name_= Z_n.b16;", "END= 0;"):
        assert key in text


def test_planted_bicluster_is_rank_one(tmp_path):
    ds = gen.make_dataset(str(tmp_path), "p", nA=200, ncols=300, nZ=0, T=1, seed=3, flip=0.0)
    A = gen.read_b16(str(tmp_path / "dir_p_in" / "A_n.b16"))
    r, c = ds.planted_rows[0], ds.planted_cols[0]
    s = np.where(A[np.ix_(r, c)], 1, -1)
    assert np.linalg.matrix_rank(s) == 1


def _pool_uniform(seed, rows, cols):
    """The library's documented counter-based uniform (lkc.h, lkc_pool_create_random): the
    splitmix64 finaliser of seed * 0x9E3779B97F4A7C15 + row * 0xD1B54A32D192ED03 +
    (col + 1) * 0xABC98388FB8FAC03, top 53 bits as a fraction; arrays broadcast."""
    u64 = np.uint64
    with np.errstate(over="ignore"):
        x = (u64(seed) * u64(0x9E3779B97F4A7C15) + np.asarray(rows, u64) * u64(0xD1B54A32D192ED03)
             + (np.asarray(cols, u64) + u64(1)) * u64(0xABC98388FB8FAC03))
        x ^= x >> u64(30); x *= u64(0xBF58476D1CE4E5B9)
        x ^= x >> u64(27); x *= u64(0x94D049BB133111EB)
        x ^= x >> u64(31)
    return (x >> u64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def test_pool_uniform_restatement():
    """The restatement of the generator: uniform on [0, 1), rows and columns independent."""
    u = _pool_uniform(5, np.arange(400)[:, None], np.arange(300)[None, :])
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.01
    assert abs(np.corrcoef(u[:, :-1].ravel(), u[:, 1:].ravel())[0, 1]) < 0.02
    assert abs(np.corrcoef(u[:-1].ravel(), u[1:].ravel())[0, 1]) < 0.02
    assert not np.array_equal(_pool_uniform(5, 0, np.arange(10)), _pool_uniform(6, 0, np.arange(10)))


@pytest.mark.gpu
def test_device_pool_generator_matches_its_definition():
    """make_pool_random (the large problems' generator, drawn by the library on the device):
    every bit equals the documented generator's draw (U(seed, r, j) < p_j with p_j = pmin +
    (pmax - pmin) U(seed, ~0, j)) except the planted blocks, which hold their rank-1 sign
    patterns; padding bits are zero; the column frequencies follow U[pmin, pmax]; the
    covariates have an all-ones first column."""
    nA, nZ, m, seed, pmin, pmax = 300, 200, 5000, 3, 0.05, 0.5
    P = gen.make_pool_random(nA, nZ, m, T=3, pmin=pmin, pmax=pmax, planted=(0.5, 0.6), flip=0.1, seed=seed)
    with P.pool as pool:
        rows = pool.read(0, nA + nZ)
    assert rows.shape == (nA + nZ, gen.bsize(m)) and P.nA == nA and P.nrows == nA + nZ
    bits = np.unpackbits(rows, axis=1, bitorder="big")
    assert not bits[:, m:].any()
    p = pmin + (pmax - pmin) * _pool_uniform(seed, 0xFFFFFFFF, np.arange(m))
    want = (_pool_uniform(seed, np.arange(nA + nZ)[:, None], np.arange(m)[None, :]) < p[None, :])
    planted = np.zeros_like(want)
    for r, c in zip(P.planted_rows, P.planted_cols):
        planted[np.ix_(r, c)] = True
    assert np.array_equal(bits[:, :m][~planted], want[~planted])
    freq = want.mean(axis=0)
    assert abs(freq.mean() - 0.275) < 0.01 and freq.min() > 0.0 and freq.max() < 0.75
    assert [len(r) for r in P.planted_rows] == [round(nA ** 0.5), round(nA ** 0.6)]
    assert all(r.max() < nA for r in P.planted_rows)          # planted in A only
    cov = np.unpackbits(P.cov, axis=1, bitorder="big")[:, : nA + nZ]
    assert cov.shape[0] == 3 and cov[0].all() and 0.4 < cov[1].mean() < 0.6
    # a planted block is a rank-1 sign pattern u_i v_j with 10% flips: two of its rows agree
    # on about 0.9^2 + 0.1^2 = 82% of its columns or on about 18% (opposite signs)
    r, c = P.planted_rows[1], P.planted_cols[1]
    assert len(c) == round(m ** 0.6)
    sub = bits[r][:, c]
    agree = (sub[:, None, :] == sub[None, :, :]).mean(axis=2)[np.triu_indices(len(r), 1)]
    assert np.abs(agree - 0.5).mean() > 0.25


def test_split_into_bins_cuts_the_rows(tmp_path):
    """A make_dataset problem split into bins by byte slicing: the bins' matrices stack back
    into the original, patient-major and SNP-major copies agree, Y/W have the asked columns."""
    import os
    from lakcluster_c_amd import gen
    ds = gen.make_dataset(str(tmp_path), "src", nA=64, ncols=300, nZ=32, T=3, gamma=0.1, seed=3)
    b = gen.split_into_bins(ds, str(tmp_path), "split", 2, ny=20, seed=4)
    d0, d1 = os.path.dirname(ds.in_path), os.path.dirname(b.in_path)
    for key, rows in (("A", 64), ("Z", 32), ("T", 64), ("S", 32)):
        whole = gen.read_b16(f"{d0}/{key}_n.b16")
        parts = [gen.read_b16(f"{d1}/{key}_n_{i}.b16") for i in range(2)]
        assert np.array_equal(np.vstack(parts), whole)
        for i in range(2):
            assert np.array_equal(gen.read_b16(f"{d1}/{key}_t_{i}.b16"), parts[i].T)
    assert gen.read_b16(f"{d1}/Y_n_0.b16").shape == (32, 20) and gen.read_b16(f"{d1}/W_n_1.b16").shape == (16, 20)
    text = open(b.in_path).read()
    assert "GLOBAL_NBINS= 2;" in text and "GLOBAL_Y_n_cols= 20;" in text

