"""The seeded .b16 generator and the .in files it writes."""
import numpy as np

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


def test_device_pool_generator_on_cpu():
    """make_pool_device (the large problems' generator) run on the CPU: the packed rows
    hold exactly the drawn bits with zero padding, the column frequencies follow
    U[pmin, pmax], the planted blocks are rank-1 sign patterns up to the flips, and the
    covariates have an all-ones first column."""
    import numpy as np
    from lakcluster_c_amd import gen
    nA, nZ, m = 300, 200, 5000
    P = gen.make_pool_device(nA, nZ, m, T=3, pmin=0.05, pmax=0.5, planted=(0.5, 0.6), flip=0.1, seed=3,
                             chunk=128, device="cpu")
    assert P.rows.shape == (nA + nZ, gen.bsize(m)) and P.nA == nA and P.nrows == nA + nZ
    bits = np.unpackbits(P.rows, axis=1, bitorder="big")
    assert not bits[:, m:].any()
    freq = bits[:, :m].mean(axis=0)
    assert abs(freq.mean() - 0.275) < 0.01 and freq.min() > 0.0 and freq.max() < 0.75
    assert [len(r) for r in P.planted_rows] == [round(nA ** 0.5), round(nA ** 0.6)]
    assert all(r.max() < nA for r in P.planted_rows)          # planted in A only
    cov = np.unpackbits(P.cov, axis=1, bitorder="big")[:, : nA + nZ]
    assert cov.shape[0] == 3 and cov[0].all() and 0.4 < cov[1].mean() < 0.6
    # a planted block is a rank-1 sign pattern u_i v_j with 10% flips: two of its rows agree
    # on about 0.9^2 + 0.1^2 = 82% of its columns or on about 18% (opposite signs); two
    # unplanted rows on the same columns agree on p^2 + (1-p)^2 of them (0.5 to 0.9)
    r, c = P.planted_rows[1], P.planted_cols[1]
    assert len(c) == round(m ** 0.6)
    sub = bits[r][:, c]
    agree = (sub[:, None, :] == sub[None, :, :]).mean(axis=2)[np.triu_indices(len(r), 1)]
    assert np.abs(agree - 0.5).mean() > 0.25
    others = np.setdiff1d(np.arange(nA), np.concatenate(P.planted_rows))[:len(r)]
    sub0 = bits[others][:, c]
    agree0 = (sub0[:, None, :] == sub0[None, :, :]).mean(axis=2)[np.triu_indices(len(r), 1)]
    assert np.abs(agree0 - 0.5).mean() < 0.15
