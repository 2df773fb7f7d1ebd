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
