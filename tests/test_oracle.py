"""Pins of the CPU oracle (the checker of the HIP path), on the CPU.

Each check is chosen so that a plausible mistake in the restatement fails it:
reference-run known answers for the drop schedule, a hand-worked tie order for
the quicksort, exact brute force (dyadic weights, so every summation order is
exact) for An_ajdk_v / AnZt_vv / the bit-plane contraction of An_ZtSn_ww, the
quantiser's defining properties, and the low-rank update against recomputing
every loop-subscore from scratch.
"""
import json
import os

import numpy as np
import pytest

import loopcount

HERE = os.path.dirname(os.path.abspath(__file__))


def _schedule(oracle, r, c, gamma, gamma_type=0):
    out = []
    while r > 0 or c > 0:
        out.append((r, c))
        rd, cd = oracle.get_xdrop(r, c, gamma, gamma_type)
        r, c = r - min(rd, r), c - min(cd, c)
    return out


def test_xdrop_schedule_reference_runs(oracle):
    g = json.load(open(os.path.join(HERE, "golden", "xdrop_schedules.json")))
    d = g["dir_test"]
    s = _schedule(oracle, 77, 300, d["gamma"])
    assert s[:3] == [tuple(x) for x in d["rows_cols"]]
    assert (77 - s[3][0]) + (300 - s[3][1]) == d["xdrop_lines_3_steps"]
    c1 = g["c1_4096x16384"]
    s = _schedule(oracle, *c1["start"], c1["gamma"])
    assert len(s) == c1["n_trace"]
    assert (len(s) - 1, *s[-1]) == tuple(c1["last"])
    assert sum(c1["start"]) == c1["xdrop_lines"]
    a = g["aaaa_1024x2049"]
    assert len(_schedule(oracle, *a["start"], a["gamma"])) == a["n_trace"]


def test_xdrop_edges(oracle):
    assert oracle.get_xdrop(2, 50, 0.1) == (2, 50)          # <= 2 rows: drop everything
    assert oracle.get_xdrop(50, 2, 0.1) == (50, 2)          # rows > cols: roles swap
    r, c = oracle.get_xdrop(100, 1000, 0.05, 3)
    assert c == 0 and r == 5                                 # gamma_type 3 keeps the columns
    r, c = oracle.get_xdrop(1000, 1000, 0.0)
    assert r == 1                                            # gamma 0: one row per step


def test_calc_ajdk_tables(oracle):
    p = np.array([0.5, 0.001, 0.2, 0.9, 0.37])
    t = oracle.calc_ajdk(p).reshape(15, p.size)            # t[j + 5k][b]
    assert np.all(t[0] == 1.0)
    assert t[5][0] == 1.0 and t[1][0] == 0.0                # p = 1/2: alpha 0, D 1
    assert t[5][1] == 100.0                                  # 4p(1-p) < 0.01 is clamped
    alpha, D = t[1], t[5]
    for j in range(5):
        for k in range(3):
            np.testing.assert_allclose(t[j + 5 * k], alpha ** j * D ** k, rtol=4e-16, atol=1e-300)


def test_quicksort_hand_worked_tie_order(oracle):
    # Hoare partition around the first element, worked by hand for [3, 1, 3, 2, 3]:
    # pass 1 swaps positions 0,4; pass 2 swaps 0,3; pass 3 swaps 0,1.
    d, ix = oracle.quicksort(np.array([3.0, 1.0, 3.0, 2.0, 3.0]))
    assert d.tolist() == [1.0, 2.0, 3.0, 3.0, 3.0]
    assert ix.tolist() == [1, 3, 2, 4, 0]


def test_quicksort_sorts_with_ties(oracle):
    rng = np.random.default_rng(0)
    v = rng.integers(0, 20, 500).astype(float)
    d, ix = oracle.quicksort(v)
    assert np.all(np.diff(d) >= 0)
    assert sorted(ix.tolist()) == list(range(500))
    assert np.array_equal(v[ix], d)


def _bits(rng, nrows, ncols, p=0.45):
    from lakcluster_c_amd.gen import BITJ  # noqa: F401
    rb = ((((ncols + 15) // 16 * 16) + 1919) // 1920) * 240
    b = rng.random((nrows, ncols)) < p
    X = np.zeros((nrows, rb), np.uint8)
    X[:, : (ncols + 7) // 8] = np.packbits(b, axis=1, bitorder="big")
    return X, b


def _dyadic_tables(rng, pcols):
    return (rng.integers(-64, 64, size=15 * pcols) / 8.0).astype(np.float64)


@pytest.mark.parametrize("ncols", [37, 1920, 4100])
def test_an_ajdk_v_brute_force(ncols, oracle):
    rng = np.random.default_rng(ncols)
    X, b = _bits(rng, 20, ncols)
    pc = X.shape[1] // 240
    m = rng.random(ncols) < 0.8
    cm = np.zeros(X.shape[1], np.uint8)
    cm[: (ncols + 7) // 8] = np.packbits(m, bitorder="big")
    ajdk = _dyadic_tables(rng, pc)
    rows = np.array([0, 3, 7, 19], np.int32)
    got = oracle.an_ajdk_v(X, rows, cm, ncols, ajdk)
    w = ajdk.reshape(15, pc)[:, np.arange(ncols) // 1920]   # weight of every column
    s = np.where(b[rows], 1.0, -1.0) * m
    want = s @ w.T
    assert np.array_equal(got, want)


@pytest.mark.parametrize("centred", [False, True])
def test_anzt_vv_brute_force(centred, oracle):
    rng = np.random.default_rng(11)
    ncols = 3900
    A, ab = _bits(rng, 9, ncols)
    Z, zb = _bits(rng, 7, ncols)
    pc = A.shape[1] // 240
    m = rng.random(ncols) < 0.7
    cm = np.zeros(A.shape[1], np.uint8)
    cm[: (ncols + 7) // 8] = np.packbits(m, bitorder="big")
    ar, zr = np.array([0, 2, 5, 8], np.int32), np.array([1, 6], np.int32)
    sa, sz = np.where(ab[ar], 1.0, -1.0), np.where(zb[zr], 1.0, -1.0)
    if not centred:
        got = oracle.anzt_vv(A, ar, Z, zr, cm, ncols)
        want = (sa * m) @ sz.T
    else:
        ajdk = np.abs(_dyadic_tables(rng, pc)) + 0.125
        xa, xz = rng.integers(-50, 50, ar.size) / 4.0, rng.integers(-50, 50, zr.size) / 4.0
        got = oracle.anzt_vv(A, ar, Z, zr, cm, ncols, ajdk, xa, xz)
        t = ajdk.reshape(15, pc)
        D = t[5][np.arange(ncols) // 1920]
        a2d1 = float(np.sum(t[7][np.arange(ncols) // 1920] * m))
        want = (sa * (D * m)) @ sz.T - xa[:, None] - xz[None, :] + a2d1
    assert np.array_equal(got, want)


@pytest.mark.parametrize("B", [8, 16, 32])
def test_xcalc_properties(B, oracle):
    rng = np.random.default_rng(B)
    nlist = np.sort(rng.permutation(300)[:200]).astype(np.int32)
    L = rng.normal(size=(200, 5))
    cz, mn, mlt, planes = oracle.xcalc(L, nlist, 300, B)
    lo, hi = L.min(), L.max()
    assert mn == lo - 0.0625 * (hi - lo)
    assert cz.min() >= 1 and cz.max() <= 2 ** B - 1
    v, c = L.T.ravel(), cz.ravel().astype(float)
    o = np.argsort(v, kind="stable")
    assert np.all(np.diff(c[o]) >= 0)
    assert np.all(np.abs(mn + c / mlt - v) <= (1 + 1e-9) / mlt)
    # plane (j, b) holds bit b of cz(j, n) at n, nothing outside nlist
    bits = np.unpackbits(planes, axis=2, bitorder="big")[:, :, :300]
    rebuilt = np.zeros((5, 300), np.uint64)
    for bb in range(B):
        rebuilt += bits[:, bb, :].astype(np.uint64) << np.uint64(bb)
    full = np.zeros((5, 300), np.uint64)
    full[:, nlist] = cz
    assert np.array_equal(rebuilt, full)


def test_an_ztsn_ww_plane_contraction(oracle):
    """sum_b 2^b pm0(A, plane_b) over the mask equals sum_n A(n) cz(n) exactly."""
    rng = np.random.default_rng(5)
    n_total, ns, B = 2500, 3, 32
    A, ab = _bits(rng, 12, n_total)
    nlist = np.sort(rng.permutation(n_total)[:1800]).astype(np.int32)
    amask = np.zeros(A.shape[1], np.uint8)
    mb = np.zeros(n_total, bool); mb[nlist] = True
    amask[: (n_total + 7) // 8] = np.packbits(mb, bitorder="big")
    L = rng.normal(size=(nlist.size, ns)) * 10
    alist = np.arange(12, dtype=np.int32)
    got = oracle.an_ztsn_ww(A, alist, amask, L, nlist, n_total, B)
    cz, mn, mlt, _ = oracle.xcalc(L, nlist, n_total, B)
    s = np.where(ab[:, nlist], 1, -1).astype(np.int64)
    ll = s.sum(axis=1).astype(np.float64)
    G = s @ cz.T.astype(np.int64)                        # exact integers
    want = (ll[:, None] * mn * mlt + G.astype(np.float64)) / mlt
    assert np.array_equal(got, want)


def _run_pair(oracle, tmp_path, **kw):
    from lakcluster_c_amd import gen
    ds = gen.make_dataset(str(tmp_path), "lr", **kw)
    text = open(ds.in_path).read()
    return oracle.run_text(text), oracle.run_text(text, recompute=True)


@pytest.mark.parametrize("kw", [
    dict(nA=80, ncols=500, nZ=60, T=3, seed=3, gamma=0.05),
    dict(nA=120, ncols=700, nZ=0, T=1, seed=4, gamma=0.05),
    dict(nA=100, ncols=2500, nZ=90, T=2, seed=5, gamma=0.05, row_mask_frac=0.1, col_mask_frac=0.1),
])
def test_low_rank_update_matches_recompute(kw, oracle, tmp_path):
    """The low-rank update (bcc_lrup) reproduces recomputing every loop-subscore from scratch."""
    a, b = _run_pair(oracle, tmp_path, **kw)
    n = int(0.6 * min(len(a.trace), len(b.trace)))
    np.testing.assert_array_equal(a.trace[:n, :3], b.trace[:n, :3])
    np.testing.assert_allclose(a.trace[:n, 3:5], b.trace[:n, 3:5], rtol=1e-6, atol=0)
    # the same rows and columns leave at every step (near-equal scores may swap places within a step)
    tr = a.trace
    cum = np.cumsum(np.r_[0, (tr[:-1, 1] - tr[1:, 1]) + (tr[:-1, 2] - tr[1:, 2])]).astype(int)
    for st in range(n):
        seg_a = {tuple(x) for x in a.xdrop_a[cum[st]:cum[st + 1]].tolist()}
        seg_b = {tuple(x) for x in b.xdrop_a[cum[st]:cum[st + 1]].tolist()}
        assert seg_a == seg_b, st


def test_run_invariants(oracle, datasets):
    """Every row and column is dropped exactly once; counts follow get_xdrop."""
    for name in ("caseonly", "masked"):
        ds = datasets[name]
        r = oracle.run_text(open(ds.in_path).read())
        rows = r.xdrop_a[r.xdrop_a[:, 0] >= 0, 0]
        cols = r.xdrop_a[r.xdrop_a[:, 1] >= 0, 1]
        assert len(set(rows.tolist())) == len(rows) == int(r.trace[0, 1])
        assert len(set(cols.tolist())) == len(cols) == int(r.trace[0, 2])
        for s in range(len(r.trace) - 1):
            rd, cd = oracle.get_xdrop(r.trace[s, 1], r.trace[s, 2], ds.gamma)
            assert r.trace[s + 1, 1] == r.trace[s, 1] - min(rd, r.trace[s, 1])
            assert r.trace[s + 1, 2] == r.trace[s, 2] - min(cd, r.trace[s, 2])
        assert np.all(np.isfinite(r.trace[:, 3:5]))


@pytest.mark.parametrize("kw,steps", [
    (dict(nA=12, ncols=20, nZ=10, T=3, seed=5, gamma=0.1, planted=()), 0),
    (dict(nA=14, ncols=30, nZ=0, T=1, seed=6, gamma=0.1), 0),
    (dict(nA=16, ncols=25, nZ=11, T=2, seed=7, gamma=0.1, row_mask_frac=0.2, col_mask_frac=0.2), 0),
    (dict(nA=16, ncols=40, nZ=12, T=3, seed=8, gamma=0.1), 3),
])
def test_scores_match_loop_count_definition(kw, steps, oracle, tmp_path):
    """The flattened-loop scores equal the reference's brute-force loop counts, at the
    initial pass and after elimination steps (so through the low-rank update)."""
    from lakcluster_c_amd import gen
    ds = gen.make_dataset(str(tmp_path), "bf", **kw)
    d = str(tmp_path / "dir_bf_in")
    S = oracle.State(open(ds.in_path).read())
    for _ in range(steps):
        S.iterate()
    A = gen.read_b16(d + "/A_n.b16")
    Z = gen.read_b16(d + "/Z_n.b16") if kw["nZ"] else None
    Tm = gen.read_b16(d + "/T_n.b16") if kw["T"] > 1 else np.ones((A.shape[0], 1), bool)
    Sm = gen.read_b16(d + "/S_n.b16") if (kw["T"] > 1 and Z is not None) else (np.ones((Z.shape[0], 1), bool) if Z is not None else None)
    ar = np.flatnonzero(S.get("mask_Ar_j")); kl = np.flatnonzero(S.get("mask_Ac_j"))
    zr = np.flatnonzero(S.get("mask_Zr_j")) if Z is not None else None
    tab = S.get("ajdk").reshape(15, -1)
    nt = kw["T"]
    bf = loopcount.loop_counts(A, Z, Tm, Sm, tab, ar, zr, kl, nt)
    for name, want in bf.items():
        rows = kl if name.startswith("QC") else ar
        got = S.get(name).reshape(-1, nt)[rows]
        scale = np.abs(want).max()
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-8 * scale, err_msg=name)
    S.close()


@pytest.mark.parametrize("kw", [
    dict(nA=14, ncols=36, nZ=11, T=3, seed=21, gamma=0.1),
    dict(nA=18, ncols=40, nZ=0, T=1, seed=22, gamma=0.1),
    dict(nA=16, ncols=30, nZ=12, T=2, seed=23, gamma=0.12, row_mask_frac=0.15, col_mask_frac=0.15),
    dict(nA=15, ncols=34, nZ=13, T=3, seed=24, gamma=0.1, kappa_squared=0.5),
    dict(nA=16, ncols=30, nZ=0, T=4, seed=25, gamma=0.1, kappa_squared=0.25),
    dict(nA=14, ncols=3900, nZ=10, T=3, seed=26, gamma=0.2, B=24),
])
def test_trace_and_drops_match_loop_count_scores(kw, oracle, tmp_path):
    """The whole elimination against the loop-count definition: at every step the trace's
    mean row/column scores equal the means of the normalised (bcc_sumscores_nrm), covariate-
    corrected (_ifT) and combined (_cmb) brute-force loop counts, and the rows and columns
    dropped are the lowest-scoring ones.  The kappa_squared = 0 cases with several covariates
    pin the reference's no-op ifT (its scale is kappa_squared itself, not the table of
    GLOBAL_kappa_squared_loop_scale_factor_); the kappa_squared > 0 cases pin the pow (QR) and
    linear (QC) corrections."""
    from lakcluster_c_amd import gen
    ds = gen.make_dataset(str(tmp_path), "ns", **kw)
    text = open(ds.in_path).read()
    n, compared = loopcount.check_run(oracle, text, str(tmp_path / "dir_ns_in"), kw["nZ"], kw["T"],
                                      kw.get("kappa_squared", 0.0), rtol=1e-6 if kw.get("B", 32) == 32 else 1e-4)
    assert n >= 5 and compared >= n // 2


def test_duplicate_columns_tie_exactly(oracle, tmp_path):
    """Duplicated SNP columns (copies of their left neighbour in the same 1920-column block)
    get bit-identical column subscores, so their scores tie exactly and the drop order among
    them is the reference quicksort's tie order; the run still drops every row and column
    once with the get_xdrop counts."""
    from lakcluster_c_amd import gen
    ds = gen.make_dataset(str(tmp_path), "dup", nA=60, ncols=300, nZ=40, T=3, gamma=0.05, seed=41, dup_cols=0.5)
    text = open(ds.in_path).read()
    A = gen.read_b16(str(tmp_path / "dir_dup_in" / "A_n.b16"))
    dup = [j for j in range(1, A.shape[1]) if np.array_equal(A[:, j], A[:, j - 1])]
    assert len(dup) > 100
    S = oracle.State(text)
    for name in ("QC_TAT", "QC_ZSZ"):
        q = S.get(name).reshape(-1, 3)
        for j in dup:
            assert q[j].tobytes() == q[j - 1].tobytes(), (name, j)
    S.close()
    r = oracle.run_text(text)
    cols = r.xdrop_a[r.xdrop_a[:, 1] >= 0, 1]
    assert sorted(cols.tolist()) == list(range(300))
