"""Bit-exact parity of the gfx950 path with the CPU oracle, through the C ABI.

Every elimination is compared on the whole trace (iteration, rows, cols, mean
row score, mean column score, Irem) and on the full xdrop lists: the order in
which rows and columns are removed must be identical, and the printed means
equal to the last bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = ["caseonly", "casectrl", "masked", "blocks", "b16_g1", "kappa", "tiny", "wide_z"]


def _read(path):
    with open(path) as f:
        return f.read()


@pytest.mark.parametrize("kern", ["mfma", "valu"])
@pytest.mark.parametrize("case", CASES)
def test_elimination_matches_oracle(case, kern, datasets, oracle):
    from lakcluster_c_amd import lkc
    ds = datasets[case]
    text = _read(ds.in_path)
    print(f"case {case} kernels {kern}: A {ds.nA}x{ds.ncols} Z {ds.nZ} T {ds.T} B {ds.B}", flush=True)
    ref = oracle.run_text(text)
    with lkc.Bcc(text) as run:
        run.set_kernels(kern)
        run.run()
        tr = run.trace()
        xa, xb = run.xdrop()
        ops = run.ops
        n_iter = run.n_iter
    assert n_iter == ref.n_iter
    assert tr.shape == ref.trace.shape
    np.testing.assert_array_equal(tr[:, [0, 1, 2, 5]], ref.trace[:, [0, 1, 2, 5]])
    # means: identical bit patterns
    assert tr[:, 3:5].tobytes() == ref.trace[:, 3:5].tobytes(), \
        np.argwhere(tr[:, 3:5] != ref.trace[:, 3:5])[:5]
    np.testing.assert_array_equal(xa, ref.xdrop_a)
    np.testing.assert_array_equal(xb, ref.xdrop_b)
    assert ops == ref.ops
    # every row and column is eventually dropped exactly once
    rows = xa[xa[:, 0] >= 0, 0]
    cols = xa[xa[:, 1] >= 0, 1]
    assert len(set(rows.tolist())) == len(rows)
    assert len(set(cols.tolist())) == len(cols)


@pytest.mark.parametrize("kw", [
    dict(nA=120, ncols=700, nZ=90, T=3, gamma=0.05, seed=41, dup_cols=0.5),
    dict(nA=150, ncols=600, nZ=0, T=1, gamma=0.05, seed=42, dup_cols=0.6),
])
def test_tied_scores_take_the_reference_quicksort(kw, oracle, tmp_path):
    """Duplicated SNP columns tie exactly; where ties reach the dropped prefix the engine
    sorts with the reference's quicksort on the host, and the run stays bit-exact."""
    from lakcluster_c_amd import gen, lkc
    ds = gen.make_dataset(str(tmp_path), "dup", **kw)
    text = _read(ds.in_path)
    print(f"tied scores: {kw}", flush=True)
    ref = oracle.run_text(text)
    with lkc.Bcc(text) as run:
        run.run()
        tr = run.trace()
        xa, xb = run.xdrop()
        fb = run.sort_fallbacks
    assert fb > 0
    assert tr.tobytes() == ref.trace.tobytes()
    np.testing.assert_array_equal(xa, ref.xdrop_a)
    np.testing.assert_array_equal(xb, ref.xdrop_b)


def test_stepwise_api_matches_whole_run(datasets, oracle):
    """lkc_bcc_init_lf + lkc_bcc_iterate reproduce lkc_bcc_run."""
    from lakcluster_c_amd import lkc
    text = _read(datasets["casectrl"].in_path)
    ref = oracle.run_text(text)
    with lkc.Bcc(text) as run:
        run.init_lf()
        steps = 0
        while not run.iterate():
            steps += 1
        tr = run.trace()
    assert tr.tobytes() == ref.trace.tobytes()


def test_max_iter_prefix(datasets, oracle):
    from lakcluster_c_amd import lkc
    text = _read(datasets["masked"].in_path)
    ref = oracle.run_text(text, max_iter=5)
    with lkc.Bcc(text) as run:
        run.run(5)
        tr = run.trace()
    assert tr.tobytes() == ref.trace.tobytes()


def test_output_files_match(datasets, oracle, tmp_path):
    """The three output files are byte-identical to the oracle's."""
    from lakcluster_c_amd import lkc
    ds = datasets["casectrl"]
    text = _read(ds.in_path)
    d_ref = tmp_path / "ref"
    d_gpu = tmp_path / "gpu"
    d_ref.mkdir()
    oracle.run_text(text, out_dir=str(d_ref), write_files=True)
    lkc.lakcluster_driver(text, str(d_gpu))
    for name in ("out_trace.txt", "out_xdrop_a.txt", "out_xdrop_b.txt"):
        assert (d_gpu / name).read_bytes() == (d_ref / name).read_bytes(), name


def _packed(rng, nrows, ncols):
    from lakcluster_c_amd import lkc
    rb = lkc.bsize(ncols)
    bits = rng.random((nrows, ncols)) < 0.4
    X = np.zeros((nrows, rb), dtype=np.uint8)
    X[:, : (ncols + 7) // 8] = np.packbits(bits, axis=1, bitorder="big")
    return X


def _mask(rng, n, frac=0.8):
    from lakcluster_c_amd import lkc
    m = rng.random(n) < frac
    out = np.zeros(lkc.bsize(n), dtype=np.uint8)
    out[: (n + 7) // 8] = np.packbits(m, bitorder="big")
    return out


@pytest.mark.parametrize("ncols", [1, 100, 1920, 1921, 5000])
def test_kernel_an_ajdk_v(ncols, oracle):
    from lakcluster_c_amd import lkc
    rng = np.random.default_rng(ncols)
    X = _packed(rng, 70, ncols)
    rows = np.sort(rng.permutation(70)[:50]).astype(np.int32)
    cm = _mask(rng, ncols)
    pc = lkc.bsize(ncols) // 240
    ajdk = oracle.calc_ajdk(rng.uniform(0.05, 0.95, pc))
    got = lkc.an_ajdk_v(X, rows, cm, ncols, ajdk)
    ref = oracle.an_ajdk_v(X, rows, cm, ncols, ajdk)
    assert got.tobytes() == ref.tobytes()


@pytest.mark.parametrize("centred", [False, True])
@pytest.mark.parametrize("shape", [(1, 1, 10), (65, 63, 2000), (130, 70, 4100)])
def test_kernel_anzt_vv(shape, centred, oracle):
    from lakcluster_c_amd import lkc
    na, nz, ncols = shape
    rng = np.random.default_rng(na * 7 + nz)
    A = _packed(rng, na + 3, ncols)
    Z = _packed(rng, nz + 2, ncols)
    ar = np.sort(rng.permutation(na + 3)[:na]).astype(np.int32)
    zr = np.sort(rng.permutation(nz + 2)[:nz]).astype(np.int32)
    cm = _mask(rng, ncols)
    if centred:
        pc = lkc.bsize(ncols) // 240
        ajdk = oracle.calc_ajdk(rng.uniform(0.05, 0.95, pc))
        xa, xz = rng.normal(size=na), rng.normal(size=nz)
    else:
        ajdk = xa = xz = None
    got = lkc.anzt_vv(A, ar, Z, zr, cm, ncols, ajdk, xa, xz)
    ref = oracle.anzt_vv(A, ar, Z, zr, cm, ncols, ajdk, xa, xz)
    assert got.tobytes() == ref.tobytes()


@pytest.mark.parametrize("B", [8, 24, 32])
def test_kernel_xcalc_quantiser(B):
    """xcalc against its defining properties: bounds, monotonicity, dequantisation error."""
    from lakcluster_c_amd import lkc
    rng = np.random.default_rng(B)
    L = rng.normal(size=(300, 7))
    cz, mn, mlt = lkc.xcalc(L, B)
    lo, hi = L.min(), L.max()
    assert mn == lo - 0.0625 * (hi - lo)
    assert mlt == (2.0 ** B) / ((hi + 0.0625 * (hi - lo)) - mn)
    assert cz.min() >= 1 and cz.max() <= 2 ** B - 1
    v = L.T.ravel()
    c = cz.ravel().astype(np.float64)
    order = np.argsort(v, kind="stable")
    assert np.all(np.diff(c[order]) >= 0)
    assert np.all(np.abs(mn + c / mlt - v) <= 1.0 / mlt * (1 + 1e-9))


@pytest.mark.parametrize("B", [16, 32])
def test_kernel_xcalc_sparse_rows_matches_oracle(B, oracle):
    """lkc_xcalc places row i of L at nrows_list[i] of an n_total-long layout, as get_xcalc
    does (xcalc.c:135-159), with zeros where no list entry points; bit-exact with the oracle."""
    from lakcluster_c_amd import lkc
    rng = np.random.default_rng(100 + B)
    nlist = np.sort(rng.permutation(500)[:300]).astype(np.int32)
    L = rng.normal(size=(300, 5)) * 3
    cz, mn, mlt = lkc.xcalc(L, B, nrows_list=nlist, n_total=500)
    rcz, rmn, rmlt, _ = oracle.xcalc(L, nlist, 500, B)
    assert cz.shape == (5, 500)
    np.testing.assert_array_equal(cz[:, nlist], rcz)
    off = np.ones(500, bool); off[nlist] = False
    assert not cz[:, off].any()
    assert (mn, mlt) == (rmn, rmlt)


@pytest.mark.parametrize("case", ["caseonly", "casectrl", "wide_z"])
def test_stage_arrays_after_initial_pass(case, datasets, oracle):
    """Every score array of the initial pass is bit-identical to the oracle's."""
    from lakcluster_c_amd import lkc
    text = _read(datasets[case].in_path)
    ref = oracle.State(text)
    bad = []
    with lkc.Bcc(text) as run:
        run.init_lf()
        for name in oracle.STATE_NAMES:
            r = ref.get(name)
            g = run.state(name)
            if r.shape != g.shape or r.tobytes() != g.tobytes():
                nd = int(np.sum(r != g)) if r.shape == g.shape else -1
                bad.append(f"{name}: {nd} of {r.size} differ")
    ref.close()
    assert not bad, "; ".join(bad)


@pytest.mark.parametrize("B", [16, 32])
@pytest.mark.parametrize("shape", [(40, 100, 1), (70, 2000, 3)])
def test_kernel_an_ztsn_ww(shape, B, oracle):
    """xcalc -> bit planes -> An_ZtSn_ww, bit-exact."""
    from lakcluster_c_amd import lkc
    nr, n_total, ns = shape
    rng = np.random.default_rng(nr + n_total + B)
    A = _packed(rng, nr, n_total)
    alist = np.sort(rng.permutation(nr)[: nr - 5]).astype(np.int32)
    nlist = np.sort(rng.permutation(n_total)[: int(n_total * 0.7)]).astype(np.int32)
    amask = np.zeros(lkc.bsize(n_total), dtype=np.uint8)
    bits = np.zeros(n_total, dtype=bool); bits[nlist] = True
    amask[: (n_total + 7) // 8] = np.packbits(bits, bitorder="big")
    L = rng.normal(size=(nlist.size, ns)) * 100
    got = lkc.an_ztsn_ww(A, alist, amask, L, nlist, n_total, B)
    ref = oracle.an_ztsn_ww(A, alist, amask, L, nlist, n_total, B)
    if got.tobytes() != ref.tobytes():
        i = np.argwhere(got != ref)[:3]
        raise AssertionError(f"{len(np.argwhere(got != ref))} differ; e.g. {[(tuple(x), got[tuple(x)], ref[tuple(x)]) for x in i]}")


@pytest.mark.parametrize("shape", [(1, 1), (64, 64), (65, 1921), (1000, 3001), (3, 5000), (2000, 65)])
def test_kernel_transpose_bits(shape):
    """Device transpose of packed bit rows equals numpy's transpose, padding zero."""
    from lakcluster_c_amd import lkc
    nrows, ncols = shape
    rng = np.random.default_rng(nrows * 7 + ncols)
    X = _packed(rng, nrows, ncols)
    Y = lkc.transpose_bits(X, nrows, ncols)
    bits = np.unpackbits(X, axis=1, bitorder="big")[:, :ncols]
    ref = np.zeros((ncols, lkc.bsize(nrows)), dtype=np.uint8)
    ref[:, : (nrows + 7) // 8] = np.packbits(bits.T, axis=1, bitorder="big")
    assert Y.shape == ref.shape
    assert np.array_equal(Y, ref)


@pytest.mark.parametrize("drop", ["n", "t"])
@pytest.mark.parametrize("case", ["casectrl", "masked", "blocks"])
def test_elimination_from_one_copy(case, drop, datasets, oracle):
    """A .in naming only one of A_n/A_t (and Z_n/Z_t) builds the other on the device;
    the trace and drop order are unchanged."""
    from lakcluster_c_amd import lkc
    text = _read(datasets[case].in_path)
    ref = oracle.run_text(text)
    keys = (f"GLOBAL_A_{drop}_name_", f"GLOBAL_Z_{drop}_name_")
    one = "\n".join(l for l in text.splitlines() if not l.strip().startswith(keys)) + "\n"
    assert one != text
    with lkc.Bcc(one) as run:
        run.run()
        tr = run.trace()
        xa, xb = run.xdrop()
    assert tr.tobytes() == ref.trace.tobytes()
    np.testing.assert_array_equal(xa, ref.xdrop_a)
    np.testing.assert_array_equal(xb, ref.xdrop_b)
