"""The dcc half-loop elimination (dexcluster_driver, dexcluster_driver.c:5-57), one bin,
no Y/W controls.

Pins of the oracle: the half-loop subscores equal the reference's brute-force
element-wise definitions (get_TAnZtS_uu / get_TAnAtT_uu, TAnZtS.c:1399-1600;
get_D_AtTn_ZtSn_uu / get_D_AtTn_AtTn_uu, D_AtTn_ZtSn.c:311-540), written out in numpy,
at the initial pass and after elimination steps:
    QR_TAnZtS[a,t] = sum_z S(z,t) sum_k Ah(a,k) D_k Zh(z,k)
    QR_TAnAtT[a,t] = sum_{a'!=a} T(a',t) sum_k Ah(a,k) D_k Ah(a',k)
    QC_TAnZtS[k,t] = D_k (sum_a Ah(a,k) T(a,t)) (sum_z Zh(z,k) S(z,t))
    QC_TAnAtT[k,t] = D_k sum_{a!=a'} Ah(a,k) T(a,t) Ah(a',k) T(a',t)
with Ah = A(+-1) - alpha, Zh = Z(+-1) - alpha (alpha, D per 1920-column block); the QR
terms go through the B_MLT-bit quantisation (relative 1e-8 at 32 bits), the QC terms are
exact products.  Whole runs: at every step the trace means equal the means of the
normalised (dcc_sumscores_nrm), covariate-corrected (_ifT) and combined (_cmb) brute-force
half-loop scores, the dropped rows and columns are the lowest-scoring ones, and every row
and column is dropped once with the get_xdrop counts.
The engine is held to the oracle bit for bit.

The reference's committed dcc fixture (dir_test/, from dir_in/dcc_sumscores_error.in)
runs two bins with Y/W controls on data drawn by dcc_init's generator; those settings are
outside this path, so it pins only the drop schedule (tests/test_golden.py)."""
import numpy as np
import pytest

import loopcount


def _dcc_text(ds):
    return open(ds.in_path).read().replace("GLOBAL_TEST_TYPE= lakcluster_driver;", "GLOBAL_TEST_TYPE= dexcluster_driver;")


@pytest.mark.parametrize("kw,steps", [
    (dict(nA=12, ncols=20, nZ=10, T=3, seed=5, gamma=0.1, planted=()), 0),
    (dict(nA=14, ncols=30, nZ=0, T=1, seed=6, gamma=0.1), 0),
    (dict(nA=16, ncols=25, nZ=11, T=2, seed=7, gamma=0.1, row_mask_frac=0.2, col_mask_frac=0.2), 0),
    (dict(nA=16, ncols=40, nZ=12, T=3, seed=8, gamma=0.1), 3),
    (dict(nA=40, ncols=4000, nZ=30, T=3, seed=9, gamma=0.1), 2),
])
def test_dcc_scores_match_halfloop_definition(kw, steps, oracle, tmp_path):
    from lakcluster_c_amd import gen
    ds = gen.make_dataset(str(tmp_path), "bf", **kw)
    d = str(tmp_path / "dir_bf_in")
    S = oracle.State(_dcc_text(ds))
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
    bf = loopcount.halfloop_counts(A, Z, Tm, Sm, tab, ar, zr, kl, nt)
    for name, want in bf.items():
        rows = kl if name.startswith("QC") else ar
        got = S.get(name).reshape(-1, nt)[rows]
        scale = np.abs(want).max()
        tol = 1e-12 if name.startswith("QC") else 1e-8
        np.testing.assert_allclose(got, want, rtol=0, atol=tol * scale, err_msg=name)
    S.close()


@pytest.mark.parametrize("kw", [
    dict(nA=14, ncols=36, nZ=11, T=3, seed=31, gamma=0.1),
    dict(nA=18, ncols=40, nZ=0, T=1, seed=32, gamma=0.1),
    dict(nA=16, ncols=30, nZ=12, T=2, seed=33, gamma=0.12, row_mask_frac=0.15, col_mask_frac=0.15),
    dict(nA=15, ncols=34, nZ=13, T=3, seed=34, gamma=0.1, kappa_squared=0.5),
    dict(nA=14, ncols=3900, nZ=10, T=3, seed=36, gamma=0.2, B=24),
])
def test_dcc_trace_and_drops_match_halfloop_scores(kw, oracle, tmp_path):
    from lakcluster_c_amd import gen
    ds = gen.make_dataset(str(tmp_path), "ns", **kw)
    n, compared = loopcount.check_run(oracle, _dcc_text(ds), str(tmp_path / "dir_ns_in"), kw["nZ"], kw["T"],
                                      kw.get("kappa_squared", 0.0), dcc=True,
                                      rtol=1e-6 if kw.get("B", 32) == 32 else 1e-4)
    assert n >= 5 and compared >= n // 2


@pytest.mark.parametrize("name", ["caseonly", "casectrl", "masked", "kappa"])
def test_dcc_run_invariants(oracle, datasets, name):
    ds = datasets[name]
    r = oracle.run_text(_dcc_text(ds))
    rows = r.xdrop_a[r.xdrop_a[:, 0] >= 0, 0]
    cols = r.xdrop_a[r.xdrop_a[:, 1] >= 0, 1]
    assert len(set(rows.tolist())) == len(rows) == int(r.trace[0, 1])
    assert len(set(cols.tolist())) == len(cols) == int(r.trace[0, 2])
    for s in range(len(r.trace) - 1):
        rd, cd = oracle.get_xdrop(r.trace[s, 1], r.trace[s, 2], ds.gamma)
        assert r.trace[s + 1, 1] == r.trace[s, 1] - min(rd, r.trace[s, 1])
        assert r.trace[s + 1, 2] == r.trace[s, 2] - min(cd, r.trace[s, 2])
    assert np.all(np.isfinite(r.trace[:, 3:5]))
    # a different path from bcc's on the same data
    b = oracle.run_text(open(ds.in_path).read())
    assert not np.array_equal(b.trace[:, 3:5], r.trace[:, 3:5])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["caseonly", "casectrl", "masked", "blocks", "b16_g1", "kappa", "tiny", "wide_z"])
def test_engine_dcc_matches_oracle(datasets, oracle, name):
    from lakcluster_c_amd import lkc
    ds = datasets[name]
    text = _dcc_text(ds)
    print(f"dcc {name}", flush=True)
    ref = oracle.run_text(text)
    with lkc.Bcc(text) as run:
        assert run.is_dcc
        run.run()
        tr = run.trace()
        xa, xb = run.xdrop()
        ops = run.ops
    assert tr.tobytes() == ref.trace.tobytes()
    np.testing.assert_array_equal(xa, ref.xdrop_a)
    np.testing.assert_array_equal(xb, ref.xdrop_b)
    assert ops == ref.ops


def _no_strategy(text, qc=None):
    lines = [l for l in text.splitlines() if "_strategy" not in l]
    if qc:
        lines.insert(0, f"GLOBAL_QC_strategy= {qc};")
    return "\n".join(lines) + "\n"


def test_oracle_dcc_reads_no_strategy(datasets, oracle):
    """dexcluster_driver reads no QR/QC strategy (dcc.c:1167): a .in without the strategy
    lines, or with the dcc files' "YnWt store one", runs the same elimination."""
    text = _dcc_text(datasets["casectrl"])
    ref = oracle.run_text(text)
    for t in (_no_strategy(text), _no_strategy(text, "YnWt store one")):
        r = oracle.run_text(t)
        assert r.trace.tobytes() == ref.trace.tobytes()


@pytest.mark.gpu
def test_engine_dcc_reads_no_strategy(datasets, oracle):
    from lakcluster_c_amd import lkc
    text = _no_strategy(_dcc_text(datasets["casectrl"]), "YnWt store one")
    print("dcc without strategy lines", flush=True)
    ref = oracle.run_text(text)
    with lkc.Bcc(text) as run:
        run.run()
        tr = run.trace()
        xa, _ = run.xdrop()
    assert tr.tobytes() == ref.trace.tobytes()
    np.testing.assert_array_equal(xa, ref.xdrop_a)


@pytest.mark.gpu
def test_engine_dcc_output_files(tmp_path, datasets, oracle):
    from lakcluster_c_amd import lkc
    ds = datasets["casectrl"]
    text = _dcc_text(ds)
    a, b = str(tmp_path / "eng"), str(tmp_path / "ora")
    lkc.dexcluster_driver(text, a)
    oracle.run_text(text, out_dir=b, write_files=True)
    for f in ("out_trace.txt", "out_xdrop_a.txt", "out_xdrop_b.txt"):
        assert open(f"{a}/{f}").read() == open(f"{b}/{f}").read(), f
    # the entry point runs dcc whatever the .in names; lkc_lakcluster_driver runs bcc
    c = str(tmp_path / "bcc")
    lkc.lakcluster_driver(text, c)
    assert open(f"{c}/out_trace.txt").read() != open(f"{a}/out_trace.txt").read()


@pytest.mark.gpu
def test_engine_dcc_scramble_matches_oracle(tmp_path, datasets, oracle):
    """wrap_dcc_scramble and dcc_scramble (R_init.c:90-163) are bcc_scramble's twins: the same load-time redraw."""
    from lakcluster_c_amd import lkc
    ds = datasets["casectrl"]
    r = oracle.run_text(_dcc_text(ds))
    f1 = str(tmp_path / "tail.txt")
    with open(f1, "w") as f:
        for a, b in r.xdrop_a.reshape(-1, 2)[len(r.xdrop_a) // 2:]:
            f.write(f"{int(a)} {int(b)}\n")
    text = _dcc_text(ds).replace("END= 0;", "").rstrip()
    text += f"\nGLOBAL_scramble_num= 1;\nGLOBAL_scramble_out_xdrop_= {f1};\nGLOBAL_scramble_rseed_= 5;\nEND= 0;\n"
    ref = oracle.run_text(text)
    assert not np.array_equal(ref.trace, r.trace)
    with lkc.Bcc(text) as run:
        run.run()
        tr = run.trace()
    assert tr.tobytes() == ref.trace.tobytes()
