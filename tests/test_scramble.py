"""bcc_scramble (R_init.c:165-193, 256-285): the bits of the (row, column) pairs named by
an out_xdrop file are redrawn before the elimination.

The oracle is pinned against an independent restatement that uses exact arithmetic:
the generator's closed form seed_n = seed_0 * 35^n mod (2^31 - 1) (R01GET,
boxmuller.c:57-66, with the constants of lakcluster_ver18.c:204-205), its value
seed_n / (2^31 - 1) rounded first to the 64-bit significand of the C long double
and then to a double, the row-major draw order over the ascending, mask-pruned
index lists, repeated indices redrawn in place, and p_k = (ones of column k over
the loaded Z rows) / their number.  The engine is checked against the oracle on
the whole elimination that follows."""
import os
from fractions import Fraction

import numpy as np
import pytest

P = 2147483647


def _rne(x: Fraction) -> int:
    """x rounded to the nearest integer, ties to even."""
    f = x.numerator // x.denominator
    r = x - f
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and f % 2 == 1):
        f += 1
    return f


def _r01(seed: int) -> float:
    """double((long double)seed / (long double)P) for 0 < seed < P."""
    v = Fraction(seed, P)
    e = 0
    while v * Fraction(2) ** (-e) < 1:
        e -= 1
    while v * Fraction(2) ** (-e) >= 2:
        e += 1
    m64 = _rne(v * Fraction(2) ** (63 - e))          # 64-bit significand
    if m64 == 1 << 64:
        m64, e = 1 << 63, e + 1
    m53 = _rne(Fraction(m64, 1 << 11))                # then to 53 bits
    return float(Fraction(m53) * Fraction(2) ** (e - 52))


def test_generator_values_exact():
    # the double-rounded quotient against the platform long double on the seeds a
    # scramble of seed 1 starts with
    s = 1
    for _ in range(2000):
        s = s * 35 % P
        r = _r01(s)
        assert 0.0 < r < 1.0
        assert r == float(np.longdouble(s) / np.longdouble(P)) or np.finfo(np.longdouble).nmant < 63


def _write_xdrop(path, rows, cols):
    with open(path, "w") as f:
        for a in rows:
            f.write(f"{a} -1\n")
        for k in cols:
            f.write(f"-1 {k}\n")


def _scrambled_text(ds, files, seeds=None):
    text = open(ds.in_path).read().replace("END= 0;", "").rstrip()
    text += f"\nGLOBAL_scramble_num= {len(files)};\nGLOBAL_scramble_out_xdrop_= {','.join(files)};\n"
    if seeds is not None:
        text += f"GLOBAL_scramble_rseed_= {','.join(str(s) for s in seeds)};\n"
    return text + "END= 0;\n"


def _expected(A0, Z, rows, cols, nA, M, seed, rmask=None, cmask=None):
    rows = sorted(a for a in rows if 0 <= a < nA and (rmask is None or rmask[a]))
    cols = sorted(k for k in cols if 0 <= k < M and (cmask is None or cmask[k]))
    A = A0.copy()
    zrb = Z.shape[0]
    zc = Z.sum(axis=0) if zrb else np.zeros(M)
    s = seed
    for a in rows:
        for k in cols:
            s = s * 35 % P
            p = float(zc[k]) / float(zrb) if zrb else float("nan")
            A[a, k] = 1 if _r01(s) < p else 0
    return A


@pytest.mark.parametrize("name", ["casectrl", "caseonly"])
def test_oracle_scramble_bits(tmp_path, datasets, oracle, name):
    ds = datasets[name]
    base = oracle.State(open(ds.in_path).read())
    nA = int(base.get("mask_Ar_j").size)
    M = int(base.get("mask_Ac_j").size)
    A0 = base.get("An_bits").reshape(nA, M).astype(np.int8)
    Zf = base.get("Zn_bits")
    Z = Zf.reshape(-1, M).astype(np.int64) if Zf.size else np.zeros((0, M), np.int64)
    base.close()
    # rows and columns with repeats, indices out of range and -1 pairs
    rows = [5, 17, 3, 17, nA + 4, 0]
    cols = [10, 2, 400 % M, 2, M - 1, M + 7]
    f1 = str(tmp_path / "x1.txt")
    _write_xdrop(f1, rows, cols)
    st = oracle.State(_scrambled_text(ds, [f1], [7]))
    A1 = st.get("An_bits").reshape(nA, M).astype(np.int8)
    At1 = st.get("At_bits").reshape(M, nA).astype(np.int8)
    st.close()
    exp = _expected(A0, Z, rows, cols, nA, M, 7)
    np.testing.assert_array_equal(A1, exp)
    np.testing.assert_array_equal(At1, exp.T)
    changed = np.argwhere(A1 != A0)
    assert all(a in rows and k in cols for a, k in changed)
    if name == "caseonly":
        # no controls: p = 0 / 0, so every redrawn bit is 0
        sel = np.ix_(sorted({a for a in rows if a < nA}), sorted({k for k in cols if k < M}))
        assert not A1[sel].any()


def test_oracle_two_scrambles_in_sequence(tmp_path, datasets, oracle):
    ds = datasets["casectrl"]
    base = oracle.State(open(ds.in_path).read())
    nA, M = int(base.get("mask_Ar_j").size), int(base.get("mask_Ac_j").size)
    A0 = base.get("An_bits").reshape(nA, M).astype(np.int8)
    Z = base.get("Zn_bits").reshape(-1, M).astype(np.int64)
    base.close()
    r1, c1 = list(range(0, 40, 3)), list(range(5, 300, 7))
    r2, c2 = list(range(1, 30, 2)), list(range(0, 200, 5))
    f1, f2 = str(tmp_path / "a.txt"), str(tmp_path / "b.txt")
    _write_xdrop(f1, r1, c1)
    _write_xdrop(f2, r2, c2)
    st = oracle.State(_scrambled_text(ds, [f1, f2]))      # default seeds 1 and 2
    A2 = st.get("An_bits").reshape(nA, M).astype(np.int8)
    st.close()
    exp = _expected(_expected(A0, Z, r1, c1, nA, M, 1), Z, r2, c2, nA, M, 2)
    np.testing.assert_array_equal(A2, exp)


def test_scramble_missing_file_is_an_error(tmp_path, datasets, oracle):
    ds = datasets["casectrl"]
    text = _scrambled_text(ds, [str(tmp_path / "absent.txt")])
    with pytest.raises(RuntimeError):
        oracle.State(text)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["casectrl", "masked", "caseonly"])
def test_engine_scramble_matches_oracle(tmp_path, datasets, oracle, name):
    from lakcluster_c_amd import lkc
    ds = datasets[name]
    # the scramble list of a null run: the tail of the elimination order of the data itself
    f1 = str(tmp_path / "tail.txt")
    r = oracle.run_text(open(ds.in_path).read())
    pairs = r.xdrop_a.reshape(-1, 2)
    with open(f1, "w") as f:
        for a, b in pairs[len(pairs) // 2:]:
            f.write(f"{int(a)} {int(b)}\n")
    f2 = str(tmp_path / "head.txt")
    with open(f2, "w") as f:
        for a, b in pairs[: len(pairs) // 5]:
            f.write(f"{int(a)} {int(b)}\n")
    text = _scrambled_text(ds, [f1, f2], [11, 3])
    print(f"scramble {name}: {len(pairs)} xdrop lines", flush=True)
    ref = oracle.run_text(text)
    with lkc.Bcc(text) as run:
        run.run()
        tr = run.trace()
        xa, xb = run.xdrop()
    assert tr.tobytes() == ref.trace.tobytes()
    np.testing.assert_array_equal(xa, ref.xdrop_a)
    np.testing.assert_array_equal(xb, ref.xdrop_b)
