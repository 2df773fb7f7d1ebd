"""Brute-force loop counts and the reference's score post-processing, in numpy.

Test infrastructure: the element-wise definitions the oracle is pinned against.
Nothing here calls the oracle's score code; every formula is written out from the
reference's definitions (cited file:line).

  loop_counts     get_QR_AnZtSWnYt_uu / get_QC_AtTYnWtSZn_uu (bcc_flattenloop.c:1033-1328)
  halfloop_counts get_TAnZtS_uu / get_TAnAtT_uu (TAnZtS.c:1399-1600),
                  get_D_AtTn_ZtSn_uu / get_D_AtTn_AtTn_uu (D_AtTn_ZtSn.c:311-540)
  final_scores    bcc_sumscores_nrm / _ifT / _cmb (bcc_sumscores.c:324-779), and the dcc
                  twins dcc_sumscores_nrm / _ifT / _cmb (dcc_sumscores.c:248-699); one bin,
                  GLOBAL_Ireq <= 0
"""
import numpy as np

from lakcluster_c_amd import gen


def _centred(X, al, rows, kl):
    return (np.where(X, 1.0, -1.0) - al)[np.ix_(rows, kl)]


def loop_counts(A, Z, Tm, Sm, tab, ar, zr, kl, nt):
    """The full-loop subscores over the active rows ar, zr and columns kl:
      QR_TAT[a,t] = sum_{a'!=a} sum_{j!=j'} Ah(a,j)D_j Ah(a',j) T(a',t) Ah(a',j')D_j' Ah(a,j')
      QR_ZSZ[a,t] = sum_z      sum_{j!=j'} Ah(a,j)D_j Zh(z,j)  S(z,t)  Zh(z,j')D_j' Ah(a,j')
      QC_TAT[k,t] = D_k sum_{a!=a'} Ah(a,k)T(a,t) Ah(a',k)T(a',t) sum_{y!=k} Ah(a,y)D_y Ah(a',y)
      QC_ZSZ[k,t] = D_k sum_{a,z}   Ah(a,k)T(a,t) Zh(z,k)S(z,t)   sum_{y!=k} Ah(a,y)D_y Zh(z,y)
    with Ah = A(+-1) - alpha, Zh = Z(+-1) - alpha (alpha, D per 1920-column block)."""
    b = np.arange(A.shape[1]) // 1920
    al, D = tab[1][b], tab[5][b]
    Ah = _centred(A, al, ar, kl)
    Zh = _centred(Z, al, zr, kl) if Z is not None else None
    D = D[kl]
    T = np.where(Tm, 1.0, -1.0)[ar][:, :nt]
    S = np.where(Sm, 1.0, -1.0)[zr][:, :nt] if Z is not None else None
    G = (Ah * D) @ Ah.T
    G2 = (Ah * D) ** 2 @ (Ah ** 2).T
    W = G ** 2 - G2
    np.fill_diagonal(W, 0.0)
    out = {"QR_TAT": W @ T}
    if Z is not None:
        out["QR_ZSZ"] = (((Ah * D) @ Zh.T) ** 2 - (Ah * D) ** 2 @ (Zh ** 2).T) @ S
    qct = np.zeros((len(kl), nt)); qcz = np.zeros((len(kl), nt))
    for i in range(len(kl)):
        m = np.ones(len(kl), bool); m[i] = False
        Gk = (Ah[:, m] * D[m]) @ Ah[:, m].T
        np.fill_diagonal(Gk, 0.0)
        u = Ah[:, i:i + 1] * T
        qct[i] = D[i] * np.einsum("at,ab,bt->t", u, Gk, u)
        if Z is not None:
            v = Zh[:, i:i + 1] * S
            qcz[i] = D[i] * np.einsum("at,az,zt->t", u, (Ah[:, m] * D[m]) @ Zh[:, m].T, v)
    out["QC_TAT"] = qct
    if Z is not None:
        out["QC_ZSZ"] = qcz
    return out


def halfloop_counts(A, Z, Tm, Sm, tab, ar, zr, kl, nt):
    """The half-loop subscores of the dcc elimination:
      QR_TAnZtS[a,t] = sum_z S(z,t) sum_k Ah(a,k) D_k Zh(z,k)
      QR_TAnAtT[a,t] = sum_{a'!=a} T(a',t) sum_k Ah(a,k) D_k Ah(a',k)
      QC_TAnZtS[k,t] = D_k (sum_a Ah(a,k) T(a,t)) (sum_z Zh(z,k) S(z,t))
      QC_TAnAtT[k,t] = D_k sum_{a!=a'} Ah(a,k) T(a,t) Ah(a',k) T(a',t)
    returned under the names of the bcc arrays that hold them (QR_ZSZ, QR_TAT, QC_ZSZ, QC_TAT)."""
    b = np.arange(A.shape[1]) // 1920
    al, D = tab[1][b], tab[5][b]
    Ah = _centred(A, al, ar, kl)
    D = D[kl]
    T = np.where(Tm, 1.0, -1.0)[ar][:, :nt]
    G = (Ah * D) @ Ah.T
    np.fill_diagonal(G, 0.0)
    out = {"QR_TAT": G @ T}
    u = Ah.T @ T                                    # [k, t]
    out["QC_TAT"] = D[:, None] * (u * u - (Ah ** 2).sum(0)[:, None])
    if Z is not None:
        Zh = _centred(Z, al, zr, kl)
        S = np.where(Sm, 1.0, -1.0)[zr][:, :nt]
        out["QR_ZSZ"] = ((Ah * D) @ Zh.T) @ S
        out["QC_ZSZ"] = D[:, None] * u * (Zh.T @ S)
    return out


def _den(x):
    # dra_plusdivequals_s___m_m (ra_stats.c:70-76): a denominator below GLOBAL_tolerance becomes 1
    return float(x) if abs(x) >= 1e-12 else 1.0


def _mds_pow(x, s):
    # dra_mds_pow_s___m_m (ra_stats.c:86-104): where x[:,0] >= 0,
    # x[:,0] <- sgn(v) sqrt|v|, v = x[:,0]^2 - sum_t>=1 x[:,t]^2 / s
    x = x.copy()
    v = x[:, 0] ** 2 - (x[:, 1:] ** 2).sum(1) / s
    r = np.where(v >= 0, np.sqrt(np.abs(v)), -np.sqrt(np.abs(v)))
    x[:, 0] = np.where(x[:, 0] >= 0, r, x[:, 0])
    return x


def _mds_nrm(x, s):
    # dra_mds_nrm_s___m_m (ra_stats.c:77-85): x[:,0] <- x[:,0] - sum_t>=1 x[:,t] / s
    x = x.copy()
    x[:, 0] = x[:, 0] - x[:, 1:].sum(1) / s
    return x


def denominators(Ar, Zr, M, dcc=False):
    """The normalisers of the four subscores (see final_scores), after the tolerance rule."""
    if dcc:
        d = dict(QR_TAT=(Ar - 1) * M, QR_ZSZ=Zr * M, QC_TAT=Ar * (Ar - 1), QC_ZSZ=Ar * Zr)
    else:
        d = dict(QR_TAT=(Ar - 1) * M * (M - 1), QR_ZSZ=Zr * M * (M - 1),
                 QC_TAT=Ar * (M - 1) * (Ar - 1), QC_ZSZ=Ar * (M - 1) * Zr)
    return {k: _den(v) for k, v in d.items()}


def final_scores(raw, Ar, Zr, cpop, nt, kappa2, dcc=False):
    """The row and column scores that are sorted (first covariate column), from the raw
    subscores over the active rows / columns.

    Normalisers (one bin, Ireq <= 0), bcc (bcc_sumscores.c:634-669):
      QR_TAT / ((Ar-1) M (M-1)),  QR_ZSZ / (Zr M (M-1)),
      QC_TAT / (Ar (M-1) (Ar-1)), QC_ZSZ / (Ar (M-1) Zr);
    dcc (dcc_sumscores.c:553-588):
      QR_TAT / ((Ar-1) M),  QR_ZSZ / (Zr M),  QC_TAT / (Ar (Ar-1)),  QC_ZSZ / (Ar Zr).
    ifT (bcc_sumscores.c:483-548, dcc_sumscores.c:411-477): n_mds = min(6, max(0, T-1)); the
    scale is GLOBAL_kappa_squared whatever its sign (the `if (...);{ ... }` statement), so with
    kappa_squared = 0 nothing changes; else the QR terms take the pow form, the QC terms the
    linear form, with denominator n_mds * kappa_squared.
    cmb (bcc_sumscores.c:324-356, dcc_sumscores.c:248-280): QC = QC_TAT - QC_ZSZ,
    QR = QR_TAT - QR_ZSZ (first covariate column)."""
    d = denominators(Ar, Zr, cpop, dcc)
    nrm = {k: v / d[k] for k, v in raw.items()}
    n_mds = min(6, max(0, nt - 1))
    s = n_mds * kappa2
    if s > 0:
        for k in nrm:
            nrm[k] = _mds_pow(nrm[k], s) if k.startswith("QR") else _mds_nrm(nrm[k], s)
    QR = nrm["QR_TAT"][:, 0].copy()
    QC = nrm["QC_TAT"][:, 0].copy()
    if "QR_ZSZ" in nrm:
        QR = QR - nrm["QR_ZSZ"][:, 0]
        QC = QC - nrm["QC_ZSZ"][:, 0]
    return QR, QC


def load_problem(d, nZ, T):
    """The bits of a generated problem (gen.make_dataset) under directory d."""
    A = gen.read_b16(d + "/A_n.b16")
    Z = gen.read_b16(d + "/Z_n.b16") if nZ else None
    Tm = gen.read_b16(d + "/T_n.b16") if T > 1 else np.ones((A.shape[0], 1), bool)
    if Z is None:
        Sm = None
    else:
        Sm = gen.read_b16(d + "/S_n.b16") if T > 1 else np.ones((Z.shape[0], 1), bool)
    return A, Z, Tm, Sm


def problem_of(S, T):
    """The bits of the problem an oracle State holds (A, Z or None, T, S or None), for
    problems that come from no files (bcc_init's)."""
    NA = S.get("mask_Ar_j").size
    A = S.get("An_bits").reshape(NA, -1).astype(bool)
    NZ = S.get("mask_Zr_j").size
    Z = S.get("Zn_bits").reshape(NZ, -1).astype(bool) if NZ else None
    Tm = S.get("Tt_bits").reshape(T, NA).T.astype(bool)
    Sm = S.get("St_bits").reshape(T, NZ).T.astype(bool) if NZ else None
    return A, Z, Tm, Sm


def check_run(oracle, text, d, nZ, T, kappa2, dcc=False, max_steps=None, rtol=1e-6, problem=None):
    """Walks a whole oracle elimination: at every step the trace's mean row and column
    scores equal the means of the brute-force final scores, and the rows and columns the step
    drops are the lowest-scoring ones.  Tolerance: rtol of the larger of the step's largest
    score and the initial pass's largest subscores under the step's normalisers.  The low-rank
    update subtracts terms quantised to B_MLT bits at the scale of the subscores they came from,
    so late steps carry absolute errors of the initial scale (about 1e-9 of the current scale at
    32 bits early on, 1e-5 at a last 2 x 2 step).
    Returns (steps checked, drop sets compared)."""
    run = oracle.run_text(text)
    S = oracle.State(text)
    A, Z, Tm, Sm = problem if problem is not None else load_problem(d, nZ, T)
    tab = S.get("ajdk").reshape(15, -1)
    counts = loop_counts if not dcc else halfloop_counts
    tr = run.trace
    n = len(tr) if max_steps is None else min(len(tr), max_steps)
    pos = 0
    compared = 0
    init = None
    n_mds = min(6, max(0, T - 1))
    amp = 1.0 + (n_mds / (n_mds * kappa2) if n_mds * kappa2 > 0 else 0.0)
    for st in range(n):
        ar = np.flatnonzero(S.get("mask_Ar_j")); kl = np.flatnonzero(S.get("mask_Ac_j"))
        zr = np.flatnonzero(S.get("mask_Zr_j")) if Z is not None else None
        assert (len(ar), len(kl)) == (int(tr[st, 1]), int(tr[st, 2])), st
        raw = counts(A, Z, Tm, Sm, tab, ar, zr, kl, T)
        nz = len(zr) if zr is not None else 0
        QR, QC = final_scores(raw, len(ar), nz, len(kl), T, kappa2, dcc)
        if init is None:
            init = {k: np.abs(v).max() for k, v in raw.items()}
        den = denominators(len(ar), nz, len(kl), dcc)
        for col, v, pre in ((3, QR, "QR"), (4, QC, "QC")):
            want = v.mean() if v.size else np.nan
            want = want if np.isfinite(want) else 1.0
            carried = amp * sum(init[k] / den[k] for k in init if k.startswith(pre))
            scale = max(np.abs(v).max() if v.size else 0.0, carried, 1e-300)
            assert abs(tr[st, col] - want) <= rtol * scale, (st, col, tr[st, col], want)
        rdrop, cdrop = oracle.get_xdrop(len(ar), len(kl), oracle_gamma(text), oracle_gamma_type(text))
        rdrop, cdrop = min(rdrop, len(ar)), min(cdrop, len(kl))
        seg = run.xdrop_a[pos:pos + rdrop + cdrop]
        pos += rdrop + cdrop
        got_r = set(seg[:rdrop, 0].tolist()); got_c = set(seg[rdrop:, 1].tolist())
        for v, idx, k, got in ((QR, ar, rdrop, got_r), (QC, kl, cdrop, got_c)):
            if k == 0 or k == len(v):
                assert len(got) == k
                continue
            o = np.sort(v)
            gap = o[k] - o[k - 1]
            if gap > 10 * rtol * max(np.abs(v).max(), 1e-300):
                assert got == set(idx[np.argsort(v)[:k]].tolist()), st
                compared += 1
        if S.iterate():
            break
    S.close()
    return n, compared


def _key(text, name, default):
    for line in text.splitlines():
        if line.strip().startswith(name + "="):
            return line.split("=", 1)[1].strip().rstrip(";").strip()
    return default


def oracle_gamma(text):
    return float(_key(text, "GLOBAL_gamma", "0.125"))


def oracle_gamma_type(text):
    return int(_key(text, "GLOBAL_gamma_type", "0"))
