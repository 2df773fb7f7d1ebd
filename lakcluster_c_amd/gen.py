"""Seeded synthetic inputs for the lakcluster elimination (``.b16`` files + ``.in``).

The ``.b16`` layout follows the reference's writer (binary_read.c:57-81,
dir_m/tutorial_binary_compress.m): an ``int32`` header ``bitj, nrows, ncols``
followed by the matrix stored column by column, each column packed most
significant bit first and padded to a multiple of ``bitj`` bits.  Bit 1 means
+1 and bit 0 means -1.

For a patients-by-SNPs matrix ``A`` the reference reads ``A_n.b16`` (``A``
itself) as its SNP-major copy and ``A_t.b16`` (``A`` transposed) as its
patient-major copy (bcc.c:281-301); both are written here.

The planted model: each SNP column ``j`` is Bernoulli(``p_j``) with
``p_j ~ U[pmin, pmax]``; a bicluster of ``m`` rows by ``m'`` columns is
overwritten with the rank-one sign pattern ``u_i v_j`` with each entry flipped
with probability ``flip``.  Covariates (``T`` for cases, ``S`` for controls)
are an all-ones intercept column plus +-1 columns with P(+1)=0.5.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

BITJ = 16


def write_b16(path: str, mat: np.ndarray) -> None:
    """Write a boolean (nrows x ncols) matrix in the reference's .b16 format."""
    mat = np.asarray(mat, dtype=bool)
    nrows, ncols = mat.shape
    ext = (BITJ - nrows % BITJ) % BITJ
    with open(path, "wb") as fp:
        np.array([BITJ, nrows, ncols], dtype=np.int32).tofile(fp)
        step = max(1, (1 << 27) // max(1, nrows))
        for c0 in range(0, ncols, step):
            cols = mat[:, c0:c0 + step].T
            if ext:
                cols = np.concatenate([cols, np.zeros((cols.shape[0], ext), dtype=bool)], axis=1)
            np.packbits(cols, axis=1, bitorder="big").tofile(fp)


def read_b16(path: str) -> np.ndarray:
    """Read a .b16 file back into a boolean (nrows x ncols) matrix."""
    with open(path, "rb") as fp:
        bitj, nrows, ncols = np.fromfile(fp, dtype=np.int32, count=3)
        wp = (nrows + (bitj - nrows % bitj) % bitj) // 8
        raw = np.fromfile(fp, dtype=np.uint8, count=int(ncols) * int(wp)).reshape(int(ncols), int(wp))
    bits = np.unpackbits(raw, axis=1, bitorder="big")[:, :nrows]
    return bits.T.astype(bool)


@dataclass
class Dataset:
    """Paths and shapes of one generated problem."""
    in_path: str
    out_dir: str
    nA: int
    nZ: int
    ncols: int
    T: int
    gamma: float
    B: int
    planted_rows: List[np.ndarray] = field(default_factory=list)
    planted_cols: List[np.ndarray] = field(default_factory=list)


def _bernoulli_columns(rng: np.random.Generator, n: int, p: np.ndarray, chunk: int = 2048) -> np.ndarray:
    out = np.empty((n, p.size), dtype=bool)
    for r0 in range(0, n, chunk):
        r1 = min(n, r0 + chunk)
        out[r0:r1] = rng.random((r1 - r0, p.size), dtype=np.float32) < p[None, :]
    return out


def _plant(rng: np.random.Generator, A: np.ndarray, m: int, mc: int, flip: float) -> Tuple[np.ndarray, np.ndarray]:
    n, M = A.shape
    rows = np.sort(rng.permutation(n)[:m])
    cols = np.sort(rng.permutation(M)[:mc])
    u = rng.random(m) < 0.5
    v = rng.random(mc) < 0.5
    block = ~(u[:, None] ^ v[None, :])          # +1 where u_i v_j = +1
    block ^= rng.random((m, mc)) < flip
    A[np.ix_(rows, cols)] = block
    return rows, cols


def _covariates(rng: np.random.Generator, n: int, T: int) -> np.ndarray:
    cov = rng.random((n, T)) < 0.5
    cov[:, 0] = True
    return cov


def make_dataset(root: str, name: str, nA: int, ncols: int, nZ: int = 0, T: int = 1,
                 pmin: float = 0.05, pmax: float = 0.5, planted: Sequence[float] = (0.55,),
                 flip: float = 0.1, gamma: float = 0.05, B: int = 32, seed: int = 1,
                 gamma_type: Optional[int] = None, kappa_squared: Optional[float] = None,
                 row_mask_frac: float = 0.0, col_mask_frac: float = 0.0,
                 write_T_file: Optional[bool] = None, dup_cols: float = 0.0) -> Dataset:
    """Generate A (and optionally Z, T, S) and the .in file that drives the elimination.

    ``planted`` lists bicluster exponents e: m = round(nA**e) rows by
    round(ncols**e) columns, planted in A only.  ``row_mask_frac`` /
    ``col_mask_frac`` > 0 write GLOBAL_A_n_rind_/GLOBAL_A_n_cind masks with that
    fraction of entries switched off.  ``dup_cols`` > 0 makes that fraction of the
    columns copies of their left neighbour in the same 1920-column block, in A and in Z
    (duplicated SNPs: their column scores tie exactly).
    """
    rng = np.random.default_rng(seed)
    d_in = os.path.join(root, f"dir_{name}_in")
    d_out = os.path.join(root, f"dir_{name}_out")
    os.makedirs(d_in, exist_ok=True)
    os.makedirs(d_out, exist_ok=True)
    p = rng.uniform(pmin, pmax, size=ncols).astype(np.float32)
    A = _bernoulli_columns(rng, nA, p)
    prow, pcol = [], []
    for e in planted:
        m, mc = int(round(nA ** e)), int(round(ncols ** e))
        r, c = _plant(rng, A, min(m, nA), min(mc, ncols), flip)
        prow.append(r)
        pcol.append(c)
    dup = np.zeros(0, np.int64)
    if dup_cols > 0:
        j = np.arange(1, ncols)
        dup = j[(j // 1920 == (j - 1) // 1920) & (rng.random(ncols - 1) < dup_cols)]
    for j in dup:
        A[:, j] = A[:, j - 1]
    write_b16(os.path.join(d_in, "A_n.b16"), A)
    write_b16(os.path.join(d_in, "A_t.b16"), A.T)
    del A
    if nZ > 0:
        Z = _bernoulli_columns(rng, nZ, p)
        for j in dup:
            Z[:, j] = Z[:, j - 1]
        write_b16(os.path.join(d_in, "Z_n.b16"), Z)
        write_b16(os.path.join(d_in, "Z_t.b16"), Z.T)
        del Z
    if write_T_file is None:
        write_T_file = T > 1
    if write_T_file:
        Tm = _covariates(rng, nA, T)
        write_b16(os.path.join(d_in, "T_n.b16"), Tm)
        write_b16(os.path.join(d_in, "T_t.b16"), Tm.T)
        if nZ > 0:
            Sm = _covariates(rng, nZ, T)
            write_b16(os.path.join(d_in, "S_n.b16"), Sm)
            write_b16(os.path.join(d_in, "S_t.b16"), Sm.T)
    lines = [
        "GLOBAL_verbose= 0;",
        "GLOBAL_thread_count= 8;",
        "GLOBAL_omp_type= 1;",
        "GLOBAL_TEST_TYPE= lakcluster_driver;",
        "GLOBAL_QR_strategy= YnWt;",
        "GLOBAL_QC_strategy= YnWt;",
        "GLOBAL_NBINS= 1;",
        f"GLOBAL_B_MLT= {B};",
        f"GLOBAL_gamma= {gamma:.4f};",
    ]
    if gamma_type is not None:
        lines.append(f"GLOBAL_gamma_type= {gamma_type};")
    if kappa_squared is not None:
        lines.append(f"GLOBAL_kappa_squared= {kappa_squared};")
    lines += [
        f"GLOBAL_DIR_XPRE= {d_in};",
        "GLOBAL_A_n_name_= A_n.b16;",
        "GLOBAL_A_t_name_= A_t.b16;",
        f"GLOBAL_A_n_rows_= {nA};",
        f"GLOBAL_A_n_cols= {ncols};",
    ]
    if row_mask_frac > 0:
        write_b16(os.path.join(d_in, "mr_A.b16"), (rng.random(nA) >= row_mask_frac)[:, None])
        lines.append("GLOBAL_A_n_rind_= mr_A.b16;")
    if col_mask_frac > 0:
        write_b16(os.path.join(d_in, "mc_A.b16"), (rng.random(ncols) >= col_mask_frac)[:, None])
        lines.append("GLOBAL_A_n_cind= mc_A.b16;")
    if nZ > 0:
        lines += ["GLOBAL_Z_n_name_= Z_n.b16;", "GLOBAL_Z_t_name_= Z_t.b16;", f"GLOBAL_Z_n_rows_= {nZ};"]
    else:
        lines.append("GLOBAL_Z_n_rows_= 0;")
    lines.append("GLOBAL_Y_n_cols= 0;")
    lines.append(f"GLOBAL_T_n_cols= {T};")
    if write_T_file:
        lines += ["GLOBAL_T_n_name_= T_n.b16;", "GLOBAL_T_t_name_= T_t.b16;"]
        if nZ > 0:
            lines += ["GLOBAL_S_n_name_= S_n.b16;", "GLOBAL_S_t_name_= S_t.b16;"]
    lines += [f"GLOBAL_DIR_NAME= {d_out};", "END= 0;"]
    in_path = os.path.join(d_in, f"{name}.in")
    with open(in_path, "w") as fp:
        fp.write("\n".join(lines) + "\n")
    return Dataset(in_path, d_out, nA, nZ, ncols, T, gamma, B, prow, pcol)


def make_bins_dataset(root: str, name: str, bins: Sequence[Tuple[int, int]], ncols: int, ny: int = 0, T: int = 1,
                      pmin: float = 0.05, pmax: float = 0.5, planted: Sequence[float] = (0.55,), flip: float = 0.1,
                      gamma: float = 0.05, B: int = 32, seed: int = 1, gamma_type: Optional[int] = None,
                      kappa_squared: Optional[float] = None, row_mask_frac: float = 0.0, col_mask_frac: float = 0.0,
                      ycol_mask_frac: float = 0.0, tcol_mask: Optional[Sequence[bool]] = None,
                      write_T_file: Optional[bool] = None) -> Dataset:
    """A dexcluster_driver problem with several bins and Y/W genetic controls.

    ``bins`` lists each bin's (case rows, control rows); a bin may have no controls.  The data are drawn for all rows at
    once (A and Y over every bin's cases, Z and W over every bin's controls, the planted
    bicluster in A) and bin b takes the next block of rows, so the same seed with the rows
    split differently gives the same patients.  Per bin the .in names A_n/A_t, Z_n/Z_t,
    Y_n/Y_t, W_n/W_t, T_n/T_t, S_n/S_t and the row masks as comma-separated lists
    (updateglobals.c:76-117); the column masks of A, Y and T are shared.
    """
    rng = np.random.default_rng(seed)
    d_in = os.path.join(root, f"dir_{name}_in")
    d_out = os.path.join(root, f"dir_{name}_out")
    os.makedirs(d_in, exist_ok=True)
    os.makedirs(d_out, exist_ok=True)
    nb = len(bins)
    nA = int(sum(b[0] for b in bins))
    nZ = int(sum(b[1] for b in bins))
    p = rng.uniform(pmin, pmax, size=ncols).astype(np.float32)
    A = _bernoulli_columns(rng, nA, p)
    prow, pcol = [], []
    for e in planted:
        m, mc = int(round(nA ** e)), int(round(ncols ** e))
        r, c = _plant(rng, A, min(m, nA), min(mc, ncols), flip)
        prow.append(r)
        pcol.append(c)
    Z = _bernoulli_columns(rng, nZ, p)
    if write_T_file is None:
        write_T_file = T > 1
    Tm = _covariates(rng, nA, T)
    Sm = _covariates(rng, nZ, T)
    mrA = rng.random(nA) >= row_mask_frac
    mrZ = rng.random(nZ) >= row_mask_frac
    mcA = rng.random(ncols) >= col_mask_frac
    # the Y/W controls last: the same seed with and without them gives the same A, Z, T, S
    py = rng.uniform(pmin, pmax, size=ny).astype(np.float32)
    Y = _bernoulli_columns(rng, nA, py)
    W = _bernoulli_columns(rng, nZ, py)
    mcY = rng.random(ny) >= ycol_mask_frac
    # per key one item per bin, empty where the bin has no such file
    keys = {k: [""] * nb for k in ("A_n", "A_t", "Z_n", "Z_t", "Y_n", "Y_t", "W_n", "W_t", "T_n", "T_t", "S_n",
                                   "S_t", "A_rind", "Z_rind")}
    a0 = z0 = 0
    for b, (na, nz) in enumerate(bins):
        ra, rz = slice(a0, a0 + na), slice(z0, z0 + nz)
        a0 += na
        z0 += nz
        mats = [("A", A[ra]), ("Y", Y[ra])]
        if nz > 0:
            mats += [("Z", Z[rz]), ("W", W[rz])]
        if write_T_file:
            mats.append(("T", Tm[ra]))
            if nz > 0:
                mats.append(("S", Sm[rz]))
        for key, X in mats:
            if key in ("Y", "W") and ny <= 0:
                continue
            write_b16(os.path.join(d_in, f"{key}_n_{b}.b16"), X)
            write_b16(os.path.join(d_in, f"{key}_t_{b}.b16"), X.T)
            keys[f"{key}_n"][b] = f"{key}_n_{b}.b16"
            keys[f"{key}_t"][b] = f"{key}_t_{b}.b16"
        if row_mask_frac > 0:
            write_b16(os.path.join(d_in, f"mr_A_{b}.b16"), mrA[ra][:, None])
            keys["A_rind"][b] = f"mr_A_{b}.b16"
            if nz > 0:
                write_b16(os.path.join(d_in, f"mr_Z_{b}.b16"), mrZ[rz][:, None])
                keys["Z_rind"][b] = f"mr_Z_{b}.b16"
    lines = [
        "GLOBAL_verbose= 0;",
        "GLOBAL_thread_count= 8;",
        "GLOBAL_TEST_TYPE= dexcluster_driver;",
        f"GLOBAL_NBINS= {nb};",
        f"GLOBAL_B_MLT= {B};",
        f"GLOBAL_gamma= {gamma:.4f};",
    ]
    if gamma_type is not None:
        lines.append(f"GLOBAL_gamma_type= {gamma_type};")
    if kappa_squared is not None:
        lines.append(f"GLOBAL_kappa_squared= {kappa_squared};")
    lines += [f"GLOBAL_DIR_XPRE= {d_in};", f"GLOBAL_A_n_cols= {ncols};", f"GLOBAL_Y_n_cols= {max(ny, 0)};",
              f"GLOBAL_T_n_cols= {T};"]
    names = {"A_n": "GLOBAL_A_n_name_", "A_t": "GLOBAL_A_t_name_", "Z_n": "GLOBAL_Z_n_name_",
             "Z_t": "GLOBAL_Z_t_name_", "Y_n": "GLOBAL_Y_n_name_", "Y_t": "GLOBAL_Y_t_name_",
             "W_n": "GLOBAL_W_n_name_", "W_t": "GLOBAL_W_t_name_", "T_n": "GLOBAL_T_n_name_",
             "T_t": "GLOBAL_T_t_name_", "S_n": "GLOBAL_S_n_name_", "S_t": "GLOBAL_S_t_name_",
             "A_rind": "GLOBAL_A_n_rind_", "Z_rind": "GLOBAL_Z_n_rind_"}
    for k, key in names.items():
        if any(keys[k]):
            lines.append(f"{key}= {', '.join(keys[k])};")
    if col_mask_frac > 0:
        write_b16(os.path.join(d_in, "mc_A.b16"), mcA[:, None])
        lines.append("GLOBAL_A_n_cind= mc_A.b16;")
    if ycol_mask_frac > 0 and ny > 0:
        write_b16(os.path.join(d_in, "mc_Y.b16"), mcY[:, None])
        lines.append("GLOBAL_Y_n_cind= mc_Y.b16;")
    if tcol_mask is not None:
        write_b16(os.path.join(d_in, "mc_T.b16"), np.asarray(tcol_mask, dtype=bool)[:, None])
        lines.append("GLOBAL_T_n_cind= mc_T.b16;")
    lines += [f"GLOBAL_DIR_NAME= {d_out};", "END= 0;"]
    in_path = os.path.join(d_in, f"{name}.in")
    with open(in_path, "w") as fp:
        fp.write("\n".join(lines) + "\n")
    return Dataset(in_path, d_out, nA, nZ, ncols, T, gamma, B, prow, pcol)


def split_into_bins(ds: Dataset, root: str, name: str, nbins: int, ny: int = 0, seed: int = 1,
                    pmin: float = 0.05, pmax: float = 0.5) -> Dataset:
    """The patients of a make_dataset problem split into nbins equal bins for
    dexcluster_driver, with ny Y/W genetic-control columns added (Bernoulli columns with
    p ~ U[pmin, pmax], drawn for every case and control row).

    Bin b takes the next block of case rows and of control rows.  The .b16 files are cut by
    byte slicing (a row block of a column-by-column file is a byte range of every column),
    so the rows per bin must be a multiple of 16 in cases and controls."""
    d_src = os.path.dirname(ds.in_path)
    nA, nZ = ds.nA, ds.nZ
    if nA % (nbins * BITJ) or nZ % (nbins * BITJ):
        raise ValueError("rows per bin must be a multiple of 16")
    d_in = os.path.join(root, f"dir_{name}_in")
    d_out = os.path.join(root, f"dir_{name}_out")
    os.makedirs(d_in, exist_ok=True)
    os.makedirs(d_out, exist_ok=True)

    def cut(key, rows, b):
        nb = rows // nbins
        r0 = b * nb
        for suffix, patients_are_rows in (("n", True), ("t", False)):
            src = os.path.join(d_src, f"{key}_{suffix}.b16")
            with open(src, "rb") as fp:
                bitj, nr, nc = (int(v) for v in np.fromfile(fp, dtype=np.int32, count=3))
            wp = (nr + (bitj - nr % bitj) % bitj) // 8
            raw = np.memmap(src, dtype=np.uint8, mode="r", offset=12, shape=(nc, wp))
            dst = os.path.join(d_in, f"{key}_{suffix}_{b}.b16")
            with open(dst, "wb") as fp:
                if patients_are_rows:      # the (patients x columns) matrix: a byte range per column
                    np.array([bitj, nb, nc], dtype=np.int32).tofile(fp)
                    np.ascontiguousarray(raw[:, r0 // 8:(r0 + nb) // 8]).tofile(fp)
                else:                      # its transpose: one file column per patient
                    np.array([bitj, nr, nb], dtype=np.int32).tofile(fp)
                    np.ascontiguousarray(raw[r0:r0 + nb]).tofile(fp)
            del raw

    keys = {k: [""] * nbins for k in ("A_n", "A_t", "Z_n", "Z_t", "Y_n", "Y_t", "W_n", "W_t", "T_n", "T_t",
                                      "S_n", "S_t")}
    has_T = os.path.exists(os.path.join(d_src, "T_n.b16"))
    for b in range(nbins):
        for key, rows, present in (("A", nA, True), ("Z", nZ, nZ > 0), ("T", nA, has_T), ("S", nZ, has_T and nZ > 0)):
            if present:
                cut(key, rows, b)
                keys[f"{key}_n"][b] = f"{key}_n_{b}.b16"
                keys[f"{key}_t"][b] = f"{key}_t_{b}.b16"
    if ny > 0:
        rng = np.random.default_rng(seed)
        py = rng.uniform(pmin, pmax, size=ny).astype(np.float32)
        for key, rows in (("Y", nA), ("W", nZ)):
            if rows == 0:
                continue
            X = _bernoulli_columns(rng, rows, py)
            nb = rows // nbins
            for b in range(nbins):
                write_b16(os.path.join(d_in, f"{key}_n_{b}.b16"), X[b * nb:(b + 1) * nb])
                write_b16(os.path.join(d_in, f"{key}_t_{b}.b16"), X[b * nb:(b + 1) * nb].T)
                keys[f"{key}_n"][b] = f"{key}_n_{b}.b16"
                keys[f"{key}_t"][b] = f"{key}_t_{b}.b16"
            del X
    lines = ["GLOBAL_verbose= 0;", "GLOBAL_TEST_TYPE= dexcluster_driver;", f"GLOBAL_NBINS= {nbins};",
             f"GLOBAL_B_MLT= {ds.B};", f"GLOBAL_gamma= {ds.gamma:.4f};", f"GLOBAL_DIR_XPRE= {d_in};",
             f"GLOBAL_A_n_cols= {ds.ncols};", f"GLOBAL_Y_n_cols= {ny};", f"GLOBAL_T_n_cols= {ds.T};"]
    names = {"A_n": "GLOBAL_A_n_name_", "A_t": "GLOBAL_A_t_name_", "Z_n": "GLOBAL_Z_n_name_",
             "Z_t": "GLOBAL_Z_t_name_", "Y_n": "GLOBAL_Y_n_name_", "Y_t": "GLOBAL_Y_t_name_",
             "W_n": "GLOBAL_W_n_name_", "W_t": "GLOBAL_W_t_name_", "T_n": "GLOBAL_T_n_name_",
             "T_t": "GLOBAL_T_t_name_", "S_n": "GLOBAL_S_n_name_", "S_t": "GLOBAL_S_t_name_"}
    for k, key in names.items():
        if any(keys[k]):
            lines.append(f"{key}= {', '.join(keys[k])};")
    lines += [f"GLOBAL_DIR_NAME= {d_out};", "END= 0;"]
    in_path = os.path.join(d_in, f"{name}.in")
    with open(in_path, "w") as fp:
        fp.write("\n".join(lines) + "\n")
    return Dataset(in_path, d_out, nA, nZ, ds.ncols, ds.T, ds.gamma, ds.B)


# the benchmark-sized problems (rows x SNP columns)
CONFIGS = {
    "case_only_4096x16384": dict(nA=4096, ncols=16384, nZ=0, T=1, pmin=0.05, pmax=0.5,
                                 planted=(0.55,), gamma=0.05, B=32),
    "case_control_4096x16384": dict(nA=4096, ncols=16384, nZ=4096, T=3, pmin=0.05, pmax=0.5,
                                    planted=(0.55,), gamma=0.02, B=32),
    "case_control_16384x65536": dict(nA=16384, ncols=65536, nZ=16384, T=3, pmin=0.05, pmax=0.5,
                                     planted=(0.55,), gamma=0.02, B=32),
}


# ---------------------------------------------------------------------------
# pooled patients and permutation-null jobs (many eliminations on one GPU)
# ---------------------------------------------------------------------------
def pack_rows(bits: np.ndarray, rb: int) -> np.ndarray:
    """Boolean (n x m) -> packed rows of rb bytes, MSB first (the device row layout)."""
    n, m = bits.shape
    out = np.zeros((n, rb), dtype=np.uint8)
    out[:, : (m + 7) // 8] = np.packbits(bits, axis=1, bitorder="big")
    return out


def bsize(n: int) -> int:
    """Bytes of one packed row of n bits: whole 1920-bit blocks (lakcluster_header.h:268)."""
    e = n + (16 - n % 16) % 16
    return ((e + 1919) // 1920) * 240


@dataclass
class Pool:
    """Patients of a case-control problem pooled: rows (packed, nrows x bsize(ncols)),
    covariate columns as rows (ncov x bsize(nrows)); the first nA rows are the cases."""
    rows: np.ndarray
    cov: np.ndarray
    nrows: int
    ncols: int
    ncov: int
    nA: int


def make_pool(nA: int, nZ: int, ncols: int, T: int = 3, pmin: float = 0.05, pmax: float = 0.5,
              planted: Sequence[float] = (0.55,), flip: float = 0.1, seed: int = 1,
              chunk: int = 2048) -> Pool:
    """The case-control generator of make_dataset, kept packed in memory: A (with the
    planted bicluster) then Z, the same column frequencies, covariates T then S."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(pmin, pmax, size=ncols).astype(np.float32)
    rb = bsize(ncols)
    n = nA + nZ
    rows = np.zeros((n, rb), dtype=np.uint8)
    plants = []
    for e in planted:
        m, mc = min(int(round(nA ** e)), nA), min(int(round(ncols ** e)), ncols)
        r = np.sort(rng.permutation(nA)[:m]); cidx = np.sort(rng.permutation(ncols)[:mc])
        u = rng.random(m) < 0.5; v = rng.random(mc) < 0.5
        blk = ~(u[:, None] ^ v[None, :]); blk ^= rng.random((m, mc)) < flip
        plants.append((r, cidx, blk))
    for r0 in range(0, n, chunk):
        r1 = min(n, r0 + chunk)
        bits = rng.random((r1 - r0, ncols), dtype=np.float32) < p[None, :]
        for (r, cidx, blk) in plants:
            sel = (r >= r0) & (r < r1)
            if sel.any():
                bits[np.ix_(r[sel] - r0, cidx)] = blk[sel]
        rows[r0:r1] = pack_rows(bits, rb)
    cov_rows = np.concatenate([_covariates(rng, nA, T), _covariates(rng, nZ, T)], axis=0) if nZ else _covariates(rng, nA, T)
    cov = pack_rows(cov_rows.T, bsize(n))
    return Pool(rows, cov, n, ncols, T, nA)


@dataclass
class NullJob:
    rowsA: np.ndarray
    rowsZ: np.ndarray
    colmask: np.ndarray      # packed, bsize(ncols) bytes
    ncols_active: int


def null_jobs(pool: Pool, n_jobs: int, seed: int, rows: Tuple[int, int] = (2048, 16384),
              cols: Tuple[int, int] = (16384, 65536)) -> List[NullJob]:
    """Permutation nulls: each job takes a uniform row subsample of the pooled patients,
    log-uniform in size, whose case/control labels are a fresh shuffle of the pooled
    labels (dir_m/lakcluster_uADZSZDA_shuffle_1.m:1-2), and a uniform column
    subset, log-uniform in size, as a column mask."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_jobs):
        nr = int(round(np.exp(rng.uniform(np.log(rows[0]), np.log(min(rows[1], pool.nrows))))))
        nc = int(round(np.exp(rng.uniform(np.log(cols[0]), np.log(min(cols[1], pool.ncols))))))
        labels = rng.permutation(np.r_[np.ones(pool.nA, bool), np.zeros(pool.nrows - pool.nA, bool)])
        pick = np.sort(rng.permutation(pool.nrows)[:nr])
        a = pick[labels[pick]].astype(np.int32)
        z = pick[~labels[pick]].astype(np.int32)
        if a.size == 0:
            a, z = z[:1], z[1:]
        cm = np.zeros(pool.ncols, bool)
        cm[rng.permutation(pool.ncols)[:nc]] = True
        out.append(NullJob(a, z, pack_rows(cm[None, :], bsize(pool.ncols))[0], nc))
    return out


def write_job(root: str, name: str, pool: Pool, job: NullJob, gamma: float, B: int, gamma_type: int = 0,
              kappa_squared: float = 0.0) -> Dataset:
    """The same job as files and a .in (rows as A/Z, covariates, the column mask as
    mc_A.b16), for the oracle and for the file-driven path; a job with no Z rows is
    written as a case-only problem."""
    d_in = os.path.join(root, f"dir_{name}_in")
    d_out = os.path.join(root, f"dir_{name}_out")
    os.makedirs(d_in, exist_ok=True)
    os.makedirs(d_out, exist_ok=True)
    unpack = lambda R: np.unpackbits(R, axis=1, bitorder="big")[:, :pool.ncols].astype(bool)
    A = unpack(pool.rows[job.rowsA])
    nz = int(job.rowsZ.size)
    covb = np.unpackbits(pool.cov, axis=1, bitorder="big")[:, :pool.nrows].astype(bool).T   # nrows x T
    write_b16(os.path.join(d_in, "A_n.b16"), A); write_b16(os.path.join(d_in, "A_t.b16"), A.T)
    Tm = covb[job.rowsA]
    write_b16(os.path.join(d_in, "T_n.b16"), Tm); write_b16(os.path.join(d_in, "T_t.b16"), Tm.T)
    if nz:
        Z = unpack(pool.rows[job.rowsZ])
        write_b16(os.path.join(d_in, "Z_n.b16"), Z); write_b16(os.path.join(d_in, "Z_t.b16"), Z.T)
        Sm = covb[job.rowsZ]
        write_b16(os.path.join(d_in, "S_n.b16"), Sm); write_b16(os.path.join(d_in, "S_t.b16"), Sm.T)
    cm = np.unpackbits(job.colmask, bitorder="big")[:pool.ncols].astype(bool)
    write_b16(os.path.join(d_in, "mc_A.b16"), cm[:, None])
    # gamma in full (shortest round-trip) precision, so that the file's run is the job's
    lines = ["GLOBAL_verbose= 0;", "GLOBAL_thread_count= 8;", "GLOBAL_omp_type= 1;",
             "GLOBAL_TEST_TYPE= lakcluster_driver;", "GLOBAL_QR_strategy= YnWt;", "GLOBAL_QC_strategy= YnWt;",
             "GLOBAL_NBINS= 1;", f"GLOBAL_B_MLT= {B};", f"GLOBAL_gamma= {float(gamma)!r};",
             f"GLOBAL_gamma_type= {gamma_type};", f"GLOBAL_kappa_squared= {float(kappa_squared)!r};",
             f"GLOBAL_DIR_XPRE= {d_in};", "GLOBAL_A_n_name_= A_n.b16;", "GLOBAL_A_t_name_= A_t.b16;",
             f"GLOBAL_A_n_rows_= {A.shape[0]};", f"GLOBAL_A_n_cols= {pool.ncols};", "GLOBAL_A_n_cind= mc_A.b16;"]
    if nz:
        lines += ["GLOBAL_Z_n_name_= Z_n.b16;", "GLOBAL_Z_t_name_= Z_t.b16;", f"GLOBAL_Z_n_rows_= {nz};"]
    else:
        lines.append("GLOBAL_Z_n_rows_= 0;")
    lines += ["GLOBAL_Y_n_cols= 0;", f"GLOBAL_T_n_cols= {pool.ncov};",
              "GLOBAL_T_n_name_= T_n.b16;", "GLOBAL_T_t_name_= T_t.b16;"]
    if nz:
        lines += ["GLOBAL_S_n_name_= S_n.b16;", "GLOBAL_S_t_name_= S_t.b16;"]
    lines += [f"GLOBAL_DIR_NAME= {d_out};", "END= 0;"]
    in_path = os.path.join(d_in, f"{name}.in")
    with open(in_path, "w") as fp:
        fp.write("\n".join(lines) + "\n")
    return Dataset(in_path, d_out, A.shape[0], nz, pool.ncols, pool.ncov, gamma, B)


@dataclass
class DevicePool:
    """A pool drawn and kept on the device (``pool``: an lkc.Pool); the first nA rows are
    the cases.  ``cov`` holds the covariate columns as rows (ncov x bsize(nrows))."""
    pool: object
    cov: np.ndarray
    nrows: int
    ncols: int
    ncov: int
    nA: int
    planted_rows: List[np.ndarray]
    planted_cols: List[np.ndarray]


def make_pool_random(nA: int, nZ: int, ncols: int, T: int = 3, pmin: float = 0.01, pmax: float = 0.5,
                     planted: Sequence[float] = (0.5, 0.55, 0.6), flip: float = 0.1, seed: int = 1) -> DevicePool:
    """make_pool's column model drawn on the device by the library (lkc_pool_create_random),
    for problems too large for the host generator (65536 + 65536 patients x 524288 SNPs is
    6.9e10 draws): column frequencies p_j ~ U[pmin, pmax] and Bernoulli(p_j) entries from
    the library's counter-based generator keyed by ``seed``; biclusters planted in A
    (m = round(nA**e) rows by round(ncols**e) columns, rank-1 sign pattern u_i v_j with each
    entry flipped with probability ``flip``) and covariates T then S (an all-ones first
    column and i.i.d. +-1 columns) from the host generator seeded with ``seed``.  The
    problem is the same on every run and device (it is not the host generator's problem)."""
    from . import lkc
    rng = np.random.default_rng(seed)
    n = nA + nZ
    plants = []
    for e in planted:
        m, mc = min(int(round(nA ** e)), nA), min(int(round(ncols ** e)), ncols)
        r = np.sort(rng.permutation(nA)[:m]); cidx = np.sort(rng.permutation(ncols)[:mc])
        u = rng.random(m) < 0.5; v = rng.random(mc) < 0.5
        blk = ~(u[:, None] ^ v[None, :]); blk ^= rng.random((m, mc)) < flip
        plants.append((r, cidx, blk))
    cov_rows = np.concatenate([_covariates(rng, nA, T), _covariates(rng, nZ, T)], axis=0) if nZ else _covariates(rng, nA, T)
    cov = pack_rows(cov_rows.T, bsize(n))
    P = lkc.Pool.random(n, ncols, pmin, pmax, seed, plants, cov, T)
    return DevicePool(P, cov, n, ncols, T, nA, [r for r, _, _ in plants], [c for _, c, _ in plants])
