"""ctypes binding of the CPU oracle (oracle/bin/liblko_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "bin", "liblko_oracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    vp, ip, dp, cp, u8p = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_char_p, C.POINTER(C.c_uint8)
    sig = {
        "lko_bsize": (C.c_size_t, [C.c_int]),
        "lko_psize": (C.c_int, [C.c_int]),
        "lko_get_xdrop": (None, [C.c_double, C.c_double, C.c_double, C.c_int, ip, ip]),
        "lko_calc_ajdk": (None, [dp, C.c_int, dp]),
        "lko_quicksort_xij": (C.c_uint, [dp, ip, ip, ip, ip, ip, ip, C.c_int, C.c_int]),
        "lko_an_ajdk_v": (None, [u8p, C.c_size_t, ip, C.c_int, u8p, C.c_int, dp, dp]),
        "lko_anzt_vv": (None, [u8p, C.c_size_t, ip, C.c_int, u8p, C.c_size_t, ip, C.c_int, u8p, C.c_int, dp, dp, dp, dp]),
        "lko_run_text": (vp, [cp, cp, C.c_int, C.c_int]),
        "lko_run_text_recompute": (vp, [cp, C.c_int]),
        "lko_xcalc": (None, [dp, ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_longlong, vp, C.POINTER(C.c_ulonglong)]),
        "lko_planes_free": (None, [vp]),
        "lko_status": (C.c_int, [vp, cp, C.c_int]),
        "lko_n_trace": (C.c_int, [vp]),
        "lko_n_xdrop": (C.c_int, [vp]),
        "lko_n_iter": (C.c_int, [vp]),
        "lko_ops": (C.c_longlong, [vp]),
        "lko_times": (None, [vp, dp]),
        "lko_get_trace": (None, [vp, dp]),
        "lko_get_drops": (None, [vp, ip, ip]),
        "lko_free": (None, [vp]),
        "lko_state_new": (vp, [cp]),
        "lko_an_ztsn_ww_L": (None, [u8p, C.c_size_t, C.c_int, ip, C.c_int, u8p, dp, ip, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_longlong, dp]),
        "lko_state_iterate": (C.c_int, [vp]),
        "lko_state_get": (C.c_long, [vp, cp, dp, C.c_long]),
        "lko_state_free": (None, [vp]),
        "lko_scorebox": (C.c_int, [cp, C.c_int, ip, ip, ip, ip, dp, dp, cp, C.c_int]),
        "lko_scorebox_xdrop_array": (None, [C.c_int, C.c_int, C.c_int, ip, ip]),
        "lko_dcc_sumscores_test": (C.c_int, [vp, cp, dp, ip, ip, ip, ip]),
        "lko_dcc_driver_bins": (C.c_int, [cp, cp, ip, ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class OracleRun:
    def __init__(self, trace, xa, xb, ops, times, n_iter):
        self.trace, self.xdrop_a, self.xdrop_b, self.ops, self.times, self.n_iter = trace, xa, xb, ops, times, n_iter


def run_text(in_text: str, max_iter: int = -1, out_dir: Optional[str] = None, write_files: bool = False,
             recompute: bool = False) -> OracleRun:
    """Whole elimination; recompute=True replaces the low-rank update by full recomputation."""
    L = lib()
    if recompute:
        h = L.lko_run_text_recompute(in_text.encode(), max_iter)
    else:
        h = L.lko_run_text(in_text.encode(), out_dir.encode() if out_dir else None, max_iter, 1 if write_files else 0)
    err = C.create_string_buffer(256)
    st = L.lko_status(h, err, 256)
    if st != 0:
        L.lko_free(h)
        raise RuntimeError(f"oracle: {err.value.decode()}")
    nt, nx = L.lko_n_trace(h), L.lko_n_xdrop(h)
    tr = np.zeros((max(nt, 1), 6)); xa = np.zeros((max(nx, 1), 2), np.int32); xb = np.zeros((max(nx, 1), 2), np.int32)
    L.lko_get_trace(h, _p(tr, C.c_double))
    L.lko_get_drops(h, _p(xa, C.c_int), _p(xb, C.c_int))
    t = np.zeros(3)
    L.lko_times(h, _p(t, C.c_double))
    r = OracleRun(tr[:nt], xa[:nx], xb[:nx], int(L.lko_ops(h)), tuple(t), L.lko_n_iter(h))
    L.lko_free(h)
    return r


def get_xdrop(r: float, c: float, gamma: float, gamma_type: int = 0) -> Tuple[int, int]:
    a, b = C.c_int(0), C.c_int(0)
    lib().lko_get_xdrop(r, c, gamma, gamma_type, C.byref(a), C.byref(b))
    return a.value, b.value


def calc_ajdk(p: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.zeros(15 * p.size)
    lib().lko_calc_ajdk(_p(p, C.c_double), p.size, _p(out, C.c_double))
    return out


class _Planes(C.Structure):
    _fields_ = [("B", C.c_int), ("n_total", C.c_int), ("rb", C.c_size_t), ("nw", C.c_int),
                ("row0", C.POINTER(C.c_uint8)), ("planes", C.POINTER(C.c_uint8)),
                ("min_d", C.c_double), ("max_d", C.c_double), ("mlt_d", C.c_double)]


def xcalc(L: np.ndarray, nlist, n_total: int, B: int, D_MLT: Optional[int] = None):
    """get_xcalc of L (len(nlist) x nw): returns cz (nw x nn), min, mlt, planes (nw x B x rb bytes)."""
    L = np.ascontiguousarray(L, np.float64)
    nlist = np.ascontiguousarray(nlist, np.int32)
    nn, nw = L.shape
    D = (1 << B) if D_MLT is None else D_MLT
    P = _Planes()
    cz = np.zeros(max(nn * nw, 1), np.uint64)
    lib().lko_xcalc(_p(L, C.c_double), _p(nlist, C.c_int), nn, n_total, nw, B, D, C.byref(P), _p(cz, C.c_ulonglong))
    rb = P.rb
    planes = np.ctypeslib.as_array(P.planes, shape=(max(nw * B * rb, 1),)).copy()[: nw * B * rb].reshape(nw, B, rb)
    out = (cz[: nn * nw].reshape(nw, nn), P.min_d, P.mlt_d, planes)
    lib().lko_planes_free(C.byref(P))
    return out


def quicksort(d: np.ndarray):
    d = np.ascontiguousarray(d, dtype=np.float64).copy()
    ix = np.arange(d.size, dtype=np.int32)
    null = C.POINTER(C.c_int)()
    depth = lib().lko_quicksort_xij(_p(d, C.c_double), _p(ix, C.c_int), null, null, null, null, null, 0, d.size - 1) \
        if d.size else 0
    quicksort.depth = int(depth)
    return d, ix


def an_ajdk_v(X, rows, colmask, ncols, ajdk):
    X = np.ascontiguousarray(X, np.uint8); rows = np.ascontiguousarray(rows, np.int32)
    colmask = np.ascontiguousarray(colmask, np.uint8); ajdk = np.ascontiguousarray(ajdk, np.float64)
    out = np.zeros((max(rows.size, 1), 15))
    lib().lko_an_ajdk_v(_p(X, C.c_uint8), X.shape[1], _p(rows, C.c_int), rows.size, _p(colmask, C.c_uint8), ncols,
                        _p(ajdk, C.c_double), _p(out, C.c_double))
    return out[:rows.size]


def anzt_vv(A, arows, Z, zrows, colmask, ncols, ajdk=None, xa=None, xz=None):
    A = np.ascontiguousarray(A, np.uint8); Z = np.ascontiguousarray(Z, np.uint8)
    arows = np.ascontiguousarray(arows, np.int32); zrows = np.ascontiguousarray(zrows, np.int32)
    colmask = np.ascontiguousarray(colmask, np.uint8)
    out = np.zeros((max(arows.size, 1), max(zrows.size, 1)))
    dnull = C.POINTER(C.c_double)()
    f = lambda v: _p(np.ascontiguousarray(v, np.float64), C.c_double) if v is not None else dnull
    keep = [np.ascontiguousarray(v, np.float64) for v in (ajdk, xa, xz) if v is not None]
    lib().lko_anzt_vv(_p(A, C.c_uint8), A.shape[1], _p(arows, C.c_int), arows.size, _p(Z, C.c_uint8), Z.shape[1],
                      _p(zrows, C.c_int), zrows.size, _p(colmask, C.c_uint8), ncols, f(ajdk), f(xa), f(xz),
                      _p(out, C.c_double))
    del keep
    return out[:arows.size, :zrows.size]


def an_ztsn_ww(A, alist, amask, L, nlist, n_total, B, D_MLT=None):
    """xcalc of L (len(nlist) x ns) then An_ZtSn_ww for rows alist of A; returns (A rows, ns)."""
    A = np.ascontiguousarray(A, np.uint8); alist = np.ascontiguousarray(alist, np.int32)
    amask = np.ascontiguousarray(amask, np.uint8); L = np.ascontiguousarray(L, np.float64)
    nlist = np.ascontiguousarray(nlist, np.int32)
    ns = L.shape[1]
    out = np.zeros((A.shape[0], max(ns, 1)))
    D = (1 << B) if D_MLT is None else D_MLT
    lib().lko_an_ztsn_ww_L(_p(A, C.c_uint8), A.shape[1], A.shape[0], _p(alist, C.c_int), alist.size, _p(amask, C.c_uint8),
                           _p(L, C.c_double), _p(nlist, C.c_int), nlist.size, n_total, ns, B, D, _p(out, C.c_double))
    return out[:, :ns]


def scorebox_xdrop_array(nxols: int, x_max: int, x_num: int):
    xval, xdrop = np.zeros(max(x_num, 1), np.int32), np.zeros(max(x_num, 1), np.int32)
    lib().lko_scorebox_xdrop_array(nxols, x_max, x_num, _p(xval, C.c_int), _p(xdrop, C.c_int))
    return xval[:x_num], xdrop[:x_num]


class Scorebox:
    """Scorebox grid (row_num x col_num): active counts, Irem and the mean row and column scores."""

    def __init__(self, rpop, cpop, irem, qr, qc):
        self.A_rpop_j_total, self.A_cpop_j, self.Irem, self.QR_avg, self.QC_avg = rpop, cpop, irem, qr, qc


def scorebox(in_text: str, recompute: bool = False) -> Scorebox:
    L = lib()
    dims = np.zeros(4, np.int32)
    err = C.create_string_buffer(256)
    if L.lko_scorebox(in_text.encode(), 0, _p(dims, C.c_int), None, None, None, None, None, err, 256) != 0:
        raise RuntimeError(f"oracle: {err.value.decode()}")
    nr, nc = int(dims[1]), int(dims[3])
    ia = [np.zeros(nr * nc, np.int32) for _ in range(3)]
    da = [np.zeros(nr * nc) for _ in range(2)]
    if L.lko_scorebox(in_text.encode(), 1 if recompute else 0, _p(dims, C.c_int), *[_p(a, C.c_int) for a in ia],
                      *[_p(a, C.c_double) for a in da], err, 256) != 0:
        raise RuntimeError(f"oracle: {err.value.decode()}")
    # [nr + nc*row_num] -> (row_num, col_num)
    sh = lambda a: a.reshape(nc, nr).T.copy()
    return Scorebox(*[sh(a) for a in ia + da])


class DccParams(C.Structure):
    """lko_dcc_params: the GLOBAL_TEST_* settings of a dcc_sumscores test input."""
    _fields_ = [(n, C.c_int) for n in ("nbins", "A_n_rows", "A_n_cols", "Z_n_rows", "Y_n_cols", "T_n_cols", "B_MLT",
                                      "niter", "sparse", "gamma_type")] + \
               [(n, C.c_double) for n in ("mrand", "gamma", "kappa2")]


def dcc_sumscores_test(out_dir: Optional[str] = None, **kw):
    """dcc_sumscores_test on dcc_init's synthetic problem (several bins, Y/W controls).
    Returns (trace (n x 6), xdrop_a, xdrop_b); writes the three output files into out_dir."""
    p = DccParams(**{**dict(nbins=1, A_n_rows=0, A_n_cols=0, Z_n_rows=0, Y_n_cols=0, T_n_cols=1, B_MLT=32, niter=0,
                            sparse=1, gamma_type=0, mrand=0.5, gamma=0.125, kappa2=0.0), **kw})
    L = lib()
    nt, nx = C.c_int(0), C.c_int(0)
    if L.lko_dcc_sumscores_test(C.byref(p), None, None, C.byref(nt), None, None, C.byref(nx)) != 0:
        raise RuntimeError("oracle: dcc_sumscores_test failed")
    tr = np.zeros((max(nt.value, 1), 6)); xa = np.zeros((max(nx.value, 1), 2), np.int32); xb = np.zeros_like(xa)
    if L.lko_dcc_sumscores_test(C.byref(p), out_dir.encode() if out_dir else None, _p(tr, C.c_double), C.byref(nt),
                                _p(xa, C.c_int), _p(xb, C.c_int), C.byref(nx)) != 0:
        raise RuntimeError("oracle: dcc_sumscores_test failed")
    return tr[:nt.value], xa[:nx.value], xb[:nx.value]


def read_outputs(out_dir: str):
    """(trace (n x 6), xdrop_a (n x 2), xdrop_b (n x 2)) from the three output files."""
    def rd(name, dt, w):
        a = np.loadtxt(os.path.join(out_dir, name), dtype=dt, ndmin=2)
        return a.reshape(-1, w)
    return rd("out_trace.txt", np.float64, 6), rd("out_xdrop_a.txt", np.int64, 2), rd("out_xdrop_b.txt", np.int64, 2)


def dcc_driver_bins(in_text: str, out_dir: str):
    """dexcluster_driver on the .in's per-bin files (several bins, Y/W controls) to the end;
    writes the three output files into out_dir and returns them as read_outputs does."""
    os.makedirs(out_dir, exist_ok=True)
    nt, nx = C.c_int(0), C.c_int(0)
    rc = lib().lko_dcc_driver_bins(in_text.encode(), out_dir.encode(), C.byref(nt), C.byref(nx))
    if rc != 0:
        raise RuntimeError(f"oracle: dcc driver over bins failed ({rc})")
    return read_outputs(out_dir)


class State:
    """Stepwise oracle state: load + initial pass, then one loop pass per iterate()."""

    def __init__(self, in_text: str):
        self._h = lib().lko_state_new(in_text.encode())
        if not self._h:
            raise RuntimeError("oracle: load failed")

    def iterate(self) -> bool:
        return bool(lib().lko_state_iterate(self._h))

    def get(self, name: str) -> np.ndarray:
        n = lib().lko_state_get(self._h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(max(n, 1))
        lib().lko_state_get(self._h, name.encode(), _p(out, C.c_double), n)
        return out[:n]

    def close(self):
        if self._h:
            lib().lko_state_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


STATE_NAMES = ["lf_An_ajdk", "lf_Zn_ajdk", "lf_AtTn", "lf_ZtSn", "lf_AnZt", "lf_AnAt", "lf_AtTAnZtSZn",
               "lf_AtTAnAtTAn", "lf_AnZtSZnAt", "lf_AnAtTAnAt", "At_An_a1d1", "An_a2d2_ZtSn", "An_a2d2_AtTn",
               "a1d2_AtTn", "a3d2_AtTn", "a1d2_ZtSn", "a3d2_ZtSn", "et_Tn", "et_Sn", "et_An", "et_An_a1d1",
               "T_AnAt_AnAt", "QR_ZSZ", "QR_TAT", "QC_ZSZ", "QC_TAT"]
