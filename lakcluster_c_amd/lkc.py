"""ctypes binding of the gfx950 library ``liblkc.so`` (C ABI in ``include/lkc.h``).

The binding mirrors the reference's driver-level interface: an elimination is
configured by the key/value text of a lakcluster ``.in`` file
(updateglobals.c:5-199) and produces the trace and xdrop lists that
``lakcluster_driver`` writes (lakcluster_driver.c:14-251,
bcc_sumscores.c:127-157).  There is no CPU fallback: loading fails with an
error when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblkc.so")
DRIVER_PATH = os.path.join(_HERE, "lkc_driver")

_lib = None


class LkcError(RuntimeError):
    pass


# the stages of one pass of lakcluster_driver's loop, in its order (lakcluster_driver.c:121-246),
# after which lf_AnZt_S_WnYt, An_ajdk, lf_ZtSn, singlestudy_ww, doublestudy_ww and
# flattenloop run again; sumscores_dmp (the output files) may follow sumscores_mxA
INITIAL_STAGES = ("An_ajdk", "lf_ZtSn", "lf_YnWt", "lf_At_T_YnWt_S_Zn", "lf_AnZt_S_WnYt",
                  "singlestudy_ww", "doublestudy_ww", "flattenloop")
LOOP_STAGES = ("sumscores_nrm", "sumscores_ifT", "sumscores_srt", "sumscores_cmb", "sumscores_xij",
               "sumscores_mxA", "lrup_QR_YnWt_stage_0", "lrup_QR_YnWt_stage_1", "lrup_QR_YnWt_stage_3",
               "lrup_QC_ZtSWn_stage_0", "lrup_QC_ZtSWn_stage_1", "lrup_QC_ZtSWn_stage_2",
               "lrup_QC_YnWt_stage_a0", "lrup_QC_YnWt_stage_a1", "lrup_QC_YnWt_stage_a2",
               "lrup_QC_YnWt_stage_a3", "lrup_QC_YnWt_stage_b0", "lrup_QC_YnWt_stage_b1",
               "lrup_QC_YnWt_stage_b2", "lrup_mxdup", "M_mxset", "M_ZtSWn", "M_YnWt",
               "lf_AtTYn____WtSZn", "lrup_QC_YnWt_stage_c")
RESCORE_STAGES = ("lf_AnZt_S_WnYt", "An_ajdk", "lf_ZtSn", "singlestudy_ww", "doublestudy_ww", "flattenloop")


def lib():
    """Load liblkc.so (built by ``make -C lakcluster_c_amd/csrc``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LkcError(f"{LIB_PATH} is missing: build it with `make -C lakcluster_c_amd/csrc`")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp, cp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_char_p
    u8p = C.POINTER(C.c_uint8)
    sig = {
        "lkc_last_error": (cp, []),
        "lkc_bcc_error": (cp, [vp]),
        "lkc_bsize": (C.c_size_t, [C.c_int]),
        "lkc_lakcluster_driver": (C.c_int, [cp, cp]),
        "lkc_bcc_load": (C.c_int, [cp, C.POINTER(vp)]),
        "lkc_bcc_init_lf": (C.c_int, [vp]),
        "lkc_bcc_iterate": (C.c_int, [vp, ip]),
        "lkc_bcc_run": (C.c_int, [vp, C.c_int]),
        "lkc_bcc_write_outputs": (C.c_int, [vp, cp]),
        "lkc_bcc_An_ajdk": (C.c_int, [vp]),
        "lkc_bcc_lf_ZtSn": (C.c_int, [vp]),
        "lkc_bcc_lf_YnWt": (C.c_int, [vp]),
        "lkc_bcc_lf_At_T_YnWt_S_Zn": (C.c_int, [vp]),
        "lkc_bcc_lf_AnZt_S_WnYt": (C.c_int, [vp]),
        "lkc_bcc_singlestudy_ww": (C.c_int, [vp]),
        "lkc_bcc_doublestudy_ww": (C.c_int, [vp]),
        "lkc_bcc_flattenloop": (C.c_int, [vp]),
        "lkc_bcc_sumscores": (C.c_int, [vp]),
        "lkc_bcc_lrup": (C.c_int, [vp]),
        "lkc_bcc_n_iter": (C.c_int, [vp]),
        "lkc_bcc_n_trace": (C.c_int, [vp]),
        "lkc_bcc_get_trace": (C.c_int, [vp, dp]),
        "lkc_bcc_n_xdrop": (C.c_int, [vp]),
        "lkc_bcc_get_xdrop": (C.c_int, [vp, ip, ip]),
        "lkc_bcc_ops": (C.c_longlong, [vp]),
        "lkc_bcc_sort_fallbacks": (C.c_long, [vp]),
        "lkc_bcc_times": (C.c_int, [vp, dp]),
        "lkc_bcc_shape": (C.c_int, [vp, ip]),
        "lkc_bcc_free": (None, [vp]),
        "lkc_bcc_debug_get": (C.c_long, [vp, cp, dp, C.c_long]),
        "lkc_bcc_reset": (C.c_int, [vp]),
        "lkc_bcc_set_profiling": (C.c_int, [vp, C.c_int]),
        "lkc_bcc_set_kernels": (C.c_int, [vp, C.c_int]),
        "lkc_bcc_kernel_stats": (C.c_int, [vp, cp, C.POINTER(C.c_longlong), dp, dp]),
        "lkc_tune_set": (C.c_int, [C.c_int]),
        "lkc_tune_get": (C.c_int, [ip, ip, C.POINTER(C.c_longlong), dp]),
        "lkc_tune_save": (C.c_int, [cp]),
        "lkc_tune_verify": (C.c_int, [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
        "lkc_tune_load": (C.c_int, [cp]),
        "lkc_tune_clear": (C.c_int, []),
        "lkc_An_ajdk_v": (C.c_int, [u8p, C.c_size_t, C.c_int, ip, C.c_int, u8p, C.c_int, dp, dp]),
        "lkc_AnZt_vv": (C.c_int, [u8p, C.c_size_t, C.c_int, ip, C.c_int, u8p, C.c_size_t, C.c_int, ip, C.c_int,
                                  u8p, C.c_int, dp, dp, dp, dp]),
        "lkc_An_ZtSn_ww": (C.c_int, [u8p, C.c_size_t, C.c_int, ip, C.c_int, u8p, dp, ip, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_longlong, dp]),
        "lkc_xcalc": (C.c_int, [dp, ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_longlong,
                                C.POINTER(C.c_ulonglong), dp]),
        "lkc_transpose_bits": (C.c_int, [u8p, C.c_size_t, C.c_int, C.c_int, u8p, C.c_size_t]),
        "lkc_pool_create_random": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_ulonglong, C.c_int, ip, ip,
                                             ip, ip, u8p, u8p, C.c_size_t, C.c_int, C.POINTER(vp)]),
        "lkc_pool_read": (C.c_int, [vp, C.c_int, C.c_int, u8p]),
        "lkc_An_ZtSn_ww_noalloc": (C.c_int, [vp, C.c_int, vp, C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                                             C.c_int, C.c_longlong, vp, vp]),
        "lkc_pool_create": (C.c_int, [u8p, C.c_size_t, C.c_int, C.c_int, u8p, C.c_size_t, C.c_int, C.POINTER(vp)]),
        "lkc_pool_free": (None, [vp]),
        "lkc_bcc_load_job": (C.c_int, [vp, ip, C.c_int, ip, C.c_int, u8p, C.POINTER(JobParams), C.POINTER(vp)]),
        "lkc_batch_run": (C.c_int, [vp, C.POINTER(Job), C.c_int, C.c_int, C.c_double, C.POINTER(JobResult)]),
        "lkc_batch_results_free": (None, [C.POINTER(JobResult), C.c_int]),
        "lkc_job_bytes": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_int]),
        "lkc_lakcluster_scorebox": (C.c_int, [cp, cp]),
        "lkc_dexcluster_scorebox": (C.c_int, [cp, cp]),
        "lkc_scorebox_xdrop_array": (C.c_int, [C.c_int, C.c_int, C.c_int, ip, ip]),
        "lkc_bcc_scorebox": (C.c_int, [vp]),
        "lkc_bcc_get_scorebox": (C.c_int, [vp, ip, ip, ip, ip, dp, dp]),
        "lkc_bcc_write_scorebox": (C.c_int, [vp, cp]),
        "lkc_bcc_is_scorebox": (C.c_int, [vp]),
        "lkc_dexcluster_driver": (C.c_int, [cp, cp]),
        "lkc_dcc_sumscores_test": (C.c_int, [cp, cp]),
        "lkc_dcc_driver_bins": (C.c_int, [cp, cp]),
        "lkc_bcc_sumscores_test": (C.c_int, [cp, cp]),
        "lkc_bcc_is_dcc": (C.c_int, [vp]),
        "lkc_bcc_active": (C.c_int, [vp]),
        "lkc_bcc_sumscores_dmp": (C.c_int, [vp, cp]),
        "lkc_get_xdrop": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int, ip, ip]),
        "lkc_get_xdrop_length": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int]),
        "lkc_dquicksort_xij": (C.c_uint, [dp, ip, C.c_int, C.c_int]),
        "lkc_get_xdrop_array": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, ip, ip, ip, ip, ip]),
    }
    sig.update({"lkc_bcc_" + name: (C.c_int, [vp]) for name in LOOP_STAGES})
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int, handle=None) -> None:
    if rc != 0:
        msg = lib().lkc_bcc_error(handle) if handle else lib().lkc_last_error()
        raise LkcError(f"status {rc}: {msg.decode() if msg else ''}")


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class JobParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("gamma_type", C.c_int), ("B_MLT", C.c_int),
                ("kappa_squared", C.c_double), ("sparse", C.c_int)]


class Job(C.Structure):
    _fields_ = [("rowsA", C.POINTER(C.c_int)), ("nA", C.c_int), ("rowsZ", C.POINTER(C.c_int)), ("nZ", C.c_int),
                ("colmask", C.POINTER(C.c_uint8)), ("params", JobParams), ("arrival_s", C.c_double)]


class JobResult(C.Structure):
    _fields_ = [("status", C.c_int), ("n_iter", C.c_int), ("ops", C.c_longlong), ("arrival_s", C.c_double),
                ("start_s", C.c_double), ("end_s", C.c_double), ("n_trace", C.c_long),
                ("trace", C.POINTER(C.c_double)), ("n_xdrop", C.c_long), ("xdrop_a", C.POINTER(C.c_int)),
                ("xdrop_b", C.POINTER(C.c_int))]


class Bcc:
    """One elimination resident on the device (lkc_bcc handle)."""

    def __init__(self, in_text: str):
        self._h = C.c_void_p()
        _check(lib().lkc_bcc_load(in_text.encode(), C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().lkc_bcc_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def init_lf(self) -> None:
        _check(lib().lkc_bcc_init_lf(self._h), self._h)

    def iterate(self) -> bool:
        done = C.c_int(0)
        _check(lib().lkc_bcc_iterate(self._h, C.byref(done)), self._h)
        return bool(done.value)

    def run(self, max_iter: int = -1) -> None:
        _check(lib().lkc_bcc_run(self._h, max_iter), self._h)

    def stage(self, name: str) -> None:
        """One stage of the driver by its reference name without the bcc_ prefix
        (INITIAL_STAGES, LOOP_STAGES, RESCORE_STAGES, "sumscores", "lrup")."""
        if name not in INITIAL_STAGES + LOOP_STAGES + ("sumscores", "lrup"):
            raise LkcError(f"unknown stage {name}")
        _check(getattr(lib(), "lkc_bcc_" + name)(self._h), self._h)

    def active(self) -> bool:
        """True while a row or a column is active (the driver's loop condition)."""
        return bool(lib().lkc_bcc_active(self._h))

    def sumscores_dmp(self, out_dir: Optional[str] = None) -> None:
        _check(lib().lkc_bcc_sumscores_dmp(self._h, out_dir.encode() if out_dir else None), self._h)

    def write_outputs(self, out_dir: Optional[str] = None) -> None:
        _check(lib().lkc_bcc_write_outputs(self._h, out_dir.encode() if out_dir else None), self._h)

    @property
    def n_iter(self) -> int:
        return lib().lkc_bcc_n_iter(self._h)

    def trace(self) -> np.ndarray:
        n = lib().lkc_bcc_n_trace(self._h)
        out = np.zeros((max(n, 1), 6), dtype=np.float64)
        _check(lib().lkc_bcc_get_trace(self._h, _ptr(out, C.c_double)), self._h)
        return out[:n]

    def xdrop(self) -> Tuple[np.ndarray, np.ndarray]:
        n = lib().lkc_bcc_n_xdrop(self._h)
        a = np.zeros((max(n, 1), 2), dtype=np.int32)
        b = np.zeros((max(n, 1), 2), dtype=np.int32)
        _check(lib().lkc_bcc_get_xdrop(self._h, _ptr(a, C.c_int), _ptr(b, C.c_int)), self._h)
        return a[:n], b[:n]

    @property
    def ops(self) -> int:
        return int(lib().lkc_bcc_ops(self._h))

    @property
    def sort_fallbacks(self) -> int:
        """Score sorts that took the reference quicksort on the host (ties / NaN)."""
        return int(lib().lkc_bcc_sort_fallbacks(self._h))

    def times(self) -> Tuple[float, float, float]:
        t = np.zeros(3, dtype=np.float64)
        _check(lib().lkc_bcc_times(self._h, _ptr(t, C.c_double)), self._h)
        return float(t[0]), float(t[1]), float(t[2])

    def reset(self) -> None:
        _check(lib().lkc_bcc_reset(self._h), self._h)

    def set_profiling(self, on: bool = True) -> None:
        _check(lib().lkc_bcc_set_profiling(self._h, 1 if on else 0), self._h)

    def set_kernels(self, which: str) -> None:
        """Contractions on 'mfma' (int8 matrix cores, default) or 'valu' (popcount)."""
        _check(lib().lkc_bcc_set_kernels(self._h, {"mfma": 0, "valu": 1}[which]), self._h)

    def kernel_stats(self, name: str) -> Tuple[int, float, float]:
        """(launches, device ms, algorithmic work) of a timed kernel group."""
        n, ms, w = C.c_longlong(0), C.c_double(0), C.c_double(0)
        _check(lib().lkc_bcc_kernel_stats(self._h, name.encode(), C.byref(n), C.byref(ms), C.byref(w)), self._h)
        return int(n.value), float(ms.value), float(w.value)

    def state(self, name: str) -> np.ndarray:
        """A named score array of the device state (row-major, covariate last)."""
        n = lib().lkc_bcc_debug_get(self._h, name.encode(), None, 0)
        if n < 0:
            raise LkcError(f"unknown state array {name}")
        out = np.zeros(max(n, 1), dtype=np.float64)
        lib().lkc_bcc_debug_get(self._h, name.encode(), _ptr(out, C.c_double), n)
        return out[:n]

    def shape(self) -> Tuple[int, int, int, int]:
        s = np.zeros(4, dtype=np.int32)
        _check(lib().lkc_bcc_shape(self._h, _ptr(s, C.c_int)), self._h)
        return tuple(int(x) for x in s)

    @property
    def is_dcc(self) -> bool:
        """The handle runs the dcc half-loop elimination (GLOBAL_TEST_TYPE= dexcluster_driver)."""
        return bool(lib().lkc_bcc_is_dcc(self._h))

    @property
    def is_scorebox(self) -> bool:
        """GLOBAL_TEST_TYPE is lakcluster_scorebox."""
        return bool(lib().lkc_bcc_is_scorebox(self._h))

    def scorebox(self) -> "Scorebox":
        """The scorebox replay of GLOBAL_scorebox_out_xdrop (lakcluster_scorebox) on this
        freshly loaded handle; the grids are (row_num, col_num) arrays."""
        L = lib()
        _check(L.lkc_bcc_scorebox(self._h), self._h)
        dims = np.zeros(4, dtype=np.int32)
        _check(L.lkc_bcc_get_scorebox(self._h, _ptr(dims, C.c_int), None, None, None, None, None), self._h)
        nr, nc = int(dims[1]), int(dims[3])
        ia = [np.zeros(nr * nc, dtype=np.int32) for _ in range(3)]
        da = [np.zeros(nr * nc, dtype=np.float64) for _ in range(2)]
        _check(L.lkc_bcc_get_scorebox(self._h, _ptr(dims, C.c_int), *[_ptr(a, C.c_int) for a in ia],
                                      *[_ptr(a, C.c_double) for a in da]), self._h)
        grid = [a.reshape(nc, nr).T.copy() for a in ia + da]     # entries at [nr + nc*row_num]
        return Scorebox(tuple(int(x) for x in dims), *grid)

    def write_scorebox(self, out_dir: Optional[str] = None) -> None:
        _check(lib().lkc_bcc_write_scorebox(self._h, out_dir.encode() if out_dir else None), self._h)


class Scorebox:
    """struct S_handle outputs: A_rpop_j_total, A_cpop_j, Irem, QR_avg, QC_avg as
    (row_num, col_num) grids; dims = (row_max, row_num, col_max, col_num)."""

    def __init__(self, dims, rpop, cpop, irem, qr, qc):
        self.dims = dims
        self.A_rpop_j_total, self.A_cpop_j, self.Irem, self.QR_avg, self.QC_avg = rpop, cpop, irem, qr, qc


def lakcluster_scorebox(in_text: str, out_dir: Optional[str] = None) -> None:
    """Whole scorebox run with its five .mda files written (lakcluster_scorebox)."""
    _check(lib().lkc_lakcluster_scorebox(in_text.encode(), out_dir.encode() if out_dir else None))


def dexcluster_scorebox(in_text: str, out_dir: Optional[str] = None) -> None:
    """Whole dcc scorebox run with its five .mda files written (dexcluster_scorebox)."""
    _check(lib().lkc_dexcluster_scorebox(in_text.encode(), out_dir.encode() if out_dir else None))


def scorebox_xdrop_array(nxols: int, x_max: int, x_num: int) -> Tuple[np.ndarray, np.ndarray]:
    xval = np.zeros(max(x_num, 1), dtype=np.int32)
    xdrop = np.zeros(max(x_num, 1), dtype=np.int32)
    _check(lib().lkc_scorebox_xdrop_array(nxols, x_max, x_num, _ptr(xval, C.c_int), _ptr(xdrop, C.c_int)))
    return xval[:x_num], xdrop[:x_num]


def read_mda(path: str) -> np.ndarray:
    """An .mda file of rank 2 (int32 rank, dimensions, then column-major entries); the
    entry type follows from the file size (int32 or float64)."""
    raw = open(path, "rb").read()
    rank, nr, nc = np.frombuffer(raw[:12], dtype=np.int32)
    if rank != 2:
        raise LkcError(f"{path}: rank {rank}")
    body = raw[12:]
    dt = np.int32 if len(body) == 4 * nr * nc else np.float64
    return np.frombuffer(body, dtype=dt).reshape(nc, nr).T.copy()


def lakcluster_driver(in_text: str, out_dir: Optional[str] = None) -> None:
    """Whole run with the three output files written (lakcluster_driver)."""
    _check(lib().lkc_lakcluster_driver(in_text.encode(), out_dir.encode() if out_dir else None))


def dexcluster_driver(in_text: str, out_dir: Optional[str] = None) -> None:
    """Whole dcc run with the three output files written (dexcluster_driver); several bins
    or Y/W files take the multi-bin elimination."""
    _check(lib().lkc_dexcluster_driver(in_text.encode(), out_dir.encode() if out_dir else None))


def dcc_driver_bins(in_text: str, out_dir: Optional[str] = None) -> None:
    """dexcluster_driver with several bins and Y/W controls from the .in's per-bin file lists
    (any bin count, one included); the three output files written."""
    _check(lib().lkc_dcc_driver_bins(in_text.encode(), out_dir.encode() if out_dir else None))


def dcc_sumscores_test(in_text: str, out_dir: Optional[str] = None) -> None:
    """The reference's dcc self-test (dcc_sumscores_test) on dcc_init's synthetic problem from
    the GLOBAL_TEST_* keys: several bins, Y/W controls; the three output files written."""
    _check(lib().lkc_dcc_sumscores_test(in_text.encode(), out_dir.encode() if out_dir else None))


def bcc_sumscores_test(in_text: str, out_dir: Optional[str] = None) -> None:
    """The reference's bcc self-test (bcc_sumscores_test) on bcc_init's synthetic problem from
    the GLOBAL_TEST_* keys (one bin, no Y/W); the three output files written."""
    _check(lib().lkc_bcc_sumscores_test(in_text.encode(), out_dir.encode() if out_dir else None))


def tune_set(mode: int) -> None:
    """Launch-shape autotuner mode: 0 off (cost model), 1 measure new shape classes, 2 frozen."""
    _check(lib().lkc_tune_set(int(mode)))


def tune_get() -> dict:
    """The autotuner's mode, stored shape classes, classes measured, and the device
    milliseconds its choices saved over the cost model's at measurement."""
    m, n = C.c_int(), C.c_int()
    meas, saved = C.c_longlong(), C.c_double()
    _check(lib().lkc_tune_get(C.byref(m), C.byref(n), C.byref(meas), C.byref(saved)))
    return {"mode": m.value, "classes": n.value, "measured": meas.value, "saved_ms": saved.value}


def tune_verify() -> Tuple[int, int]:
    """(launches compared, 8-byte words that differed) over the tuning measurements made in
    mode 3, where every form's output is compared with the first form's."""
    c, m = C.c_longlong(), C.c_longlong()
    _check(lib().lkc_tune_verify(C.byref(c), C.byref(m)))
    return c.value, m.value


def tune_save(path: str) -> int:
    rc = lib().lkc_tune_save(path.encode())
    _check(min(rc, 0))
    return rc


def tune_load(path: str) -> int:
    rc = lib().lkc_tune_load(path.encode())
    _check(min(rc, 0))
    return rc


def tune_clear() -> None:
    _check(lib().lkc_tune_clear())


def dquicksort_xij(d: np.ndarray) -> Tuple[np.ndarray, np.ndarray, int]:
    """dQuickSort_xij (Quicksort.c:124-141) on a copy of d: (sorted values, permutation,
    deepest recursion level)."""
    d = np.ascontiguousarray(d, dtype=np.float64).copy()
    ix = np.arange(d.size, dtype=np.int32)
    depth = lib().lkc_dquicksort_xij(_ptr(d, C.c_double), _ptr(ix, C.c_int), 0, d.size - 1) if d.size else 0
    return d, ix, int(depth)


def get_xdrop(lrij: float, lcij: float, gamma: float, gamma_type: int = 0) -> Tuple[int, int]:
    """get_xdrop (xdrop.c:5-64): (rows, columns) one pass drops."""
    r, c = C.c_int(0), C.c_int(0)
    _check(lib().lkc_get_xdrop(lrij, lcij, gamma, gamma_type, C.byref(r), C.byref(c)))
    return r.value, c.value


def get_xdrop_array(lrij: float, lcij: float, gamma: float, gamma_type: int = 0) -> np.ndarray:
    """get_xdrop_array (xdrop.c:85-111): rows [rdrop, cdrop, rkeep, ckeep] per pass."""
    n = lib().lkc_get_xdrop_length(lrij, lcij, gamma, gamma_type)
    if n < 0:
        _check(n)
    out = np.zeros((4, max(n, 1)), dtype=np.int32)
    ln = C.c_int(0)
    _check(lib().lkc_get_xdrop_array(lrij, lcij, gamma, gamma_type, n, C.byref(ln), *(_ptr(out[i], C.c_int) for i in range(4))))
    return out[:, :ln.value].T.copy()


def bsize(ncols: int) -> int:
    return int(lib().lkc_bsize(ncols))


def transpose_bits(X: np.ndarray, nrows: int, ncols: int) -> np.ndarray:
    """wrap_transpose on the device: packed rows X (nrows x rb) -> (ncols x bsize(nrows))."""
    X = np.ascontiguousarray(X, dtype=np.uint8)
    Y = np.zeros((max(ncols, 1), bsize(nrows)), dtype=np.uint8)
    _check(lib().lkc_transpose_bits(_ptr(X, C.c_uint8), X.shape[1], nrows, ncols, _ptr(Y, C.c_uint8), Y.shape[1]))
    return Y[:ncols]


def an_ajdk_v(X: np.ndarray, rows: Sequence[int], colmask: np.ndarray, ncols: int, ajdk: np.ndarray) -> np.ndarray:
    """get_An_ajdk_v on packed rows X (nrows_total x rb bytes); returns (len(rows), 15)."""
    X = np.ascontiguousarray(X, dtype=np.uint8)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    colmask = np.ascontiguousarray(colmask, dtype=np.uint8)
    ajdk = np.ascontiguousarray(ajdk, dtype=np.float64)
    out = np.zeros((max(rows.size, 1), 15), dtype=np.float64)
    _check(lib().lkc_An_ajdk_v(_ptr(X, C.c_uint8), X.shape[1], X.shape[0], _ptr(rows, C.c_int), rows.size,
                               _ptr(colmask, C.c_uint8), ncols, _ptr(ajdk, C.c_double), _ptr(out, C.c_double)))
    return out[:rows.size]


def anzt_vv(A: np.ndarray, arows, Z: np.ndarray, zrows, colmask: np.ndarray, ncols: int,
            ajdk: Optional[np.ndarray] = None, xa=None, xz=None) -> np.ndarray:
    """get_AnZt_vv on packed rows; returns (len(arows), len(zrows))."""
    A = np.ascontiguousarray(A, dtype=np.uint8)
    Z = np.ascontiguousarray(Z, dtype=np.uint8)
    arows = np.ascontiguousarray(arows, dtype=np.int32)
    zrows = np.ascontiguousarray(zrows, dtype=np.int32)
    colmask = np.ascontiguousarray(colmask, dtype=np.uint8)
    out = np.zeros((max(arows.size, 1), max(zrows.size, 1)), dtype=np.float64)
    dnull = C.POINTER(C.c_double)()
    aj = _ptr(np.ascontiguousarray(ajdk, dtype=np.float64), C.c_double) if ajdk is not None else dnull
    xa_ = np.ascontiguousarray(xa, dtype=np.float64) if xa is not None else None
    xz_ = np.ascontiguousarray(xz, dtype=np.float64) if xz is not None else None
    _check(lib().lkc_AnZt_vv(_ptr(A, C.c_uint8), A.shape[1], A.shape[0], _ptr(arows, C.c_int), arows.size,
                             _ptr(Z, C.c_uint8), Z.shape[1], Z.shape[0], _ptr(zrows, C.c_int), zrows.size,
                             _ptr(colmask, C.c_uint8), ncols, aj,
                             _ptr(xa_, C.c_double) if xa_ is not None else dnull,
                             _ptr(xz_, C.c_double) if xz_ is not None else dnull, _ptr(out, C.c_double)))
    return out[:arows.size, :zrows.size]


def xcalc(L: np.ndarray, B: int, D_MLT: Optional[int] = None, nrows_list=None,
          n_total: Optional[int] = None) -> Tuple[np.ndarray, float, float]:
    """get_xcalc quantiser on L (nn x nw), row i of L placed at nrows_list[i] (default i) of an
    n_total-long layout (default nn); returns cz (nw x n_total), min, mlt."""
    L = np.ascontiguousarray(L, dtype=np.float64)
    nn, nw = L.shape
    D = (1 << B) if D_MLT is None else D_MLT
    rows = np.arange(nn, dtype=np.int32) if nrows_list is None else np.ascontiguousarray(nrows_list, dtype=np.int32)
    nt = nn if n_total is None else int(n_total)
    cz = np.zeros((max(nw, 1), max(nt, 1)), dtype=np.uint64)
    mm = np.zeros(2, dtype=np.float64)
    _check(lib().lkc_xcalc(_ptr(L, C.c_double), _ptr(rows, C.c_int), nn, nt, nw, B, D,
                           _ptr(cz, C.c_ulonglong), _ptr(mm, C.c_double)))
    return cz[:nw, :nt], float(mm[0]), float(mm[1])


def an_ztsn_ww(A: np.ndarray, alist, amask: np.ndarray, L: np.ndarray, nlist, n_total: int, B: int,
               D_MLT: Optional[int] = None) -> np.ndarray:
    """xcalc of L (len(nlist) x ns) then An_ZtSn_ww for rows alist of A; returns (A rows, ns)."""
    A = np.ascontiguousarray(A, dtype=np.uint8)
    alist = np.ascontiguousarray(alist, dtype=np.int32)
    amask = np.ascontiguousarray(amask, dtype=np.uint8)
    L = np.ascontiguousarray(L, dtype=np.float64)
    nlist = np.ascontiguousarray(nlist, dtype=np.int32)
    ns = L.shape[1]
    out = np.zeros((A.shape[0], max(ns, 1)), dtype=np.float64)
    D = (1 << B) if D_MLT is None else D_MLT
    _check(lib().lkc_An_ZtSn_ww(_ptr(A, C.c_uint8), A.shape[1], A.shape[0], _ptr(alist, C.c_int), alist.size,
                                _ptr(amask, C.c_uint8), _ptr(L, C.c_double), _ptr(nlist, C.c_int), nlist.size,
                                n_total, ns, B, D, _ptr(out, C.c_double)))
    return out[:, :ns]


def _params(gamma: float, B: int, gamma_type: int = 0, kappa_squared: float = 0.0, sparse: int = 1) -> JobParams:
    return JobParams(gamma, gamma_type, B, kappa_squared, sparse)


class Pool:
    """Patients kept on the device (lkc_pool): packed rows (nrows x rb) and the
    covariate columns as rows (ncov x rbc)."""

    def __init__(self, rows: np.ndarray, nrows: int, ncols: int, cov: np.ndarray, ncov: int):
        self._keep = (np.ascontiguousarray(rows, dtype=np.uint8), np.ascontiguousarray(cov, dtype=np.uint8))
        r, c = self._keep
        self._h = C.c_void_p()
        _check(lib().lkc_pool_create(_ptr(r, C.c_uint8), r.shape[1], nrows, ncols, _ptr(c, C.c_uint8), c.shape[1],
                                     ncov, C.byref(self._h)))
        self.nrows, self.ncols, self.ncov = nrows, ncols, ncov

    @classmethod
    def random(cls, nrows: int, ncols: int, pmin: float, pmax: float, seed: int, plants, cov: np.ndarray,
               ncov: int) -> "Pool":
        """A pool drawn on the device (lkc_pool_create_random): Bernoulli(p_j) entries with
        p_j ~ U[pmin, pmax] from the library's counter-based generator keyed by seed, then the
        planted blocks, a list of (rows, cols, bits m x mc bool), written over them."""
        self = cls.__new__(cls)
        c = np.ascontiguousarray(cov, dtype=np.uint8)
        pm = np.array([len(r) for r, _, _ in plants], np.int32)
        pmc = np.array([len(cc) for _, cc, _ in plants], np.int32)
        pr = np.ascontiguousarray(np.concatenate([np.asarray(r, np.int32) for r, _, _ in plants]) if plants
                                  else np.zeros(1, np.int32), dtype=np.int32)
        pc = np.ascontiguousarray(np.concatenate([np.asarray(cc, np.int32) for _, cc, _ in plants]) if plants
                                  else np.zeros(1, np.int32), dtype=np.int32)
        pb = np.ascontiguousarray(np.concatenate([np.asarray(b, np.uint8).ravel() for _, _, b in plants]) if plants
                                  else np.zeros(1, np.uint8), dtype=np.uint8)
        if not plants:
            pm = pmc = np.zeros(1, np.int32)
        self._keep = (c,)
        self._h = C.c_void_p()
        _check(lib().lkc_pool_create_random(nrows, ncols, pmin, pmax, C.c_ulonglong(seed), len(plants),
                                            _ptr(pm, C.c_int), _ptr(pmc, C.c_int), _ptr(pr, C.c_int),
                                            _ptr(pc, C.c_int), _ptr(pb, C.c_uint8), _ptr(c, C.c_uint8), c.shape[1],
                                            ncov, C.byref(self._h)))
        self.nrows, self.ncols, self.ncov = nrows, ncols, ncov
        return self

    def read(self, r0: int, n: int) -> np.ndarray:
        """Rows r0 .. r0+n-1 of the device-resident pool, packed (n x bsize(ncols))."""
        out = np.zeros((max(n, 1), bsize(self.ncols)), dtype=np.uint8)
        _check(lib().lkc_pool_read(self._h, r0, n, _ptr(out, C.c_uint8)))
        return out[:n]

    def close(self) -> None:
        if self._h:
            lib().lkc_pool_free(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def job(self, rowsA, rowsZ, colmask: Optional[np.ndarray], gamma: float, B: int, **kw) -> "Bcc":
        """One job as a Bcc handle (lkc_bcc_load_job)."""
        a = np.ascontiguousarray(rowsA, dtype=np.int32)
        z = np.ascontiguousarray(rowsZ, dtype=np.int32)
        cm = _ptr(np.ascontiguousarray(colmask, dtype=np.uint8), C.c_uint8) if colmask is not None else C.POINTER(C.c_uint8)()
        h = C.c_void_p()
        prm = _params(gamma, B, **kw)
        _check(lib().lkc_bcc_load_job(self._h, _ptr(a, C.c_int), a.size, _ptr(z, C.c_int), z.size, cm,
                                      C.byref(prm), C.byref(h)))
        b = Bcc.__new__(Bcc)
        b._h = h
        return b

    def batch(self, jobs, max_concurrent: int, budget_bytes: float = 0.0, arrivals=None):
        """Run jobs [(rowsA, rowsZ, colmask, gamma, B[, {gamma_type, kappa_squared}])] with at
        most max_concurrent at a time (lkc_batch_run).  Returns a list of dicts: status,
        n_iter, ops, arrival/start/end seconds, trace (n x 6), xdrop_a, xdrop_b (n x 2)."""
        n = len(jobs)
        keep = []
        arr = (Job * max(n, 1))()
        for j, job in enumerate(jobs):
            ra, rz, cm, gamma, B = job[:5]
            kw = job[5] if len(job) > 5 else {}
            a = np.ascontiguousarray(ra, dtype=np.int32); z = np.ascontiguousarray(rz, dtype=np.int32)
            keep += [a, z]
            arr[j].rowsA = _ptr(a, C.c_int); arr[j].nA = a.size
            arr[j].rowsZ = _ptr(z, C.c_int); arr[j].nZ = z.size
            if cm is not None:
                m = np.ascontiguousarray(cm, dtype=np.uint8); keep.append(m)
                arr[j].colmask = _ptr(m, C.c_uint8)
            arr[j].params = _params(gamma, B, **kw)
            arr[j].arrival_s = float(arrivals[j]) if arrivals is not None else 0.0
        res = (JobResult * max(n, 1))()
        _check(lib().lkc_batch_run(self._h, arr, n, max_concurrent, budget_bytes, res))
        out = []
        for j in range(n):
            r = res[j]
            tr = np.ctypeslib.as_array(r.trace, shape=(max(r.n_trace, 1) * 6,))[: r.n_trace * 6].reshape(-1, 6).copy()
            xa = np.ctypeslib.as_array(r.xdrop_a, shape=(max(r.n_xdrop, 1) * 2,))[: r.n_xdrop * 2].reshape(-1, 2).copy()
            xb = np.ctypeslib.as_array(r.xdrop_b, shape=(max(r.n_xdrop, 1) * 2,))[: r.n_xdrop * 2].reshape(-1, 2).copy()
            out.append(dict(status=r.status, n_iter=r.n_iter, ops=r.ops, arrival_s=r.arrival_s, start_s=r.start_s,
                            end_s=r.end_s, trace=tr, xdrop_a=xa, xdrop_b=xb))
        lib().lkc_batch_results_free(res, n)
        return out


def job_bytes(nA: int, nZ: int, ncols: int, B: int) -> float:
    return float(lib().lkc_job_bytes(nA, nZ, ncols, B))
