"""ctypes binding of ``libreno_hip.so`` (declared in ``include/reno_hip.h``).

This module is the only place that talks to the C ABI.  Tensors are torch
tensors on the current HIP device; raw device pointers and the current
stream handle are handed to the library, which launches asynchronously on
that stream.  There is no fallback: if the library is missing, or no GPU is
present, every compute entry point raises.
"""
import collections
import ctypes
import os
import threading

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libreno_hip.so")

RENO_OK = 0
RENO_EINVAL = -1
RENO_EHIP = -2
RENO_ENOSPACE = -3
RENO_ESOLVER = -4
RENO_EQN = -5

PREC_FP64 = 0
PREC_MIXED = 1

# exported symbols of include/reno_hip.h
EXPORTS = (
    "reno_version", "reno_status_string", "reno_last_error",
    "reno_hop_workspace_bytes", "reno_krylov_workspace_bytes",
    "reno_env_workspace_bytes", "reno_qr_workspace_bytes",
    "reno_hop_apply", "reno_hop_apply_noalloc", "reno_env_update", "reno_env_close", "reno_expm_krylov",
    "reno_expm_krylov_async", "reno_expm_krylov_mixed_async", "reno_block_qr", "reno_heff_diag", "reno_heff_dense", "reno_zgemm",
    "reno_svd_workspace_bytes", "reno_block_svd", "reno_eigh_workspace_bytes", "reno_block_eigh",
    "reno_dav_workspace_bytes", "reno_dav_orth", "reno_dav_dots", "reno_dav_ritz", "reno_dav_precond",
    "reno_batch_scratch_bytes", "reno_batch_site_step", "reno_batch_normalize",
    "reno_autotune", "reno_autotune_reset", "reno_autotune_entries",
    "reno_probe_mfma16", "reno_prof_enable", "reno_prof_collect", "reno_prof_kind_info",
    "reno_prof_cancelled",
)

# demangled names of the GEMM kernel instantiations, indexed like
# reno_prof_collect's kinds
def gemm_kind_names():
    """Names (as profilers print them) of the GEMM kernel instantiations, in
    the order of reno_prof_collect's kinds."""
    l = lib()
    n = l.reno_prof_kind_info(-1, None)
    names = []
    info = (ctypes.c_int * 7)()
    for k in range(n):
        l.reno_prof_kind_info(k, info)
        names.append("void reno::%s<%s, %s, %s, %d, %d, %d>(reno::GemmCall)" % (
            "k_cgemm" if info[6] == 64 else "k_cgemm_s",
            "double" if info[0] == PREC_FP64 else "float", "true" if info[1] else "false",
            "true" if info[2] else "false", info[3], info[4], info[5]))
    return names


class TileCfg(ctypes.Structure):
    """reno_tile_cfg: one tuned GEMM stage."""
    _fields_ = [("M", ctypes.c_int), ("N", ctypes.c_int), ("K", ctypes.c_int),
                ("batch", ctypes.c_int), ("prec", ctypes.c_int), ("tile", ctypes.c_int),
                ("nsplit", ctypes.c_int), ("planned_tile", ctypes.c_int),
                ("planned_nsplit", ctypes.c_int), ("us_tuned", ctypes.c_float),
                ("us_planned", ctypes.c_float)]


class RenoNativeError(RuntimeError):
    """Raised when the native library is unavailable or reports a failure."""


class HopDesc(ctypes.Structure):
    _fields_ = [("nsite", ctypes.c_int), ("prec", ctypes.c_int),
                ("Ml", ctypes.c_int), ("d", ctypes.c_int), ("Mr", ctypes.c_int),
                ("Dl", ctypes.c_int), ("Dr", ctypes.c_int),
                ("L", ctypes.c_void_p), ("W", ctypes.c_void_p), ("R", ctypes.c_void_p)]


class QnBlock(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int), ("ncols", ctypes.c_int),
                ("rows", ctypes.c_void_p), ("cols", ctypes.c_void_p)]


_lib = None
_lib_lock = threading.Lock()


def load_library(path: str = LIB_PATH):
    """Load the shared library and declare the C signatures (no GPU needed)."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RenoNativeError(
                f"{path} not found: build it with `python -m renormalizer_amd.csrc.build`")
        lib = ctypes.CDLL(path)
        c_int, c_size, vp, c_double = ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_double
        lib.reno_version.restype = c_int
        lib.reno_status_string.restype = ctypes.c_char_p
        lib.reno_status_string.argtypes = [c_int]
        lib.reno_last_error.restype = ctypes.c_char_p
        lib.reno_hop_workspace_bytes.restype = c_size
        lib.reno_hop_workspace_bytes.argtypes = [ctypes.POINTER(HopDesc)]
        lib.reno_krylov_workspace_bytes.restype = c_size
        lib.reno_krylov_workspace_bytes.argtypes = [ctypes.POINTER(HopDesc), c_int]
        lib.reno_env_workspace_bytes.restype = c_size
        lib.reno_env_workspace_bytes.argtypes = [c_int] * 6
        lib.reno_qr_workspace_bytes.restype = c_size
        lib.reno_qr_workspace_bytes.argtypes = [c_int, c_int, c_int, ctypes.POINTER(QnBlock)]
        lib.reno_hop_apply.restype = c_int
        lib.reno_hop_apply.argtypes = [ctypes.POINTER(HopDesc), vp, vp, vp, c_size, vp]
        lib.reno_hop_apply_noalloc.restype = c_int
        lib.reno_hop_apply_noalloc.argtypes = [ctypes.POINTER(HopDesc), vp, vp, vp]
        lib.reno_env_update.restype = c_int
        lib.reno_env_update.argtypes = [c_int] * 7 + [vp, vp, vp, vp, vp, vp, c_size, vp]
        lib.reno_expm_krylov.restype = c_int
        lib.reno_expm_krylov.argtypes = [ctypes.POINTER(HopDesc), c_double, c_double, vp, vp,
                                         ctypes.POINTER(c_int), ctypes.POINTER(c_size),
                                         vp, c_size, vp]
        lib.reno_expm_krylov_async.restype = c_int
        lib.reno_expm_krylov_async.argtypes = [ctypes.POINTER(HopDesc), c_double, c_double, vp, vp,
                                               c_int, vp, vp, c_size, vp]
        lib.reno_expm_krylov_mixed_async.restype = c_int
        lib.reno_expm_krylov_mixed_async.argtypes = [ctypes.POINTER(HopDesc), ctypes.POINTER(HopDesc),
                                                     c_int, c_double, c_double, c_double, vp, vp,
                                                     c_int, vp, vp, c_size, vp]
        lib.reno_heff_diag.restype = c_int
        lib.reno_heff_diag.argtypes = [c_int] * 5 + [vp, vp, vp, vp, vp]
        lib.reno_heff_dense.restype = c_int
        lib.reno_heff_dense.argtypes = [c_int] * 5 + [vp, vp, vp, vp, c_int, vp, vp]
        lib.reno_block_qr.restype = c_int
        lib.reno_block_qr.argtypes = [c_int, c_int, c_int, vp, c_int, ctypes.POINTER(QnBlock),
                                      vp, vp, c_int, vp, c_size, vp]
        lib.reno_svd_workspace_bytes.restype = c_size
        lib.reno_svd_workspace_bytes.argtypes = [c_int, c_int, c_int, ctypes.POINTER(QnBlock)]
        lib.reno_block_svd.restype = c_int
        lib.reno_block_svd.argtypes = [c_int, c_int, vp, c_int, ctypes.POINTER(QnBlock), vp, vp, vp,
                                       vp, c_size, vp]
        lib.reno_eigh_workspace_bytes.restype = c_size
        lib.reno_eigh_workspace_bytes.argtypes = [c_int, c_int, ctypes.POINTER(QnBlock)]
        lib.reno_block_eigh.restype = c_int
        lib.reno_block_eigh.argtypes = [c_int, vp, c_int, ctypes.POINTER(QnBlock), vp, vp, vp, c_size, vp]
        c_i64 = ctypes.c_int64
        lib.reno_dav_workspace_bytes.restype = c_size
        lib.reno_dav_workspace_bytes.argtypes = [c_i64, c_int]
        lib.reno_dav_orth.restype = c_int
        lib.reno_dav_orth.argtypes = [c_i64, c_int, vp, c_i64, vp, c_double, vp, vp, c_size, vp]
        lib.reno_dav_dots.restype = c_int
        lib.reno_dav_dots.argtypes = [c_i64, c_int, vp, c_i64, vp, vp, vp, c_size, vp]
        lib.reno_dav_ritz.restype = c_int
        lib.reno_dav_ritz.argtypes = [c_i64, c_int, c_int, vp, vp, c_i64, vp, vp, vp, vp, vp, vp, vp, c_size, vp]
        lib.reno_dav_precond.restype = c_int
        lib.reno_dav_precond.argtypes = [c_i64, vp, vp, vp, c_double, vp, vp]
        lib.reno_env_close.restype = c_int
        lib.reno_env_close.argtypes = [c_int, c_int, c_int, c_int, vp, vp, vp, vp]
        lib.reno_zgemm.restype = c_int
        lib.reno_zgemm.argtypes = [c_int, c_int, c_int, vp, vp, vp, vp]
        lib.reno_batch_scratch_bytes.restype = c_size
        lib.reno_batch_scratch_bytes.argtypes = [c_int] * 6
        lib.reno_batch_site_step.restype = c_int
        lib.reno_batch_site_step.argtypes = [c_int, vp, c_double, vp]
        lib.reno_batch_normalize.restype = c_int
        lib.reno_batch_normalize.argtypes = [c_int, vp, vp, vp, vp]
        lib.reno_autotune.restype = c_int
        lib.reno_autotune.argtypes = [ctypes.POINTER(HopDesc), vp, vp, vp, c_size, c_int,
                                      ctypes.POINTER(TileCfg), c_int, ctypes.POINTER(c_int), vp]
        lib.reno_autotune_reset.restype = c_int
        lib.reno_autotune_reset.argtypes = []
        lib.reno_autotune_entries.restype = c_int
        lib.reno_autotune_entries.argtypes = []
        lib.reno_probe_mfma16.restype = c_int
        lib.reno_probe_mfma16.argtypes = [c_int, vp, vp, vp, vp]
        lib.reno_prof_enable.restype = c_int
        lib.reno_prof_enable.argtypes = [c_int]
        lib.reno_prof_kind_info.restype = c_int
        lib.reno_prof_kind_info.argtypes = [c_int, ctypes.POINTER(c_int)]
        lib.reno_prof_cancelled.restype = c_int
        lib.reno_prof_cancelled.argtypes = [c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_int)]
        lib.reno_prof_collect.restype = c_int
        lib.reno_prof_collect.argtypes = [c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_double),
                                          ctypes.POINTER(c_int), ctypes.POINTER(c_double),
                                          ctypes.POINTER(c_int), ctypes.POINTER(c_int)]
        _lib = lib
        return lib


def lib():
    return load_library()


def _raise(status: int, what: str):
    l = lib()
    msg = f"{what}: {l.reno_status_string(status).decode()} ({l.reno_last_error().decode()})"
    if status == RENO_EQN:
        raise ValueError("Invalid quantum number. " + msg)
    if status == RENO_EINVAL:
        raise ValueError(msg)
    raise RenoNativeError(msg)


def check(status: int, what: str):
    if status != RENO_OK:
        _raise(status, what)


def _require_device(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise RenoNativeError("libreno_hip operates on device tensors only; got a CPU tensor")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class _Workspace:
    """Grow-only device scratch, one buffer per (device, stream).

    The Krylov state block, the basis V and the split-K partials live in
    this buffer, so two streams must never share one: each stream gets its
    own.  A buffer that grows is released to torch's caching allocator,
    which hands its memory out again only to allocations on the same
    stream, i.e. after the kernels already queued there that read it."""

    def __init__(self):
        self._buf = {}
        self._lock = threading.Lock()

    def get(self, nbytes: int, device, stream=None):
        dev = device.index if device.index is not None else torch.cuda.current_device()
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        key = (dev, int(stream))
        with self._lock:
            buf = self._buf.get(key)
            if buf is None or buf.numel() < nbytes:
                size = max(int(nbytes), 1 << 20)
                if buf is not None:
                    size = max(size, int(buf.numel() * 1.25))
                buf = torch.empty(size, dtype=torch.uint8, device=device)
                self._buf[key] = buf
            return buf

    def keys(self):
        with self._lock:
            return list(self._buf)


workspace = _Workspace()


def prec_dtypes(prec: int):
    """(environment dtype, MPO dtype) for a precision mode."""
    if prec == PREC_FP64:
        return torch.complex128, torch.float64
    return torch.complex64, torch.float32


def _hop_desc(nsite, prec, L, W, R, Ml, d, Mr, Dl, Dr):
    return HopDesc(nsite, prec, Ml, d, Mr, Dl, Dr, _ptr(L).value, _ptr(W).value if W is not None else None,
                   _ptr(R).value)


def _check_env_dtype(prec, *ts):
    cdt, rdt = prec_dtypes(prec)
    for t in ts:
        if t is None:
            continue
        want = rdt if not t.is_complex() else cdt
        if t.dtype != want:
            raise ValueError(f"tensor dtype {t.dtype} does not match precision mode {prec} ({want})")
        if not t.is_contiguous():
            raise ValueError("tensors must be contiguous")


def hop_apply(L, W, R, c, prec=PREC_FP64):
    """One effective-Hamiltonian matvec (reference hop_expr(...)(c)).

    L (Ml, Dl, Ml), R (Mr, Dr, Mr), W (Dl, d, d, Dr) for a site vector
    c (Ml, d, Mr); W is None for the bond (0-site) operator c (Ml, Mr).
    """
    _require_device(L, R, c, W)
    _check_env_dtype(prec, L, R, W)
    if c.dtype != torch.complex128 or not c.is_contiguous():
        raise ValueError("vector must be contiguous complex128")
    if W is None:
        Ml, Mr = c.shape
        D = L.shape[1]
        assert L.shape == (Ml, D, Ml) and R.shape == (Mr, D, Mr), "hop0 shape mismatch"
        desc = _hop_desc(0, prec, L, None, R, Ml, 1, Mr, D, D)
    else:
        Ml, d, Mr = c.shape
        Dl, Dr = W.shape[0], W.shape[3]
        assert L.shape == (Ml, Dl, Ml) and R.shape == (Mr, Dr, Mr), "hop1 shape mismatch"
        assert W.shape == (Dl, d, d, Dr), "hop1 MPO shape mismatch"
        desc = _hop_desc(1, prec, L, W, R, Ml, d, Mr, Dl, Dr)
    l = lib()
    nbytes = l.reno_hop_workspace_bytes(ctypes.byref(desc))
    ws = workspace.get(nbytes, c.device)
    out = torch.empty_like(c)
    check(l.reno_hop_apply(ctypes.byref(desc), _ptr(c), _ptr(out), _ptr(ws), nbytes, _stream()),
          "reno_hop_apply")
    return out


_TUNED = set()
_TUNED_LOCK = threading.Lock()
# every stage record the tuner returned in this process (bench reports them)
tuned_records = []


def autotune(L, W, R, c, prec=PREC_FP64, reps=8):
    """Tune the GEMM stages of the matvec hop_apply(L, W, R, c, prec) on the
    device (reno_autotune): every tile / split-K candidate is timed on the
    current stream and the fastest is used by all later launches of the same
    GEMM signatures.  Operators already tuned in this process return at once.
    Returns the records of the stages tuned by this call."""
    _require_device(L, R, c, W)
    _check_env_dtype(prec, L, R, W)
    if W is None:
        Ml, Mr = c.shape
        D = L.shape[1]
        desc = _hop_desc(0, prec, L, None, R, Ml, 1, Mr, D, D)
    else:
        Ml, d, Mr = c.shape
        desc = _hop_desc(1, prec, L, W, R, Ml, d, Mr, W.shape[0], W.shape[3])
    key = (desc.nsite, prec, desc.Ml, desc.d, desc.Mr, desc.Dl, desc.Dr, c.device.index)
    with _TUNED_LOCK:
        if key in _TUNED:
            return []
    l = lib()
    nbytes = l.reno_hop_workspace_bytes(ctypes.byref(desc))
    ws = workspace.get(nbytes, c.device)
    out = torch.empty_like(c)
    recs = (TileCfg * 4)()
    n = ctypes.c_int(0)
    check(l.reno_autotune(ctypes.byref(desc), _ptr(c), _ptr(out), _ptr(ws), nbytes, int(reps),
                          recs, 4, ctypes.byref(n), _stream()), "reno_autotune")
    got = [{f: getattr(recs[i], f) for f, _ in TileCfg._fields_} for i in range(n.value)]
    with _TUNED_LOCK:
        _TUNED.add(key)
        tuned_records.extend(got)
    return got


def autotune_reset():
    """Forget every tuned choice (library table and this module's record)."""
    check(lib().reno_autotune_reset(), "reno_autotune_reset")
    with _TUNED_LOCK:
        _TUNED.clear()
        tuned_records.clear()


def autotune_entries():
    return int(lib().reno_autotune_entries())


def env_update(env, ms, W, domain, ms_conj=None, prec=PREC_FP64):
    """contract_one_site (reference lib.py:169-250) on the device."""
    _require_device(env, ms, W, ms_conj)
    _check_env_dtype(prec, env, W)
    if ms.dtype != torch.complex128 or not ms.is_contiguous():
        raise ValueError("site tensor must be contiguous complex128")
    Ml, d, Mr = ms.shape
    Dl, Dr = W.shape[0], W.shape[3]
    assert W.shape == (Dl, d, d, Dr)
    cdt, _ = prec_dtypes(prec)
    if domain == "L":
        assert env.shape == (Ml, Dl, Ml), f"L environment shape {tuple(env.shape)} vs site {tuple(ms.shape)}"
        out = torch.empty((Mr, Dr, Mr), dtype=cdt, device=ms.device)
        dom = 0
    elif domain == "R":
        assert env.shape == (Mr, Dr, Mr), f"R environment shape {tuple(env.shape)} vs site {tuple(ms.shape)}"
        out = torch.empty((Ml, Dl, Ml), dtype=cdt, device=ms.device)
        dom = 1
    else:
        raise ValueError(f"domain {domain}")
    if ms_conj is not None:
        assert ms_conj.shape == ms.shape and ms_conj.dtype == torch.complex128
    l = lib()
    nbytes = l.reno_env_workspace_bytes(prec, Ml, d, Mr, Dl, Dr)
    ws = workspace.get(nbytes, ms.device)
    check(l.reno_env_update(dom, prec, Ml, d, Mr, Dl, Dr, _ptr(env), _ptr(ms), _ptr(ms_conj),
                            _ptr(W), _ptr(out), _ptr(ws), nbytes, _stream()),
          "reno_env_update")
    return out


_KRYLOV_KCAP = 20


def _krylov_desc(L, W, R, v, prec):
    _require_device(L, R, v, W)
    _check_env_dtype(prec, L, R, W)
    if v.dtype != torch.complex128 or not v.is_contiguous():
        raise ValueError("vector must be contiguous complex128")
    if W is None:
        Ml, Mr = v.shape
        D = L.shape[1]
        assert L.shape == (Ml, D, Ml) and R.shape == (Mr, D, Mr), "hop0 shape mismatch"
        return _hop_desc(0, prec, L, None, R, Ml, 1, Mr, D, D)
    Ml, d, Mr = v.shape
    Dl, Dr = W.shape[0], W.shape[3]
    assert L.shape == (Ml, Dl, Ml) and R.shape == (Mr, Dr, Mr), "hop1 shape mismatch"
    return _hop_desc(1, prec, L, W, R, Ml, d, Mr, Dl, Dr)


def expm_krylov(L, W, R, v, dt, prec=PREC_FP64):
    """expm(dt * H_eff) v by Lanczos (reference lib/krylov/krylov.py:27-84).

    Returns (vector, number of matvecs) like the reference.  Waits for the
    device once per chunk of iterations (normally once per call).
    """
    desc = _krylov_desc(L, W, R, v, prec)
    dt = complex(dt)
    l = lib()
    return _expm_krylov_sync(desc, dt, v, l.reno_krylov_workspace_bytes, l.reno_expm_krylov,
                             lambda nbytes: workspace.get(nbytes, v.device), _stream())


def _expm_krylov_sync(desc, dt, v, ws_bytes_fn, call_fn, get_ws, stream):
    """The retry loop of the synchronous Krylov call: on RENO_ENOSPACE the
    workspace grows to the size the library asks for (twice the Krylov
    vectors) until KRYLOV_MAX vectors; a call that has not converged by then
    raises instead of retrying forever."""
    kcap = _KRYLOV_KCAP
    out = torch.empty_like(v)
    nmv = ctypes.c_int(0)
    needed = ctypes.c_size_t(0)
    while True:
        nbytes = ws_bytes_fn(ctypes.byref(desc), kcap)
        if nbytes == 0:
            raise RenoNativeError(f"expm_krylov: no workspace size for {kcap} Krylov vectors")
        nbytes = max(int(nbytes), int(needed.value))
        ws = get_ws(nbytes)
        st = call_fn(ctypes.byref(desc), dt.real, dt.imag, _ptr(v), _ptr(out),
                     ctypes.byref(nmv), ctypes.byref(needed), _ptr(ws), ws.numel(), stream)
        if st == RENO_ENOSPACE:
            if kcap >= KRYLOV_MAX:
                raise RenoNativeError(
                    f"expm_krylov: no convergence within {KRYLOV_MAX} Lanczos vectors "
                    f"(non-finite environment or vector?)")
            kcap = min(2 * kcap, KRYLOV_MAX)
            continue
        check(st, "reno_expm_krylov")
        return out, nmv.value


# largest Krylov space the library supports (kMaxKrylov in krylov.hip)
KRYLOV_MAX = 1024


# largest Krylov space of one asynchronous call
KRYLOV_ASYNC_MAX = 1024


def expm_krylov_async(L, W, R, v, dt, prec, max_matvec, info):
    """Queue expm(dt * H_eff) v without waiting for the device.

    At most ``max_matvec`` Lanczos iterations are queued.  ``info`` is a
    device int32 tensor of two elements that receives (matvecs used, flag):
    flag 1 when the reference's stopping rule ended the iteration, 0 when
    the budget ran out first (the returned vector is then not the result and
    the caller must redo the call with ``expm_krylov``), -1 for a zero v.
    """
    desc = _krylov_desc(L, W, R, v, prec)
    if info.dtype != torch.int32 or info.numel() < 2 or not info.is_cuda:
        raise ValueError("info must be a device int32 tensor of two elements")
    max_matvec = int(max(2, min(max_matvec, KRYLOV_ASYNC_MAX)))
    dt = complex(dt)
    l = lib()
    out = torch.empty_like(v)
    nbytes = l.reno_krylov_workspace_bytes(ctypes.byref(desc), max_matvec)
    ws = workspace.get(nbytes, v.device)
    check(l.reno_expm_krylov_async(ctypes.byref(desc), dt.real, dt.imag, _ptr(v), _ptr(out),
                                   max_matvec, _ptr(info), _ptr(ws), ws.numel(), _stream()),
          "reno_expm_krylov_async")
    return out


# switch point meaning "every iteration in FP64"
SWITCH_NEVER = 1 << 30


def expm_krylov_mixed_async(L, W, R, L32, W32, R32, v, dt, max_matvec, j_switch, tau, info):
    """Queue expm(dt * H_eff) v with adaptive precision, without waiting.

    Lanczos iterations j < ``j_switch`` apply the FP64 operator (L, W, R:
    complex128 / float64), later ones the same operator held in FP32
    (L32, W32, R32: complex64 / float32).  ``info`` is a device int32 tensor
    of three elements: (matvecs used, flag) as for ``expm_krylov_async``, and
    the switch point measured on this call's result with threshold ``tau``
    (``SWITCH_NEVER`` when no iteration qualifies)."""
    desc = _krylov_desc(L, W, R, v, PREC_FP64)
    desc_lo = _krylov_desc(L32, W32, R32, v, PREC_MIXED)
    if info.dtype != torch.int32 or info.numel() < 3 or not info.is_cuda:
        raise ValueError("info must be a device int32 tensor of three elements")
    max_matvec = int(max(2, min(max_matvec, KRYLOV_ASYNC_MAX)))
    dt = complex(dt)
    l = lib()
    out = torch.empty_like(v)
    nbytes = l.reno_krylov_workspace_bytes(ctypes.byref(desc), max_matvec)
    ws = workspace.get(nbytes, v.device)
    check(l.reno_expm_krylov_mixed_async(ctypes.byref(desc), ctypes.byref(desc_lo),
                                         int(min(j_switch, SWITCH_NEVER)), float(tau),
                                         dt.real, dt.imag, _ptr(v), _ptr(out), max_matvec,
                                         _ptr(info), _ptr(ws), ws.numel(), _stream()),
          "reno_expm_krylov_mixed_async")
    return out


def _heff_args(L, W, R, shape):
    _require_device(L, W, R)
    Ml, d, Mr = shape
    Dl, Dr = W.shape[0], W.shape[3]
    if L.dtype != torch.complex128 or R.dtype != torch.complex128 or W.dtype != torch.float64:
        raise ValueError("H_eff helpers run in FP64: complex128 environments, float64 MPO")
    assert L.shape == (Ml, Dl, Ml) and R.shape == (Mr, Dr, Mr) and W.shape == (Dl, d, d, Dr), \
        "H_eff shape mismatch"
    for t in (L, W, R):
        if not t.is_contiguous():
            raise ValueError("H_eff operands must be contiguous")
    return Ml, d, Mr, Dl, Dr


def heff_diag(L, W, R, shape):
    """Diagonal of the one-site effective Hamiltonian, shape (Ml, d, Mr)
    (reference gs.py:424-436)."""
    Ml, d, Mr, Dl, Dr = _heff_args(L, W, R, shape)
    out = torch.empty((Ml, d, Mr), dtype=torch.complex128, device=L.device)
    check(lib().reno_heff_diag(Ml, d, Mr, Dl, Dr, _ptr(L), _ptr(W), _ptr(R), _ptr(out), _stream()),
          "reno_heff_diag")
    return out


def heff_dense(L, W, R, shape, idx):
    """Dense effective Hamiltonian on the flat site indices ``idx`` (device
    int64), shape (len(idx), len(idx)) (reference gs.py:302-330)."""
    Ml, d, Mr, Dl, Dr = _heff_args(L, W, R, shape)
    _require_device(idx)
    if idx.dtype != torch.int64 or not idx.is_contiguous():
        raise ValueError("idx must be a contiguous int64 device tensor")
    n = idx.numel()
    out = torch.empty((n, n), dtype=torch.complex128, device=L.device)
    check(lib().reno_heff_dense(Ml, d, Mr, Dl, Dr, _ptr(L), _ptr(W), _ptr(R), _ptr(idx), n, _ptr(out),
                                _stream()), "reno_heff_dense")
    return out


def block_qr(system, coef, blocks, v_transposed=False):
    """QN-blocked QR ("L") / RQ ("R") of a (rows x cols) complex128 matrix.

    ``blocks`` is a list of (row_index_array, col_index_array) numpy int64
    arrays in the reference's block order.  Returns (u, v) as in
    svd_qn(..., QR=True, full_matrices=False); with ``v_transposed`` the
    second factor is returned as v^T (k x cols).
    """
    _require_device(coef)
    if coef.dtype != torch.complex128 or coef.dim() != 2 or not coef.is_contiguous():
        raise ValueError("coefficient matrix must be a contiguous 2-D complex128 tensor")
    rows, cols = coef.shape
    if len(blocks) == 0:
        raise ValueError("Invalid quantum number")
    dev = coef.device
    cblocks, keep, kt = qn_block_descs(blocks, dev)
    u = torch.empty((rows, kt), dtype=torch.complex128, device=dev)
    v = torch.empty((kt, cols) if v_transposed else (cols, kt), dtype=torch.complex128, device=dev)
    l = lib()
    nbytes = l.reno_qr_workspace_bytes(rows, cols, len(blocks), cblocks)
    ws = workspace.get(nbytes, dev)
    sysc = {"L": 0, "R": 1}[system]
    check(l.reno_block_qr(sysc, rows, cols, _ptr(coef), len(blocks), cblocks, _ptr(u), _ptr(v),
                          int(bool(v_transposed)), _ptr(ws), nbytes, _stream()), "reno_block_qr")
    return u, v


def block_svd(coef, blocks):
    """QN-blocked SVD with complete bases (svd_qn's full_matrices branch) of
    a (rows x cols) complex128 device matrix.  ``blocks`` is a list of
    (row_index_array, col_index_array) in the reference's block order.
    Returns (u, v, s) on the device: u (rows x sum nrows), v (cols x sum
    ncols), first the singular columns of every block then the
    completions, and s the singular values of every block."""
    _require_device(coef)
    if coef.dtype != torch.complex128 or coef.dim() != 2 or not coef.is_contiguous():
        raise ValueError("coefficient matrix must be a contiguous 2-D complex128 tensor")
    rows, cols = coef.shape
    if len(blocks) == 0:
        raise ValueError("Invalid quantum number")
    dev = coef.device
    cblocks, keep, kt = qn_block_descs(blocks, dev)
    ut = sum(len(r) for r, _ in blocks)
    vt = sum(len(c) for _, c in blocks)
    u = torch.empty((rows, ut), dtype=torch.complex128, device=dev)
    v = torch.empty((cols, vt), dtype=torch.complex128, device=dev)
    s = torch.empty((kt,), dtype=torch.float64, device=dev)
    l = lib()
    nbytes = l.reno_svd_workspace_bytes(rows, cols, len(blocks), cblocks)
    ws = workspace.get(nbytes, dev)
    check(l.reno_block_svd(rows, cols, _ptr(coef), len(blocks), cblocks, _ptr(u), _ptr(v), _ptr(s),
                           _ptr(ws), nbytes, _stream()), "reno_block_svd")
    return u, v, s


def block_eigh(dm, sectors):
    """Eigen-decomposition of the Hermitian parts of the diagonal blocks
    ``sectors`` (row index arrays) of a complex128 (n x n) device matrix.
    Returns (u, w) on the device: u (n x sum len) with every block's
    eigenvectors at its rows, w the eigenvalues, ascending per block."""
    _require_device(dm)
    if dm.dtype != torch.complex128 or dm.dim() != 2 or dm.shape[0] != dm.shape[1] or not dm.is_contiguous():
        raise ValueError("density matrix must be a contiguous square complex128 tensor")
    if len(sectors) == 0:
        raise ValueError("Invalid quantum number")
    n = dm.shape[0]
    dev = dm.device
    cblocks, keep, _ = qn_block_descs([(r, r) for r in sectors], dev)
    ut = sum(len(r) for r in sectors)
    u = torch.empty((n, ut), dtype=torch.complex128, device=dev)
    w = torch.empty((ut,), dtype=torch.float64, device=dev)
    l = lib()
    nbytes = l.reno_eigh_workspace_bytes(n, len(sectors), cblocks)
    ws = workspace.get(nbytes, dev)
    check(l.reno_block_eigh(n, _ptr(dm), len(sectors), cblocks, _ptr(u), _ptr(w), _ptr(ws), nbytes,
                            _stream()), "reno_block_eigh")
    return u, w


def qn_block_descs(blocks, dev):
    """ctypes descriptors of the QN blocks with device index arrays.

    Returns (descriptors, keep, kt): ``keep`` holds every index tensor the
    descriptors point to, so that none is freed (and its memory handed to
    another tensor) before the call that reads them has been queued, even
    if the cache below evicts it meanwhile."""
    cblocks = (QnBlock * len(blocks))()
    keep = []
    kt = 0
    for i, (r, c) in enumerate(blocks):
        rt = _device_index(r, dev)
        ct = _device_index(c, dev)
        keep += [rt, ct]
        cblocks[i] = QnBlock(len(r), len(c), rt.data_ptr(), ct.data_ptr())
        kt += min(len(r), len(c))
    return cblocks, keep, kt


_INDEX_CACHE = collections.OrderedDict()
_INDEX_CACHE_MAX = 4096
_INDEX_LOCK = threading.Lock()


def _device_index(a, dev):
    """Device copy of a host int64 index array, cached by content and stream
    with least-recently-used eviction.  Callers keep the returned tensor alive for
    as long as a queued kernel may read it (an evicted tensor is freed when
    its last reference goes)."""
    a = np.ascontiguousarray(a, dtype=np.int64)
    # per stream: an evicted tensor returns to the allocator pool of the
    # stream it was made on, so only that stream's later work may reuse it
    stream = torch.cuda.current_stream(dev).cuda_stream if dev.type == "cuda" else 0
    key = (str(dev), int(stream), a.tobytes())
    with _INDEX_LOCK:
        t = _INDEX_CACHE.get(key)
        if t is not None:
            _INDEX_CACHE.move_to_end(key)
            return t
    t = torch.from_numpy(a).to(dev)
    with _INDEX_LOCK:
        _INDEX_CACHE[key] = t
        _INDEX_CACHE.move_to_end(key)
        while len(_INDEX_CACHE) > _INDEX_CACHE_MAX:
            _INDEX_CACHE.popitem(last=False)
    return t


def env_close(X, Y):
    """out[g] = sum_{f,h} X[f,g,h] Y[f,h] for X (F,G,H), Y (F,H) or (F,1,H)."""
    _require_device(X, Y)
    if X.dtype != Y.dtype or X.dtype not in (torch.complex128, torch.complex64):
        raise ValueError("env_close needs complex operands of one dtype")
    X = X.contiguous()
    Y = Y.contiguous()
    F, G, H = X.shape
    assert Y.numel() == F * H and Y.shape[0] == F and Y.shape[-1] == H, "env_close shape mismatch"
    prec = PREC_FP64 if X.dtype == torch.complex128 else PREC_MIXED
    out = torch.empty(G, dtype=torch.complex128, device=X.device)
    check(lib().reno_env_close(prec, F, G, H, _ptr(X), _ptr(Y), _ptr(out), _stream()), "reno_env_close")
    return out


def _dav_rows(V):
    if V.dtype != torch.complex128 or V.dim() != 2 or V.stride(1) != 1 or V.shape[0] > 64:
        raise ValueError("Davidson basis: a 2-D complex128 tensor of at most 64 contiguous rows")
    return V.shape[0], V.stride(0)


def _dav_vec(x, n):
    if x.dtype != torch.complex128 or not x.is_contiguous() or x.numel() != n:
        raise ValueError("Davidson vector: a contiguous complex128 tensor of the basis length")


def dav_orth(V, x, thresh):
    """Orthogonalise x (in place) against the rows of V twice, then normalise
    it when |x|^2 >= thresh.  Returns |x|^2 before normalisation (a host
    float: the caller's linear-dependence decision)."""
    _require_device(V, x)
    n = x.numel()
    k, ldv = _dav_rows(V) if V.shape[0] else (0, n)
    _dav_vec(x, n)
    if k:
        assert V.shape[1] == n, "Davidson basis length mismatch"
    nrm2 = torch.empty(1, dtype=torch.float64, device=x.device)
    nb = lib().reno_dav_workspace_bytes(n, k)
    ws = workspace.get(nb, x.device)
    check(lib().reno_dav_orth(n, k, _ptr(V) if k else None, ldv, _ptr(x), float(thresh), _ptr(nrm2), _ptr(ws),
                              nb, _stream()), "reno_dav_orth")
    return float(nrm2.item())


def dav_dots(V, y):
    """conj(V) @ y for the rows of V (device complex128 vector)."""
    _require_device(V, y)
    k, ldv = _dav_rows(V)
    n = V.shape[1]
    _dav_vec(y, n)
    out = torch.empty(k, dtype=torch.complex128, device=y.device)
    nb = lib().reno_dav_workspace_bytes(n, k)
    ws = workspace.get(nb, y.device)
    check(lib().reno_dav_dots(n, k, _ptr(V), ldv, _ptr(y), _ptr(out), _ptr(ws), nb, _stream()), "reno_dav_dots")
    return out


def dav_ritz(V, AV, U, e):
    """Ritz vectors X = U^T V, AX = U^T AV, residuals R = AX - e X (k x n
    each) and |R_j|^2 (host array) for U (m x k) and e (k) from the
    subspace eigenproblem."""
    _require_device(V, AV)
    m, ldv = _dav_rows(V)
    m2, ldv2 = _dav_rows(AV)
    assert m == m2 and ldv == ldv2 and V.shape[1] == AV.shape[1], "V / AV mismatch"
    n = V.shape[1]
    k = U.shape[1]
    dev = V.device
    Ud = torch.from_numpy(np.ascontiguousarray(U, dtype=np.complex128)).to(dev)
    ed = torch.from_numpy(np.ascontiguousarray(e, dtype=np.float64)).to(dev)
    X = torch.empty((k, n), dtype=torch.complex128, device=dev)
    AX = torch.empty_like(X)
    R = torch.empty_like(X)
    rn2 = torch.empty(k, dtype=torch.float64, device=dev)
    nb = lib().reno_dav_workspace_bytes(n, k)
    ws = workspace.get(nb, dev)
    check(lib().reno_dav_ritz(n, m, k, _ptr(V), _ptr(AV), ldv, _ptr(Ud), _ptr(ed), _ptr(X), _ptr(AX), _ptr(R),
                              _ptr(rn2), _ptr(ws), nb, _stream()), "reno_dav_ritz")
    return X, AX, R, rn2.cpu().numpy()


def dav_precond(r, hdiag, mask, e):
    """r / (hdiag - e + 1e-4) on the entries where ``mask`` (bool) is set."""
    _require_device(r, hdiag, mask)
    n = r.numel()
    _dav_vec(r, n)
    if hdiag.dtype != torch.float64 or hdiag.numel() != n or mask.dtype != torch.bool or mask.numel() != n:
        raise ValueError("preconditioner: float64 diagonal and bool mask of the vector length")
    out = torch.empty_like(r)
    check(lib().reno_dav_precond(n, _ptr(r), _ptr(hdiag.contiguous()), _ptr(mask.contiguous()), float(e),
                                 _ptr(out), _stream()), "reno_dav_precond")
    return out


def zgemm(A, B):
    """C = A @ B for 2-D contiguous complex128 device matrices."""
    _require_device(A, B)
    if A.dtype != torch.complex128 or B.dtype != torch.complex128:
        raise ValueError("zgemm needs complex128 operands")
    A = A.contiguous()
    B = B.contiguous()
    M, K = A.shape
    K2, N = B.shape
    assert K == K2, "zgemm inner dimension mismatch"
    C = torch.empty((M, N), dtype=torch.complex128, device=A.device)
    check(lib().reno_zgemm(M, N, K, _ptr(A), _ptr(B), _ptr(C), _stream()), "reno_zgemm")
    return C


def prof_enable(period: int):
    """Start (period >= 1: time a pseudo-random 1-in-period sample of the
    GEMM launches with HIP events), stop (0) or resume without clearing
    (-period) the GEMM launch record."""
    check(lib().reno_prof_enable(int(period)), "reno_prof_enable")


def prof_collect():
    """{kernel name: {"timed", "timed_ms", "timed_flop": executed launches of
    the event-timed sample, their GPU ms and algorithmic FLOP; "executed",
    "flop": all executed launches and their FLOP; "queued": all launches,
    including those cancelled on the device; "timed_cancelled",
    "timed_cancelled_ms": event-timed launches that were cancelled on the
    device and their summed event time}}."""
    names = gemm_kind_names()
    n = len(names)
    ms = (ctypes.c_double * n)()
    flt = (ctypes.c_double * n)()
    tm = (ctypes.c_int * n)()
    fl = (ctypes.c_double * n)()
    ex = (ctypes.c_int * n)()
    qd = (ctypes.c_int * n)()
    check(lib().reno_prof_collect(n, ms, flt, tm, fl, ex, qd), "reno_prof_collect")
    cms = (ctypes.c_double * n)()
    cct = (ctypes.c_int * n)()
    check(lib().reno_prof_cancelled(n, cms, cct), "reno_prof_cancelled")
    return {names[k]: {"timed": tm[k], "timed_ms": ms[k], "timed_flop": flt[k], "executed": ex[k],
                       "flop": fl[k], "queued": qd[k], "timed_cancelled": cct[k],
                       "timed_cancelled_ms": cms[k]}
            for k in range(n) if qd[k]}


def probe_mfma16(a, b, prec):
    """Run one 16x16x4 MFMA on row-major a (16x4), b (4x16)."""
    _require_device(a, b)
    c = torch.zeros((16, 16), dtype=a.dtype, device=a.device)
    check(lib().reno_probe_mfma16(prec, _ptr(a), _ptr(b), _ptr(c), _stream()), "reno_probe_mfma16")
    return c
