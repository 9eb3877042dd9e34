"""ctypes binding of ``libdcx.so`` (the C ABI declared in ``include/dcx.h``).

The product path: every call goes to the HIP engine.  There is no CPU fallback -- if the library or
a GPU is missing, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdcx.so")

DCX_OK = 0
DCX_E_INVALID_X = -1
DCX_E_ALPHABET = -2
DCX_E_NOMEM = -3
DCX_E_HIP = -4
DCX_E_IO = -5
DCX_E_ARG = -6
DCX_E_WIDTH = -7
DCX_E_ADMISSION = -8
DCX_E_UNSUPPORTED = -9
DCX_E_CANCELLED = -10
DCX_E_NO_GPU = -11
MAX_LEVELS = 32

PHASES = ["phase_00_preparation", "phase_01_02_samples_ranks", "phase_03_recursion", "phase_04_merge", "output"]

SUPPORTED_X = (3, 7, 13, 21, 31, 39, 57, 73, 91, 95, 133)


class DCXError(RuntimeError):
    def __init__(self, code, message=""):
        self.code = code
        super().__init__(f"DCX error {code}: {message}")


class Options(ctypes.Structure):
    _fields_ = [
        ("X", ctypes.c_uint32),
        ("pack_samples", ctypes.c_uint32),
        ("pack_merging", ctypes.c_uint32),
        ("text_on_device", ctypes.c_uint32),
        ("sa_on_device", ctypes.c_uint32),
        ("profile_kernels", ctypes.c_uint32),
        ("discarding_threshold", ctypes.c_double),
        ("base_case_chars", ctypes.c_uint64),
        ("seed", ctypes.c_uint64),
        ("batch_max_chars", ctypes.c_uint64),
        ("reserved", ctypes.c_uint64 * 5),
    ]


class LevelStats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint64),
        ("samples", ctypes.c_uint64),
        ("names", ctypes.c_uint64),
        ("reduced", ctypes.c_uint64),
        ("base_case", ctypes.c_uint32),
        ("recursed", ctypes.c_uint32),
        ("discarding", ctypes.c_uint32),
        ("key_chars", ctypes.c_uint32),
        ("merge_chars", ctypes.c_uint32),
        ("bits_per_char", ctypes.c_uint32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("levels", ctypes.c_uint32),
        ("sigma", ctypes.c_uint32),
        ("level", LevelStats * MAX_LEVELS),
        ("ms_total", ctypes.c_double),
        ("ms_device", ctypes.c_double),
        ("ms_phase", ctypes.c_double * 8),
        ("alg_bytes", ctypes.c_uint64),
        ("peak_device_bytes", ctypes.c_uint64),
        ("sort_passes", ctypes.c_uint64),
        ("sort_pass_items", ctypes.c_uint64),
        ("sort_pass_ms", ctypes.c_double),
        ("sort_pass_bytes", ctypes.c_uint64),
    ]

    def to_dict(self):
        lv = []
        for i in range(self.levels):
            l = self.level[i]
            lv.append({f: int(getattr(l, f)) for f, _ in LevelStats._fields_})
        return {
            "levels": lv,
            "sigma": int(self.sigma),
            "ms_total": float(self.ms_total),
            "ms_device": float(self.ms_device),
            "ms_phase": {PHASES[i]: float(self.ms_phase[i]) for i in range(len(PHASES))},
            "alg_bytes": int(self.alg_bytes),
            "peak_device_bytes": int(self.peak_device_bytes),
            "sort_passes": int(self.sort_passes),
            "sort_pass_items": int(self.sort_pass_items),
            "sort_pass_ms": float(self.sort_pass_ms),
            "sort_pass_bytes": int(self.sort_pass_bytes),
        }


_lib = None
_lib_lock = threading.Lock()

EXPORTED = ["dcx_default_options", "dcx_create", "dcx_destroy", "dcx_build_sa", "dcx_build_sa_batch",
            "dcx_build_sa_noalloc", "dcx_check_sa", "dcx_strerror", "dcx_last_error"]

# dcx_build_sa_noalloc: longest text, and the device array per GPU it builds in (include/dcx.h)
NOALLOC_MAX_CHARS = 1 << 20
NOALLOC_SCRATCH_BYTES = 128 << 20


def load_library(path: str = LIB_PATH):
    """Loads libdcx.so and declares the C ABI.  Raises if the HIP build is missing."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build the HIP engine first (python -m "
                              "distributed_suffix_sorting_evaluation_amd.build)")
        # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64 (same soname).  Loading torch
        # first makes libdcx bind to that copy; loading the system copy first would leave torch without GPUs.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.dcx_default_options.argtypes = [ctypes.POINTER(Options)]
        L.dcx_default_options.restype = None
        L.dcx_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_uint64]
        L.dcx_create.restype = ctypes.c_int
        L.dcx_destroy.argtypes = [vp]
        L.dcx_destroy.restype = ctypes.c_int
        L.dcx_build_sa.argtypes = [vp, vp, ctypes.c_uint64, ctypes.POINTER(Options), vp, ctypes.c_uint32,
                                   ctypes.POINTER(Stats)]
        L.dcx_build_sa.restype = ctypes.c_int
        L.dcx_build_sa_batch.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32,
                                         ctypes.POINTER(Options), vp, ctypes.POINTER(Stats)]
        L.dcx_build_sa_batch.restype = ctypes.c_int
        L.dcx_build_sa_noalloc.argtypes = [vp, vp, ctypes.c_uint64, ctypes.POINTER(Options), vp, ctypes.c_uint32,
                                           ctypes.POINTER(Stats)]
        L.dcx_build_sa_noalloc.restype = ctypes.c_int
        L.dcx_check_sa.argtypes = [vp, vp, ctypes.c_uint64, vp, ctypes.c_uint32, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int)]
        L.dcx_check_sa.restype = ctypes.c_int
        L.dcx_strerror.argtypes = [ctypes.c_int]
        L.dcx_strerror.restype = ctypes.c_char_p
        L.dcx_last_error.argtypes = [vp]
        L.dcx_last_error.restype = ctypes.c_char_p
        _lib = L
        return L


def default_options() -> Options:
    o = Options()
    load_library().dcx_default_options(ctypes.byref(o))
    return o


def _as_u8(text) -> np.ndarray:
    if isinstance(text, np.ndarray):
        return np.ascontiguousarray(text.reshape(-1).view(np.uint8) if text.dtype != np.uint8 else text.reshape(-1))
    if isinstance(text, str):
        text = text.encode("latin-1")
    return np.frombuffer(bytes(text), dtype=np.uint8)


_WIDTH_DTYPE = {4: np.uint32, 8: np.uint64}


class Engine:
    """One engine handle: a HIP device, a stream and the device memory it caches."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self._L = load_library()
        h = ctypes.c_void_p()
        rc = self._L.dcx_create(ctypes.byref(h), device, hbm_budget_bytes)
        if rc != DCX_OK:
            raise DCXError(rc, self._L.dcx_strerror(rc).decode())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.dcx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _raise(self, rc):
        detail = self._L.dcx_last_error(self._h).decode(errors="replace")
        raise DCXError(rc, f"{self._L.dcx_strerror(rc).decode()}: {detail}")

    def make_options(self, X=39, pack_samples=True, pack_merging=True, discarding_threshold=0.7,
                     base_case_chars=10000, seed=0, profile_kernels=False) -> Options:
        o = default_options()
        o.X = int(X)
        o.pack_samples = int(bool(pack_samples))
        o.pack_merging = int(bool(pack_merging))
        o.discarding_threshold = float(discarding_threshold)
        o.base_case_chars = int(base_case_chars)
        o.seed = int(seed)
        o.profile_kernels = int(bool(profile_kernels))
        return o

    def build_sa(self, text, sa_width: int = 8, stats: bool = False, **kw):
        """SA of a host text; returns numpy (uint32 / uint64, or uint8 bytes of uint40 for width 5)."""
        t = _as_u8(text)
        n = t.size
        o = self.make_options(**kw)
        if sa_width == 5:
            out = np.zeros(max(5 * n, 1), dtype=np.uint8)
        else:
            out = np.zeros(max(n, 1), dtype=_WIDTH_DTYPE[sa_width])
        st = Stats()
        rc = self._L.dcx_build_sa(self._h, t.ctypes.data if n else None, n, ctypes.byref(o),
                                  out.ctypes.data if n else None, sa_width, ctypes.byref(st))
        if rc != DCX_OK:
            self._raise(rc)
        out = out[: 5 * n] if sa_width == 5 else out[:n]
        return (out, st.to_dict()) if stats else out

    def build_sa_device(self, text_ptr: int, n: int, sa_ptr: int, sa_width: int = 8, **kw) -> dict:
        """SA of a device-resident text into a device buffer (both raw HIP pointers)."""
        o = self.make_options(**kw)
        o.text_on_device = 1
        o.sa_on_device = 1
        st = Stats()
        rc = self._L.dcx_build_sa(self._h, ctypes.c_void_p(text_ptr), n, ctypes.byref(o), ctypes.c_void_p(sa_ptr),
                                  sa_width, ctypes.byref(st))
        if rc != DCX_OK:
            self._raise(rc)
        return st.to_dict()

    def build_sa_noalloc(self, text_ptr: int, n: int, sa_ptr: int, sa_width: int = 8, **kw) -> dict:
        """dcx_build_sa_noalloc: SA of a device-resident text (at most NOALLOC_MAX_CHARS characters) into a
        device buffer, with every device buffer of the build carved from the library's per-GPU array."""
        o = self.make_options(**kw)
        st = Stats()
        rc = self._L.dcx_build_sa_noalloc(self._h, ctypes.c_void_p(text_ptr), n, ctypes.byref(o),
                                          ctypes.c_void_p(sa_ptr), sa_width, ctypes.byref(st))
        if rc != DCX_OK:
            self._raise(rc)
        return st.to_dict()

    def build_sa_batch(self, texts, stats: bool = False, batch_max_chars: int = 0, **kw):
        """Per-text 32-bit SAs of a list of texts in one batched call; returns a list of uint32 arrays."""
        arrs = [_as_u8(t) for t in texts]
        offsets = np.zeros(len(arrs) + 1, dtype=np.uint64)
        if arrs:
            offsets[1:] = np.cumsum([a.size for a in arrs])
        joined = np.concatenate(arrs) if arrs and offsets[-1] else np.zeros(1, dtype=np.uint8)
        out = np.zeros(max(int(offsets[-1]), 1), dtype=np.uint32)
        o = self.make_options(**kw)
        o.batch_max_chars = int(batch_max_chars)
        st = Stats()
        rc = self._L.dcx_build_sa_batch(self._h, joined.ctypes.data, offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                        len(arrs), ctypes.byref(o), out.ctypes.data, ctypes.byref(st))
        if rc != DCX_OK:
            self._raise(rc)
        res = [out[int(offsets[k]):int(offsets[k + 1])] for k in range(len(arrs))]
        return (res, st.to_dict()) if stats else res

    def build_sa_batch_device(self, texts_ptr: int, offsets: np.ndarray, sa_ptr: int, batch_max_chars: int = 0,
                              **kw) -> dict:
        """Batched SAs of device-resident joined texts (offsets on the host) into a device uint32 buffer."""
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        o = self.make_options(**kw)
        o.text_on_device = 1
        o.sa_on_device = 1
        o.batch_max_chars = int(batch_max_chars)
        st = Stats()
        rc = self._L.dcx_build_sa_batch(self._h, ctypes.c_void_p(texts_ptr),
                                        off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), off.size - 1,
                                        ctypes.byref(o), ctypes.c_void_p(sa_ptr), ctypes.byref(st))
        if rc != DCX_OK:
            self._raise(rc)
        return st.to_dict()

    def check_sa(self, text, sa, sa_width: int = 8) -> bool:
        t = _as_u8(text)
        s = np.ascontiguousarray(sa)
        ok = ctypes.c_int(0)
        rc = self._L.dcx_check_sa(self._h, t.ctypes.data if t.size else None, t.size,
                                  s.ctypes.data if t.size else None, sa_width, 0, ctypes.byref(ok))
        if rc != DCX_OK:
            self._raise(rc)
        return bool(ok.value)

    def check_sa_device(self, text_ptr: int, n: int, sa_ptr: int, sa_width: int = 8) -> bool:
        ok = ctypes.c_int(0)
        rc = self._L.dcx_check_sa(self._h, ctypes.c_void_p(text_ptr), n, ctypes.c_void_p(sa_ptr), sa_width, 1,
                                  ctypes.byref(ok))
        if rc != DCX_OK:
            self._raise(rc)
        return bool(ok.value)


def uint40_to_u64(b: np.ndarray) -> np.ndarray:
    """Decodes a little-endian uint40 stream (uint_types.hpp:262) to uint64."""
    a = np.frombuffer(bytes(b), dtype=np.uint8).reshape(-1, 5).astype(np.uint64)
    return a[:, 0] | (a[:, 1] << 8) | (a[:, 2] << 16) | (a[:, 3] << 24) | (a[:, 4] << 32)
