"""ctypes bindings of the CPU checker (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this package.  It binds

* ``liboracle.so`` -- ``dcx_oracle.c``, the C restatement of lightweight_dcx's PDCX path
  (pdcx.hpp, sample_string.hpp, compute_ranks.hpp, merge_samples.hpp, cli.cpp:417-539);
* ``libhostdcx.so`` -- ``host_dcx.cpp``, the multi-threaded C++ restatement of the same path (the
  host-DCX baseline bench.py times; checked here against libsais like the scalar restatement);
* ``_ref/libsais.so`` -- libsais 2.10.0 built by ``oracle/Makefile`` from the reference's own
  sources (libsais/src/libsais.c, libsais64.c), used as an independent SACA.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_SAIS = None
_HOST = None

MAX_LEVELS = 64


class LevelStat(ctypes.Structure):
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
    _fields_ = [("levels", ctypes.c_uint32), ("sigma", ctypes.c_uint32), ("lv", LevelStat * MAX_LEVELS)]

    def as_list(self):
        out = []
        for i in range(self.levels):
            l = self.lv[i]
            out.append({f: int(getattr(l, f)) for f, _ in LevelStat._fields_})
        return out


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.oracle_dcx_sa.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_uint64, u64p, ctypes.POINTER(Stats)]
        L.oracle_dcx_sa.restype = ctypes.c_int
        L.oracle_dcx_sa_int.argtypes = [u64p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64,
                                        u64p, ctypes.POINTER(Stats)]
        L.oracle_dcx_sa_int.restype = ctypes.c_int
        L.oracle_dc_params.argtypes = [ctypes.c_uint32, u32p, u32p, u32p, u32p]
        L.oracle_dc_params.restype = ctypes.c_int
        L.oracle_pack_word.argtypes = [u64p, ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_pack_word.restype = ctypes.c_uint64
        L.oracle_packing.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, u32p,
                                     ctypes.POINTER(ctypes.c_double)]
        L.oracle_check_sa.argtypes = [ctypes.c_void_p, ctypes.c_uint64, u64p]
        L.oracle_check_sa.restype = ctypes.c_int
        L.oracle_slow_sa.argtypes = [u64p, ctypes.c_uint64, u64p]
        L.oracle_compress_alphabet.argtypes = [ctypes.c_void_p, ctypes.c_uint64, u64p, u32p]
        L.oracle_write_uint40.argtypes = [u64p, ctypes.c_uint64, ctypes.c_void_p]
        _LIB = L
    return _LIB


def _u64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def dcx_sa(text, X=39, pack_samples=True, pack_merging=True, discarding_threshold=0.7,
           base_case_chars=10000, with_stats=False):
    """PDCX::compute_sa restated on the CPU; returns the SA as uint64 (and per-level stats)."""
    t = np.ascontiguousarray(np.frombuffer(bytes(text), dtype=np.uint8) if not isinstance(text, np.ndarray) else text,
                             dtype=np.uint8)
    n = t.size
    sa = np.zeros(max(n, 1), dtype=np.uint64)
    st = Stats()
    rc = lib().oracle_dcx_sa(t.ctypes.data if n else None, n, X, int(pack_samples), int(pack_merging),
                             float(discarding_threshold), int(base_case_chars), _u64p(sa), ctypes.byref(st))
    if rc != 0:
        raise ValueError(f"oracle_dcx_sa failed rc={rc}")
    sa = sa[:n]
    return (sa, st.as_list()) if with_stats else sa


def dcx_sa_int(T, X=39, discarding_threshold=0.7, base_case_chars=10000, with_stats=False):
    T = np.ascontiguousarray(T, dtype=np.uint64)
    sa = np.zeros(max(T.size, 1), dtype=np.uint64)
    st = Stats()
    rc = lib().oracle_dcx_sa_int(_u64p(T), T.size, X, float(discarding_threshold), int(base_case_chars),
                                 _u64p(sa), ctypes.byref(st))
    if rc != 0:
        raise ValueError(f"oracle_dcx_sa_int failed rc={rc}")
    sa = sa[:T.size]
    return (sa, st.as_list()) if with_stats else sa


def dc_params(X):
    D = ctypes.c_uint32()
    dc = (ctypes.c_uint32 * 16)()
    nr = (ctypes.c_uint32 * 256)()
    cmp = np.zeros(3 * X * X, dtype=np.uint32)
    rc = lib().oracle_dc_params(X, ctypes.byref(D), dc, nr, cmp.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    if rc != 0:
        raise ValueError(f"unsupported X={X}")
    return {"X": X, "D": D.value, "DC": [dc[i] for i in range(D.value)], "NEXT_RANK": [nr[i] for i in range(X)],
            "cmp": cmp.reshape(X, X, 3)}


def packing(X, sigma):
    b, cpw, w = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    r = ctypes.c_double()
    lib().oracle_packing(X, sigma, ctypes.byref(b), ctypes.byref(cpw), ctypes.byref(w), ctypes.byref(r))
    return {"bits": b.value, "chars_per_word": cpw.value, "words": w.value, "ratio": r.value}


def pack_word(chars, bits):
    c = np.ascontiguousarray(chars, dtype=np.uint64)
    return int(lib().oracle_pack_word(_u64p(c), c.size, bits))


def slow_sa(T):
    T = np.ascontiguousarray(T, dtype=np.uint64)
    sa = np.zeros(max(T.size, 1), dtype=np.uint64)
    lib().oracle_slow_sa(_u64p(T), T.size, _u64p(sa))
    return sa[:T.size]


def check_sa(text, sa):
    t = np.ascontiguousarray(np.frombuffer(bytes(text), dtype=np.uint8) if not isinstance(text, np.ndarray) else text,
                             dtype=np.uint8)
    s = np.ascontiguousarray(sa, dtype=np.uint64)
    if s.size != t.size:
        return False
    if t.size == 0:
        return True
    return lib().oracle_check_sa(t.ctypes.data, t.size, _u64p(s)) == 1


def uint40_bytes(sa):
    s = np.ascontiguousarray(sa, dtype=np.uint64)
    out = np.zeros(5 * s.size, dtype=np.uint8)
    if s.size:
        lib().oracle_write_uint40(_u64p(s), s.size, out.ctypes.data)
    return out.tobytes()


# ---------------------------------------------------------------------------------------------
# libsais (reference sources, built into oracle/_ref)
# ---------------------------------------------------------------------------------------------
def sais_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libsais.so"))


def sais():
    global _SAIS
    if _SAIS is None:
        path = os.path.join(HERE, "_ref", "libsais.so")
        if not os.path.exists(path):
            raise FileNotFoundError("oracle/_ref/libsais.so missing (built from the reference by oracle/Makefile)")
        L = ctypes.CDLL(path)
        L.libsais64.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p]
        L.libsais64.restype = ctypes.c_int64
        L.libsais64_omp.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_void_p, ctypes.c_int32]
        L.libsais64_omp.restype = ctypes.c_int64
        _SAIS = L
    return _SAIS


def libsais_sa(text, threads=1):
    """SA by libsais64 (libsais/include/libsais64.h) as uint64."""
    t = np.ascontiguousarray(np.frombuffer(bytes(text), dtype=np.uint8) if not isinstance(text, np.ndarray) else text,
                             dtype=np.uint8)
    n = t.size
    sa = np.zeros(max(n, 1), dtype=np.int64)
    if n:
        if threads > 1:
            rc = sais().libsais64_omp(t.ctypes.data, sa.ctypes.data, n, 0, None, threads)
        else:
            rc = sais().libsais64(t.ctypes.data, sa.ctypes.data, n, 0, None)
        if rc != 0:
            raise RuntimeError(f"libsais64 rc={rc}")
    return sa[:n].astype(np.uint64)


# ---------------------------------------------------------------------------------------------
# host DCX: multi-threaded restatement (host_dcx.cpp)
# ---------------------------------------------------------------------------------------------
def host_lib():
    global _HOST
    if _HOST is None:
        path = os.path.join(HERE, "libhostdcx.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.host_dcx_sa.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p]
        L.host_dcx_sa.restype = ctypes.c_int
        _HOST = L
    return _HOST


def host_dcx_sa(text, X=39, threads=1, pack_samples=True, pack_merging=True, discarding_threshold=0.7,
                base_case_chars=10000):
    """SA (uint64) by the multi-threaded host restatement of PDCX with `threads` threads."""
    t = np.ascontiguousarray(np.frombuffer(bytes(text), dtype=np.uint8) if not isinstance(text, np.ndarray) else text,
                             dtype=np.uint8)
    sa = np.zeros(max(t.size, 1), dtype=np.uint64)
    rc = host_lib().host_dcx_sa(t.ctypes.data, t.size, X, int(pack_samples), int(pack_merging),
                                float(discarding_threshold), int(base_case_chars), int(threads), sa.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"host_dcx_sa rc={rc}")
    return sa[:t.size]
