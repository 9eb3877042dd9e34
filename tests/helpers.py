"""Shared test inputs and brute-force references."""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ALL_X = (3, 7, 13, 21, 31, 39, 57, 73, 91, 95, 133)


def brute_sa(text) -> np.ndarray:
    """SA by sorting all suffixes as byte strings (tiny inputs only)."""
    b = bytes(text)
    return np.array(sorted(range(len(b)), key=lambda i: b[i:]), dtype=np.uint64)


def read(name) -> bytes:
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return f.read()


def random_text(n, sigma, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, sigma, n, dtype=np.uint16).astype(np.uint8)


def periodic_text(n, period, sigma, seed):
    rng = np.random.default_rng(seed)
    p = rng.integers(0, sigma, period, dtype=np.uint16).astype(np.uint8)
    return np.resize(p, n)


def repeat_words_text(words, reps, seed):
    """psac test_psac.cpp:178-224 style: random concatenation of a few words."""
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(words), reps)
    return np.frombuffer(b"".join(words[i] for i in idx), dtype=np.uint8).copy()


def copies_text(n, seed, sigma=4, copy_frac=0.3, max_len=2000):
    """Random text with copied segments (long LCPs) and point mutations."""
    rng = np.random.default_rng(seed)
    t = rng.integers(0, sigma, n, dtype=np.uint16).astype(np.uint8)
    done = 0
    while done < copy_frac * n:
        ln = int(rng.integers(20, max_len))
        if 2 * ln + 1 >= n:
            break
        dst = int(rng.integers(ln, n - ln))
        src = int(rng.integers(0, dst - ln + 1))
        seg = t[src:src + ln].copy()
        mut = rng.random(ln) < 0.01
        seg[mut] = rng.integers(0, sigma, int(mut.sum()))
        t[dst:dst + ln] = seg
        done += ln
    return t


def edge_cases():
    """(name, text) pairs covering the reference's edge cases: empty, single, runs, alphabet extremes."""
    cases = [
        ("empty", b""),
        ("one", b"a"),
        ("two_same", b"aa"),
        ("two_diff", b"ba"),
        ("mississippi", b"mississippi"),
        ("run_1000", b"a" * 1000),
        ("run_30000", b"z" * 30000),
        ("alt_20000", b"ab" * 10000),
        ("all_256_bytes", bytes(range(256)) * 50),
        ("bytes_255", bytes(range(255)) * 60),
        ("zeros", bytes(5000)),
        ("descending", bytes(range(255, -1, -1)) * 40),
    ]
    return cases
