"""Choice of the difference cover X by measurement.

Every X yields the same suffix array (a suffix array is unique): X only moves work between the sample
sort (a larger X samples fewer suffixes and recurses on a shorter string) and the merge (its keys are
longer and its comparator reaches deeper).  Which X is fastest depends on the alphabet, the lengths and
the repetitiveness of the texts, so the batched path picks it by timing the candidates on a sample of the
workload.  The reference takes X as a command-line choice (``--dcx``, lightweight_dcx/src/executables/
cli.cpp:475-653); this module is for callers that leave it open.
"""
from __future__ import annotations

import time

import numpy as np

CANDIDATES = (7, 13, 21, 31, 39, 57, 73, 133)


def choose_x_batch(engine, texts_ptr: int, offsets: np.ndarray, sa_ptr: int, candidates=CANDIDATES,
                   reps: int = 2, **opts):
    """Times ``build_sa_batch_device`` on the given device-resident texts for every candidate X (one
    untimed call each, then the best of ``reps``) and returns (best X, {X: ms})."""
    timings = {}
    for X in candidates:
        engine.build_sa_batch_device(texts_ptr, offsets, sa_ptr, X=X, **opts)
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            engine.build_sa_batch_device(texts_ptr, offsets, sa_ptr, X=X, **opts)
            best = min(best, time.perf_counter() - t0)
        timings[int(X)] = round(best * 1e3, 3)
    x = min(timings, key=timings.get)
    return x, timings


class XTuner:
    """Caches the chosen X per (alphabet class, log2 size class) of the work it is asked about."""

    def __init__(self, engine, candidates=CANDIDATES, reps: int = 2, **opts):
        self.engine, self.candidates, self.reps, self.opts = engine, tuple(candidates), reps, opts
        self.choice: dict = {}

    def x_for(self, key, texts_ptr: int, offsets: np.ndarray, sa_ptr: int) -> int:
        if key not in self.choice:
            self.choice[key] = choose_x_batch(self.engine, texts_ptr, offsets, sa_ptr, self.candidates,
                                              self.reps, **self.opts)[0]
        return self.choice[key]
