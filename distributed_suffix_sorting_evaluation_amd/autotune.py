"""Choice of the difference cover X and of the batched sub-batch size by measurement.

Every X yields the same suffix array (a suffix array is unique): X only moves work between the sample
sort (a larger X samples fewer suffixes and recurses on a shorter string) and the merge (its keys are
longer and its comparator reaches deeper).  Which X is fastest depends on the alphabet, the lengths and
the repetitiveness of the texts, so the batched path picks it by timing the candidates on a sample of the
workload.  The reference takes X as a command-line choice (``--dcx``, lightweight_dcx/src/executables/
cli.cpp:475-653); this module is for callers that leave it open.

The batched path joins texts into sub-batches of at most ``batch_max_chars`` characters.  The sub-batch
size is the batched path's radix key width: round 0 of a sub-batch of N characters keys about
log2(N) + 5 bits of information (under 2^24 characters a fixed number of characters of the code width),
and its texts are split apart by one radix pass per 8 bits of text index.  Smaller sub-batches take
fewer radix passes over the same bytes but pay the fixed costs of a pipeline more often; the tuner times
the candidate sizes too.  The SAs do not depend on it either.
"""
from __future__ import annotations

import time

import numpy as np

CANDIDATES = (7, 13, 21, 31, 39, 57, 73, 133)
SIZE_CANDIDATES = (1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 28)


def _time_batch(engine, texts_ptr, offsets, sa_ptr, reps, **kw):
    """One untimed call, then the best of ``reps`` timed calls, in ms."""
    engine.build_sa_batch_device(texts_ptr, offsets, sa_ptr, **kw)
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        engine.build_sa_batch_device(texts_ptr, offsets, sa_ptr, **kw)
        best = min(best, time.perf_counter() - t0)
    return round(best * 1e3, 3)


def choose_x_batch(engine, texts_ptr: int, offsets: np.ndarray, sa_ptr: int, candidates=CANDIDATES,
                   reps: int = 2, **opts):
    """Times ``build_sa_batch_device`` on the given device-resident texts for every candidate X (one
    untimed call each, then the best of ``reps``) and returns (best X, {X: ms})."""
    timings = {int(X): _time_batch(engine, texts_ptr, offsets, sa_ptr, reps, X=X, **opts) for X in candidates}
    x = min(timings, key=timings.get)
    return x, timings


def choose_batch_params(engine, texts_ptr: int, offsets: np.ndarray, sa_ptr: int, candidates=CANDIDATES,
                        sizes=SIZE_CANDIDATES, reps: int = 2, **opts):
    """X and the sub-batch size by coordinate search: X timed at the engine's default size, then every
    candidate size at that X.  Returns (X, batch_max_chars, {X: ms}, {size: ms})."""
    opts = {k: v for k, v in opts.items() if k != "batch_max_chars"}
    x, x_ms = choose_x_batch(engine, texts_ptr, offsets, sa_ptr, candidates, reps, **opts)
    size_ms = {int(s): _time_batch(engine, texts_ptr, offsets, sa_ptr, reps, X=x, batch_max_chars=int(s), **opts)
               for s in sizes}
    size = min(size_ms, key=size_ms.get)
    return x, size, x_ms, size_ms


class XTuner:
    """Caches the chosen X (and, with ``sizes``, the sub-batch size) per key, e.g. (alphabet class, log2
    size class) of the work it is asked about."""

    def __init__(self, engine, candidates=CANDIDATES, reps: int = 2, sizes=None, **opts):
        self.engine, self.candidates, self.reps, self.opts = engine, tuple(candidates), reps, opts
        self.sizes = tuple(sizes) if sizes else None
        self.choice: dict = {}
        self.size_choice: dict = {}

    def x_for(self, key, texts_ptr: int, offsets: np.ndarray, sa_ptr: int) -> int:
        if key not in self.choice:
            if self.sizes:
                x, size, _, _ = choose_batch_params(self.engine, texts_ptr, offsets, sa_ptr, self.candidates,
                                                    self.sizes, self.reps, **self.opts)
                self.size_choice[key] = size
            else:
                x = choose_x_batch(self.engine, texts_ptr, offsets, sa_ptr, self.candidates, self.reps,
                                   **self.opts)[0]
            self.choice[key] = x
        return self.choice[key]

    def size_for(self, key) -> int:
        """The sub-batch size chosen with the key's X (0, the engine's default, when sizes are not tuned)."""
        return self.size_choice.get(key, 0)
