"""MI355X DCX suffix-array engine -- host-side mirror of lightweight_dcx's PDCX interface.

``PDCX(config, dcx="dc39").compute_sa(text)`` follows ``PDCX<...>(PDCXConfig const&, Communicator&)`` and
``compute_sa(std::vector<char>&) -> std::vector<index>`` (lightweight_dcx/src/pdcx/pdcx.hpp:67-80, 529), with
the alphabet compression and packed-variant dispatch of the cli (cli.cpp:417-653) folded in.  The work runs in
``libdcx.so`` (hand-written gfx950 kernels); there is no CPU fallback.

Error behaviour: where the reference prints a message on rank 0 and calls ``exit(1)`` (missing file
cli.cpp:394-399, unsupported options), this mirror raises :class:`DCXError` carrying the C status code.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .engine import (DCXError, Engine, Stats, SUPPORTED_X, load_library, uint40_to_u64,  # noqa: F401
                     DCX_E_INVALID_X, DCX_E_ARG, DCX_E_WIDTH, DCX_E_ADMISSION, DCX_E_UNSUPPORTED)

__all__ = ["PDCXConfig", "PDCX", "DCXError", "Engine", "parse_dcx_variant", "compute_sa"]


@dataclass
class PDCXConfig:
    """Mirror of ``PDCXConfig`` (lightweight_dcx/src/pdcx/config.hpp:27-56).

    Fields that only steer the distributed sorters or bucket rounds of the MPI implementation are accepted
    (so configurations written for the reference load unchanged) but cannot change the output: a suffix
    array is unique.
    """
    discarding_threshold: float = 0.7
    use_char_packing_samples: bool = False
    use_char_packing_merging: bool = False
    seed: int = 0
    pack_extra_words: int = 0
    # accepted for interface parity; no effect on the single-GPU pipeline
    buckets_samples: list = field(default_factory=list)
    buckets_merging: list = field(default_factory=list)
    buckets_phase3: int = 1
    num_samples_phase3: int = 10000
    ams_levels: int = 1
    num_samples_splitters: int = 1000
    avg_chunks_pe: int = 1000
    min_imbalance: float = 0.25
    atomic_sorter: str = "sample_sort"
    use_string_sort: bool = False
    use_random_sampling_splitters: bool = False
    use_randomized_chunks: bool = False
    rearrange_buckets_balanced: bool = False
    use_robust_tie_break: bool = False
    use_compressed_buckets: bool = False
    # sequential base case threshold: n <= max(2 * X, base_case_chars)  (pdcx.hpp:623-624 with p = 1)
    base_case_chars: int = 10000

    def validate(self):
        if not (0.0 <= self.discarding_threshold <= 1.0):
            raise DCXError(DCX_E_ARG, "--discarding-threshold must lie in [0, 1]")
        if self.pack_extra_words != 0:
            # cli.cpp:671-679 only instantiates EXTRA_WORDS = 0
            raise DCXError(DCX_E_ARG, "--pack-extra-words supports only 0")
        for b in list(self.buckets_samples) + list(self.buckets_merging):
            if b > 65535:
                raise DCXError(DCX_E_ARG, "bucket counts must be <= 65535 (cli.cpp:190)")


def parse_dcx_variant(name) -> int:
    """``--dcx dcX`` -> X (cli.cpp:180-182, 544-652)."""
    if isinstance(name, int):
        X = name
    else:
        s = str(name).strip().lower()
        if not s.startswith("dc"):
            raise DCXError(DCX_E_INVALID_X, f"invalid --dcx {name!r}")
        try:
            X = int(s[2:])
        except ValueError:
            raise DCXError(DCX_E_INVALID_X, f"invalid --dcx {name!r}") from None
    if X not in SUPPORTED_X:
        raise DCXError(DCX_E_INVALID_X, f"dc{X} is not available; choose one of " +
                       ", ".join(f"dc{x}" for x in SUPPORTED_X))
    return X


class PDCX:
    """DCX suffix-array construction on one MI355X (one engine handle)."""

    def __init__(self, config: PDCXConfig | None = None, dcx="dc39", device: int = 0, engine: Engine | None = None):
        self.config = config or PDCXConfig()
        self.config.validate()
        self.X = parse_dcx_variant(dcx)
        self.engine = engine or Engine(device)
        self.last_stats = None

    def compute_sa(self, text, sa_width: int = 8) -> np.ndarray:
        """Suffix array of ``text`` (bytes / uint8 array).  Width 8 -> uint64, 4 -> uint32, 5 -> uint64 decoded
        from the uint40 stream."""
        c = self.config
        sa, st = self.engine.build_sa(text, sa_width=sa_width, stats=True, X=self.X,
                                      pack_samples=c.use_char_packing_samples,
                                      pack_merging=c.use_char_packing_merging,
                                      discarding_threshold=c.discarding_threshold,
                                      base_case_chars=c.base_case_chars, seed=c.seed)
        self.last_stats = st
        if sa_width == 5:
            return uint40_to_u64(sa)
        return sa


def compute_sa(text, dcx="dc39", packing=True, **cfg) -> np.ndarray:
    """One-shot helper: SA of ``text`` with the packed DC39 configuration of the reference's README run."""
    config = PDCXConfig(use_char_packing_samples=packing, use_char_packing_merging=packing, **cfg)
    return PDCX(config, dcx=dcx).compute_sa(text)
