"""Parity of the HIP engine (through the C ABI) with the oracle, the golden fixtures and libsais.

Integer output: the bar is bit-exact.  Per-level statistics (sizes, names, discarding decisions) are compared
with the oracle as well, which pins the intermediate phases, not only the final SA.
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle
from helpers import (ALL_X, GOLDEN, brute_sa, copies_text, edge_cases, periodic_text, random_text, read,
                     repeat_words_text)

pytestmark = pytest.mark.gpu

STAT_KEYS = ("n", "samples", "names", "reduced", "base_case", "recursed", "discarding", "key_chars", "merge_chars",
             "bits_per_char")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_sa(engine, text, **kw):
    sa, st = engine.build_sa(text, sa_width=8, stats=True, **kw)
    return sa, st


def same_stats(gpu_levels, oracle_levels):
    g = [{k: l[k] for k in STAT_KEYS} for l in gpu_levels]
    o = [{k: l[k] for k in STAT_KEYS} for l in oracle_levels]
    return g == o


def check_case(engine, name, t, X, ps=True, pm=True, base=0, thr=0.7, stats=True):
    say(f"case {name} n={len(t)} X={X} pack=({ps},{pm}) base={base} thr={thr}")
    sa, st = gpu_sa(engine, t, X=X, pack_samples=ps, pack_merging=pm, base_case_chars=base, discarding_threshold=thr)
    ref, ost = oracle.dcx_sa(t, X=X, pack_samples=ps, pack_merging=pm, base_case_chars=base,
                             discarding_threshold=thr, with_stats=True)
    assert np.array_equal(sa, ref), f"SA mismatch {name} X={X}"
    if stats:
        assert same_stats(st["levels"], ost), (name, X, st["levels"], ost)
    return st


def test_golden_fixtures(engine):
    for name in ["abracadabra4", "abracadabra8"]:
        text = read(name)
        expect = np.frombuffer(read(name + ".sa"), dtype="<u4").astype(np.uint64)
        for X in (3, 7, 39):
            say(f"fixture {name} X={X}")
            sa, _ = gpu_sa(engine, text, X=X, base_case_chars=0)
            assert np.array_equal(sa, expect)
    text = read("the_three_brothers.txt")
    expect = np.frombuffer(read("the_three_brothers_size_t_sa"), dtype="<u8")
    for X in (3, 13, 39, 133):
        for packing in (True, False):
            say(f"fixture three_brothers X={X} packing={packing}")
            sa, _ = gpu_sa(engine, text, X=X, pack_samples=packing, pack_merging=packing, base_case_chars=0)
            assert np.array_equal(sa, expect)
    sa, _ = gpu_sa(engine, b"mississippi", X=3, base_case_chars=0)
    assert sa.tolist() == [10, 7, 4, 1, 0, 9, 8, 6, 3, 5, 2]
    aaa = read("aaa")
    sa, _ = gpu_sa(engine, aaa, X=39, base_case_chars=0)
    assert sa.tolist() == list(range(len(aaa) - 1, -1, -1))


def test_base_case_kernel(engine):
    # default threshold 10000: the whole text is solved by the single-workgroup kernel
    for n, sigma in ((1, 4), (2, 2), (17, 3), (1000, 2), (9999, 256), (10000, 4)):
        t = random_text(n, sigma, n)
        say(f"base case n={n}")
        sa, st = gpu_sa(engine, t, X=39)
        assert st["levels"][0]["base_case"] == 1
        assert np.array_equal(sa, oracle.dcx_sa(t, X=39))
    t = periodic_text(16384, 1, 4, 1)
    sa, st = gpu_sa(engine, t, X=39, base_case_chars=16384)
    assert np.array_equal(sa, np.arange(16383, -1, -1, dtype=np.uint64))


def test_large_base_case(engine):
    """base_case_chars above the single-workgroup kernel's 16,384 characters: such a level is sorted by the
    refinement over its whole length (prefix doubling).  Random, periodic and copied texts solved whole by
    the base case, and a text whose deeper levels (32-bit names) fall under a large threshold; SAs and
    per-level statistics (which level took the base case) against the oracle."""
    # the oracle's base case is the reference's comparison sort of whole suffixes: repetitive texts stay short
    cases = [("random4", random_text(60000, 4, 31), 100000), ("random256", random_text(40000, 256, 32), 40000),
             ("run", periodic_text(20000, 1, 4, 33), 1 << 20), ("period7", periodic_text(20000, 7, 3, 34), 20000),
             ("copies", copies_text(60000, 35, copy_frac=0.5, max_len=3000), 200000),
             ("copies_deep", copies_text(400000, 36, copy_frac=0.4, max_len=5000), 60000)]
    for name, t, base in cases:
        for X in (7, 39):
            st = check_case(engine, name, t, X, base=base)
            assert any(l["base_case"] for l in st["levels"]), name
    t = random_text(1 << 20, 4, 37)
    say("base case of 2^20 characters")
    sa, st = gpu_sa(engine, t, X=39, base_case_chars=1 << 20)
    assert st["levels"][0]["base_case"] == 1 and len(st["levels"]) == 1
    assert np.array_equal(sa, oracle.libsais_sa(t))


@pytest.mark.parametrize("X", ALL_X)
def test_random_all_x(engine, X):
    for n, sigma in ((3 * X + 5, 2), (5000, 4), (40000, 4), (40000, 26), (30000, 255), (30000, 256)):
        t = random_text(n, sigma, 1000 * X + n + sigma)
        for ps, pm in ((True, True), (False, False), (True, False), (False, True)):
            check_case(engine, f"random s={sigma}", t, X, ps, pm)


@pytest.mark.parametrize("X", ALL_X)
def test_alignments_brute(engine, X):
    # every n mod X (lightweight_dcx main.cpp:100-123), tiny texts against brute force
    for n in range(2 * X + 1, 2 * X + 2 + X):
        t = random_text(n, 3, n)
        say(f"align X={X} n={n}")
        sa, _ = gpu_sa(engine, t, X=X, base_case_chars=0)
        assert np.array_equal(sa, brute_sa(t)), (X, n)


def test_repetitive(engine):
    cases = [
        ("period17_256", periodic_text(60000, 17, 256, 1)),
        ("run", periodic_text(50000, 1, 4, 2)),
        ("period3", periodic_text(70000, 3, 4, 3)),
        ("period2", periodic_text(65537, 2, 2, 4)),
        ("copies", copies_text(120000, 4)),
        ("copies256", copies_text(90000, 5, sigma=256)),
        ("copies_long", copies_text(200000, 6, copy_frac=0.6, max_len=20000)),
    ]
    for name, t in cases:
        for X in (3, 7, 21, 39, 133):
            check_case(engine, name, t, X)


def test_discarding_threshold(engine):
    t = copies_text(150000, 11, copy_frac=0.5)
    for thr in (0.0, 0.3, 0.7, 1.0):
        for X in (7, 39):
            check_case(engine, "copies", t, X, thr=thr)


def test_default_base_case_multilevel(engine):
    t = copies_text(400000, 12, copy_frac=0.4, max_len=5000)
    for X in (13, 39, 133):
        check_case(engine, "copies400k", t, X, base=10000)


def test_edge_cases(engine):
    for name, t in edge_cases():
        for X in (3, 39, 133):
            check_case(engine, name, np.frombuffer(t, dtype=np.uint8), X)


def test_sa_widths(engine):
    t = random_text(50000, 4, 77)
    ref = oracle.dcx_sa(t, X=39)
    s8 = engine.build_sa(t, sa_width=8, X=39)
    s4 = engine.build_sa(t, sa_width=4, X=39)
    s5 = engine.build_sa(t, sa_width=5, X=39)
    assert s8.dtype == np.uint64 and np.array_equal(s8, ref)
    assert s4.dtype == np.uint32 and np.array_equal(s4.astype(np.uint64), ref)
    assert s5.tobytes() == oracle.uint40_bytes(ref)


def test_device_output_64bit_from_the_level(engine):
    """A 64-bit SA in device memory is written by level 0 itself (the sample scan's second pass writes every
    entry, the tied groups are written again after the DC-comparator merge) when the text takes the two-pass
    scan (more than 2^20 characters): a text with copies and a periodic run (large tied groups the merge
    reorders) against libsais, into a 16-byte aligned output (vector stores) and into one 8 bytes off.  A
    batched build and a 32-bit build on the same engine afterwards must not touch that output again."""
    import torch
    t = copies_text(3 << 20, 41, sigma=4, copy_frac=0.5, max_len=4000)
    t[1_000_000:1_200_000] = np.tile(t[:13], 200_000 // 13 + 1)[:200_000]
    ref = oracle.libsais_sa(t, threads=8)
    n = t.size
    dt = torch.from_numpy(t).cuda()
    for X in (7, 39):
        for off in (0, 8):
            say(f"64-bit device output X={X} offset={off}")
            buf = torch.zeros(8 * n + 16, dtype=torch.uint8, device="cuda")
            engine.build_sa_device(dt.data_ptr(), n, buf.data_ptr() + off, 8, X=X)
            got = buf[off:off + 8 * n].cpu().numpy().view("<u8")
            assert np.array_equal(got, ref), (X, off)
    # the same engine next: a batch (level 0 of every sub-batch) and a 32-bit device build
    say("batch and 32-bit build after the 64-bit device output")
    off = np.array([0, 1 << 20, 2 << 20, n], dtype=np.uint64)
    ds = torch.zeros(n, dtype=torch.int32, device="cuda")
    engine.build_sa_batch_device(dt.data_ptr(), off, ds.data_ptr(), X=39)
    got = ds.cpu().numpy().view(np.uint32)
    for k in range(3):
        a, b = int(off[k]), int(off[k + 1])
        assert np.array_equal(got[a:b].astype(np.uint64), oracle.libsais_sa(t[a:b], threads=8)), k
    engine.build_sa_device(dt.data_ptr(), n, ds.data_ptr(), 4, X=39)
    assert np.array_equal(ds.cpu().numpy().view(np.uint32).astype(np.uint64), ref)
    del dt, ds
    torch.cuda.empty_cache()


def test_width5_device_output_tails_and_alignment(engine):
    """40-bit output written on the device: lengths that are not a multiple of 16 and an output pointer
    that is not 16-byte aligned take the per-entry path, the rest the 16-entry vector path."""
    import torch
    for n in (100003, 16 * 4096 + 7, 16 * 5000):
        t = random_text(n, 4, n % 1000)
        want = oracle.uint40_bytes(oracle.dcx_sa(t, X=39))
        dt = torch.from_numpy(t).cuda()
        for off in (0, 3):
            buf = torch.zeros(5 * n + 32, dtype=torch.uint8, device="cuda")
            engine.build_sa_device(dt.data_ptr(), n, buf.data_ptr() + off, 5, X=39)
            torch.cuda.synchronize()
            assert buf[off:off + 5 * n].cpu().numpy().tobytes() == want, (n, off)
            assert int(buf[:off].sum()) == 0 and int(buf[off + 5 * n:].sum()) == 0


def test_device_output_all_widths_two_pass(engine):
    """A 3 MiB text with copies and a periodic run (small and large tied groups for the merge) built into
    device outputs of widths 4, 5 and 8, aligned and misaligned, against the oracle; the bytes around
    the output stay untouched."""
    import torch
    n = (3 << 20) + 13
    t = copies_text(n - 300007, 11)
    t = np.concatenate([t[: n // 2], periodic_text(300007, 7, 4, 3), t[n // 2:]])[:n].copy()
    assert len(t) == n
    ref = oracle.dcx_sa(t, X=39)
    dt = torch.from_numpy(t).cuda()
    for width in (4, 5, 8):
        if width == 5:
            want = oracle.uint40_bytes(ref)
        else:
            want = ref.astype("<u4" if width == 4 else "<u8").tobytes()
        for off in (0, 3):
            say(f"device output width={width} off={off}")
            buf = torch.zeros(width * n + 32, dtype=torch.uint8, device="cuda")
            engine.build_sa_device(dt.data_ptr(), n, buf.data_ptr() + off, width, X=39)
            torch.cuda.synchronize()
            assert buf[off:off + width * n].cpu().numpy().tobytes() == want, (width, off)
            assert int(buf[:off].sum()) == 0 and int(buf[off + width * n:].sum()) == 0


def test_device_checker(engine):
    t = copies_text(100000, 3)
    sa = engine.build_sa(t, sa_width=8, X=39)
    assert engine.check_sa(t, sa, 8)
    bad = sa.copy()
    bad[[100, 101]] = bad[[101, 100]]
    assert not engine.check_sa(t, bad, 8)
    bad = sa.copy()
    bad[7] = bad[8]
    assert not engine.check_sa(t, bad, 8)
    bad = sa.copy()
    bad[0] = len(t) + 5
    assert not engine.check_sa(t, bad, 8)
    s5 = engine.build_sa(t, sa_width=5, X=39)
    assert engine.check_sa(t, s5, 5)


def test_errors(engine):
    from distributed_suffix_sorting_evaluation_amd import DCXError
    t = random_text(1000, 4, 1)
    for kw, code in (({"X": 40}, -1), ({"base_case_chars": (1 << 31) + 1}, -6), ({"discarding_threshold": 2.0}, -6)):
        with pytest.raises(DCXError) as e:
            engine.build_sa(t, **kw)
        assert e.value.code == code
    with pytest.raises(KeyError):
        engine.build_sa(t, sa_width=3)
    # texts beyond the longest the build takes (2^32 characters: 32-bit positions) are refused before
    # anything is read: the zero pages of these arrays stay unmapped
    from distributed_suffix_sorting_evaluation_amd import textgen
    with pytest.raises(DCXError) as e:
        engine.build_sa(np.zeros(textgen.MAX_TEXT + 1, dtype=np.uint8), sa_width=8)
    assert e.value.code == -9
    # the handle still works after errors
    assert np.array_equal(engine.build_sa(t, X=39), oracle.dcx_sa(t, X=39))


def test_dna_8mib_against_libsais_and_digest(engine):
    from distributed_suffix_sorting_evaluation_amd import textgen
    t = textgen.generate("uniform_acgt", 8 << 20, 1)
    say("dna 8MiB")
    sa, st = gpu_sa(engine, t, X=39)
    if oracle.sais_available():
        assert np.array_equal(sa, oracle.libsais_sa(t, threads=8))
    path = os.path.join(GOLDEN, "sa_digests.json")
    if os.path.exists(path):
        import hashlib
        d = json.load(open(path))["dna_8mib"]
        assert hashlib.sha256(sa.astype("<u8").tobytes()).hexdigest() == d["sa_u64_sha256"]
    assert st["levels"][0]["samples"] == 1505647


def test_genome_64mib_digest(engine):
    import hashlib
    from distributed_suffix_sorting_evaluation_amd import textgen
    path = os.path.join(GOLDEN, "sa_digests.json")
    if not os.path.exists(path):
        pytest.skip("digests not generated")
    rec = json.load(open(path))["genome_64mib"]
    t = textgen.generate(**rec["params"])
    say("genome 64MiB")
    sa = engine.build_sa(t, sa_width=8, X=39)
    assert hashlib.sha256(sa.astype("<u8").tobytes()).hexdigest() == rec["sa_u64_sha256"]
    assert engine.check_sa(t, sa, 8)


def _coded_round0_texts():
    """Texts of 2^24 + a few characters whose first sort round is keyed by the alphabetic code of their
    characters: Zipf bytes with a periodic run (32 code bits above the position, keys-only passes), bytes that
    are 95% one value (codewords of one bit: 40 code bits reach past the key length) and a skewed 4-letter
    text (the code resolves 30 characters in 5 radix passes instead of 8), both keyed by 40 code bits beside
    split positions, and a skewed 5-letter text (sentinel-coded characters: the padding is a code of its
    own)."""
    rng = np.random.default_rng(77)
    n = (1 << 24) + 4097
    r = np.arange(1, 257, dtype=np.float64) ** -1.1
    z = rng.choice(256, size=n, p=r / r.sum()).astype(np.uint8)
    z[5_000_000:5_000_000 + 170_000] = np.tile(z[:17], 10_000)
    z[-300:] = z[1000:1300]  # the text's end repeats an earlier stretch: boundary suffixes in tied groups
    skew = np.where(rng.random(n) < 0.95, 7, rng.integers(0, 256, size=n)).astype(np.uint8)
    skew[:256] = np.arange(256, dtype=np.uint8)  # all 256 byte values occur
    dna = rng.choice(4, size=n, p=[0.82, 0.1, 0.05, 0.03]).astype(np.uint8) + ord("A")
    five = np.frombuffer(b"ACGTN", dtype=np.uint8)[rng.choice(5, size=n, p=[0.4, 0.25, 0.2, 0.1, 0.05])]
    return [("zipf", z), ("one_value_95", skew), ("dna_skewed", dna), ("five_letters", five)]


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_coded_round0_against_oracle(engine, case):
    """Round 0 keyed by an alphabetic code (texts of at least 2^24 characters with dense character words):
    the SA and every level's statistics (names, reduced, recursion) equal the oracle's, so the groups after
    the code-keyed round and the group rounds that continue from each member's decoded character count are
    exactly the groups of equal L-character keys."""
    name, t = _coded_round0_texts()[case]
    for X in (39, 7):
        say(f"coded round 0 {name} n={t.size} X={X}")
        sa, st = gpu_sa(engine, t, X=X)
        ref, ost = oracle.dcx_sa(t, X=X, with_stats=True)
        assert np.array_equal(sa, ref), (name, X)
        assert same_stats(st["levels"], ost), (name, X, st["levels"], ost)


def test_coded_round0_every_x_against_libsais(engine):
    """The code-keyed round 0 under every difference cover: the byte-alphabet text (32 code bits above the
    position) and the sentinel-coded 5-letter text of 2^24 + 4,097 characters, their SAs against libsais for
    all eleven X (L from 3 to 136 characters: the code's character count capped at L, look-ahead partitions
    of the periodic run over up to 19 rounds)."""
    if not oracle.sais_available():
        pytest.skip("libsais not built")
    texts = _coded_round0_texts()
    for name, t in (texts[0], texts[3]):
        ref = oracle.libsais_sa(t, threads=8)
        for X in ALL_X:
            say(f"coded round 0 {name} X={X} vs libsais")
            sa, _ = gpu_sa(engine, t, X=X)
            assert np.array_equal(sa, ref), (name, X)


def test_zipf_16mib_digest_dc21_dc39_dc133(engine):
    """The byte-alphabet workload (Zipf(1.1) over 256 values plus a period-17 block), sigma = 256 (9-bit
    packing), at 16 MiB: every X gives the SA whose libsais digest is committed."""
    import hashlib
    from distributed_suffix_sorting_evaluation_amd import textgen
    path = os.path.join(GOLDEN, "sa_digests.json")
    if not os.path.exists(path):
        pytest.skip("digests not generated")
    rec = json.load(open(path))["zipf_16mib"]
    t = textgen.generate(**rec["params"])
    assert hashlib.sha256(t.tobytes()).hexdigest() == rec["text_sha256"]
    for X in (21, 39, 133):
        say(f"zipf 16MiB X={X}")
        sa = engine.build_sa(t, sa_width=8, X=X)
        assert hashlib.sha256(sa.astype("<u8").tobytes()).hexdigest() == rec["sa_u64_sha256"], X


def test_genome_1gib_digest(engine):
    """The 1 GiB genome-like workload at full size against its committed libsais digest (64-bit SA)."""
    import hashlib
    from distributed_suffix_sorting_evaluation_amd import textgen
    path = os.path.join(GOLDEN, "sa_digests.json")
    if not os.path.exists(path):
        pytest.skip("digests not generated")
    rec = json.load(open(path))["genome_1gib"]
    t = textgen.generate(**rec["params"])
    say("genome 1GiB")
    sa = engine.build_sa(t, sa_width=8, X=39)
    assert hashlib.sha256(sa.tobytes()).hexdigest() == rec["sa_u64_sha256"]


def test_genome_max_text_device_check(engine):
    """The longest text the build takes, 2^32 characters over ACGT: every count of level 0 (positions,
    tied positions, radix bins, scan totals) reaches 2^32, past 32 bits, while the positions 0 .. 2^32 - 1
    still fit them; round 0 sorts 40-bit keys whose positions need all 32 bits (24 in the free key bits, 8
    in the byte beside).  The 40-bit SA stays in HBM and is checked there (permutation, suffix order by
    first character and the rank of the next)."""
    import torch
    from distributed_suffix_sorting_evaluation_amd import textgen
    n = textgen.MAX_TEXT
    t = textgen.generate("genome", n, 7)
    say(f"genome max text n={n}")
    dt = torch.from_numpy(t).to("cuda")
    del t
    ds = torch.empty(n * 5, dtype=torch.uint8, device="cuda")
    st = engine.build_sa_device(dt.data_ptr(), n, ds.data_ptr(), 5, X=39)
    assert st["levels"][0]["n"] == n
    assert engine.check_sa_device(dt.data_ptr(), n, ds.data_ptr(), 5)
    del dt, ds
    torch.cuda.empty_cache()


def test_large_groups_partition(engine):
    """Large tied groups whose members share most continuations: a period-13 run (its groups split off only
    the members near the run's end in each round) and 30,000 copies of one 60-character string behind
    random 12-character prefixes (every member of a group has the same next characters for several
    rounds).  Both take the three-way partition around each group's first key; against the oracle."""
    rng = np.random.default_rng(21)
    s60 = rng.integers(0, 4, 60, dtype=np.uint8)
    copies = np.concatenate([np.concatenate([rng.integers(0, 4, 12, dtype=np.uint8), s60]) for _ in range(30000)])
    t = np.concatenate([random_text(2 << 20, 4, 5), periodic_text(1 << 20, 13, 4, 6), copies,
                        random_text(1 << 20, 4, 7)])
    say(f"large groups partition n={len(t)}")
    sa = engine.build_sa(t, sa_width=8, X=39)
    assert np.array_equal(sa, oracle.dcx_sa(t, X=39))


def test_large_groups_mixed_continuations(engine):
    """Four 50-character words repeated 40,000 times in random order: the large groups of a word's
    occurrences share their next characters inside the word (the partition's all-equal case) and split
    four ways at its end, where most members differ from the first (the partition declines and the
    members are sorted); against the oracle."""
    rng = np.random.default_rng(33)
    words = [rng.integers(0, 4, 50, dtype=np.uint8).tobytes() for _ in range(4)]
    t = np.concatenate([repeat_words_text(words, 40000, 34), random_text(1 << 20, 4, 35)])
    say(f"large groups mixed continuations n={len(t)}")
    sa = engine.build_sa(t, sa_width=8, X=39)
    assert np.array_equal(sa, oracle.dcx_sa(t, X=39))


def _periodic_block(length, period, follower, seed):
    """A periodic run over {0..3} whose period continues with character 1 after `length` characters, then
    `follower`: the run's tied groups compare at its end, all members the same way (ascending positions for
    a follower above 1, descending below)."""
    rng = np.random.default_rng(seed)
    p = rng.integers(0, 4, period, dtype=np.uint8)
    p[length % period] = 1
    return np.concatenate([np.resize(p, length), np.array([follower], dtype=np.uint8)])


@pytest.mark.parametrize("X", [21, 39, 133])
def test_tie_break_monotone_groups(engine, X):
    """Large tied groups already in comparator order, in reverse order, and mixed, in one text: periodic runs
    whose groups all ascend (follower above the period's next character) or all descend (follower below),
    groups just above and below the 16-member threshold, repeated words (groups whose members compare both
    ways, which take the merge sort), and a periodic run reaching the text end (boundary suffixes in its
    groups).  Against libsais."""
    rng = np.random.default_rng(90 + X)
    words = [rng.integers(0, 4, 60, dtype=np.uint8).tobytes() for _ in range(3)]
    parts = [random_text(1 << 18, 4, 91), _periodic_block(40000, 11, 3, 92), random_text(1 << 16, 4, 93),
             _periodic_block(30000, 7, 0, 94), random_text(1 << 16, 4, 95),
             _periodic_block(17 * 23 + 150, 23, 2, 96), random_text(4096, 4, 97),
             _periodic_block(15 * 29 + 150, 29, 0, 98), random_text(4096, 4, 99),
             repeat_words_text(words, 3000, 100), random_text(1 << 16, 4, 101),
             periodic_text(20000, 5, 4, 102)]
    t = np.concatenate(parts)
    say(f"tie-break monotone groups X={X} n={len(t)}")
    sa = engine.build_sa(t, sa_width=8, X=X)
    assert np.array_equal(sa, oracle.libsais_sa(t, threads=8))


def test_dense_codes_text_end(engine):
    # alphabets of 2, 4 and 256 symbols make the merge use dense character codes, where the padding
    # reads like the smallest character; texts ending in long runs of that character exercise the
    # tie-break rule for windows that reach past the end
    cases = []
    for sigma, tail in ((4, 300), (2, 150), (256, 200), (4, 7)):
        t = random_text(50000, sigma, sigma + tail)
        t = np.concatenate([t, np.full(tail, t.min(), dtype=np.uint8)])
        cases.append((f"s{sigma}_tail{tail}", t))
    p = periodic_text(30000, 5, 4, 9)
    cases.append(("period_then_min_run", np.concatenate([p, np.full(500, p.min(), dtype=np.uint8)])))
    for name, t in cases:
        for X in (3, 21, 39, 133):
            for packing in (True, False):
                check_case(engine, name, t, X, packing, packing)


def _batch_cases():
    rng = np.random.default_rng(42)
    texts = []
    for k in range(60):
        n = int(rng.integers(0, 30000)) if k % 7 else int(rng.integers(0, 40))
        sigma = int(rng.choice([2, 4, 5, 26, 200]))
        texts.append(random_text(n, sigma, 1000 + k))
    texts.append(copies_text(50000, 77))
    texts.append(periodic_text(40000, 17, 4, 78))
    texts.append(b"")
    texts.append(b"a")
    return texts


def test_batch_matches_per_text(engine):
    texts = _batch_cases()
    for X in (7, 39):
        for limit in (0, 60000):
            say(f"batch X={X} count={len(texts)} limit={limit}")
            sas = engine.build_sa_batch(texts, X=X, batch_max_chars=limit)
            for k, (t, sa) in enumerate(zip(texts, sas)):
                ref = oracle.dcx_sa(t, X=X)
                assert sa.dtype == np.uint32 and np.array_equal(sa.astype(np.uint64), ref), (X, limit, k)


def test_autotuned_x_batch(engine):
    """The autotuner times real batched builds and returns one of its candidates for X, and for X with the
    sub-batch size; the batches built with the chosen parameters equal the oracle per text (the SAs
    depend on neither)."""
    import torch
    from distributed_suffix_sorting_evaluation_amd.autotune import XTuner, choose_batch_params, choose_x_batch
    texts = [np.frombuffer(t, dtype=np.uint8) if isinstance(t, bytes) else np.asarray(t, dtype=np.uint8)
             for t in _batch_cases()]
    off = np.zeros(len(texts) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([t.size for t in texts])
    dt = torch.from_numpy(np.concatenate(texts)).cuda()
    ds = torch.empty(int(off[-1]), dtype=torch.int32, device="cuda")
    say("autotune X")
    x, ms = choose_x_batch(engine, dt.data_ptr(), off, ds.data_ptr(), candidates=(7, 21, 39), reps=1)
    assert x in (7, 21, 39) and set(ms) == {7, 21, 39} and ms[x] == min(ms.values())
    tuner = XTuner(engine, candidates=(7, 39), reps=1)
    assert tuner.x_for(("acgt", 16), dt.data_ptr(), off, ds.data_ptr()) in (7, 39)
    assert ("acgt", 16) in tuner.choice and tuner.size_for(("acgt", 16)) == 0
    say("autotune X and sub-batch size")
    x2, size, x_ms, size_ms = choose_batch_params(engine, dt.data_ptr(), off, ds.data_ptr(), candidates=(7, 39),
                                                  sizes=(60000, 1 << 24), reps=1)
    assert x2 in (7, 39) and size in (60000, 1 << 24) and size_ms[size] == min(size_ms.values())
    tuner = XTuner(engine, candidates=(39,), reps=1, sizes=(60000, 1 << 24))
    assert tuner.x_for("k", dt.data_ptr(), off, ds.data_ptr()) == 39 and tuner.size_for("k") in (60000, 1 << 24)
    for X, limit in ((x, 0), (x2, size)):
        ds.fill_(-1)
        engine.build_sa_batch_device(dt.data_ptr(), off, ds.data_ptr(), X=X, batch_max_chars=limit)
        got = ds.cpu().numpy().view(np.uint32)
        for k, t in enumerate(texts):
            assert np.array_equal(got[int(off[k]):int(off[k + 1])].astype(np.uint64), oracle.dcx_sa(t, X=39)), (X, k)


def test_batch_byte_alphabet(engine):
    # all 256 byte values: joined characters need a 32-bit text inside the engine
    texts = [random_text(20000, 256, 5), bytes(range(256)) * 30, random_text(5000, 3, 6), b"\x00\x00\x00"]
    sas = engine.build_sa_batch(texts, X=21)
    for t, sa in zip(texts, sas):
        assert np.array_equal(sa.astype(np.uint64), oracle.dcx_sa(t, X=21))


def test_batch_many_short_texts(engine):
    """The per-text partition of a joined SA looks each entry's text up in LDS tables for sub-batches of
    at most 12,288 texts and in a global owner table beyond that: 12,288 and 13,000 short texts (every
    97th one empty) at host and device output, and joined strings short enough for one table entry per
    position."""
    import torch
    rng = np.random.default_rng(7)
    for count in (12288, 13000):
        lens = rng.integers(0, 40, count)
        lens[::97] = 0
        texts = [random_text(int(n), 4, 5000 + k) for k, n in enumerate(lens)]
        refs = [brute_sa(t) for t in texts]
        say(f"batch of {count} short texts, host output")
        sas = engine.build_sa_batch(texts, X=7)
        for k, (sa, ref) in enumerate(zip(sas, refs)):
            assert np.array_equal(sa.astype(np.uint64), ref), (count, k)
        say(f"batch of {count} short texts, device output")
        off = np.zeros(count + 1, dtype=np.uint64)
        off[1:] = np.cumsum(lens)
        dt = torch.from_numpy(np.concatenate(texts)).cuda()
        ds = torch.full((int(off[-1]) + 64,), -1, dtype=torch.int32, device="cuda")
        engine.build_sa_batch_device(dt.data_ptr(), off, ds.data_ptr(), X=7)
        got = ds.cpu().numpy().view(np.uint32)
        for k, ref in enumerate(refs):
            assert np.array_equal(got[int(off[k]):int(off[k + 1])].astype(np.uint64), ref), (count, k)
        assert (got[int(off[-1]):] == 0xFFFFFFFF).all(), "the partition wrote past the batch's SA"
    for texts in ([b"banana", b"", b"mississippi"], [b"a"], [b"", b"ab"], [b"ba", b"b"]):
        say(f"short joined string {texts}")
        for sa, t in zip(engine.build_sa_batch(texts, X=3), texts):
            assert np.array_equal(sa.astype(np.uint64), brute_sa(t)), texts


def test_batch_dense_codes(engine):
    """Batches whose joined alphabet (the separator plus sigma letters) is a power of two are packed with
    dense codes (character c as c - 1, the separator as 0 like the padding): sigma 1, 3 and 7, with
    texts that end in runs of their smallest letter, split into sub-batches of about 20,000 characters."""
    rng = np.random.default_rng(11)
    for sigma in (1, 3, 7):
        texts = []
        for k in range(40):
            t = random_text(int(rng.integers(0, 3000)), sigma, 300 + 40 * sigma + k)
            if k % 3 == 0:
                t = np.concatenate([t, np.zeros(int(rng.integers(1, 60)), dtype=np.uint8)])
            texts.append(t + ord("a"))
        for limit in (0, 20000):
            say(f"dense batch sigma={sigma} limit={limit}")
            sas = engine.build_sa_batch(texts, X=13, batch_max_chars=limit)
            for k, (t, sa) in enumerate(zip(texts, sas)):
                assert np.array_equal(sa.astype(np.uint64), oracle.dcx_sa(t, X=13)), (sigma, limit, k)


def test_batch_empty_texts(engine):
    """Batches of empty texts only (a joined string of separators: short enough for the base case, and
    20,000 separators through the full pipeline with no text bytes to pack or partition), and 16,000 empty
    texts around one short text (the full pipeline, one text to partition)."""
    import torch
    for count in (3, 20000):
        say(f"batch of {count} empty texts")
        sas = engine.build_sa_batch([b""] * count, X=7)
        assert len(sas) == count and all(sa.size == 0 for sa in sas)
        off = np.zeros(count + 1, dtype=np.uint64)
        ds = torch.full((64,), -1, dtype=torch.int32, device="cuda")
        dt = torch.zeros(16, dtype=torch.uint8, device="cuda")
        engine.build_sa_batch_device(dt.data_ptr(), off, ds.data_ptr(), X=7)
        assert (ds.cpu().numpy() == -1).all()
    texts = [b""] * 8000 + [b"abracadabra"] + [b""] * 8000
    sas = engine.build_sa_batch(texts, X=7)
    assert all(sa.size == 0 for k, sa in enumerate(sas) if k != 8000)
    assert np.array_equal(sas[8000].astype(np.uint64), brute_sa(b"abracadabra"))


def test_batch_errors(engine):
    from distributed_suffix_sorting_evaluation_amd import DCXError
    with pytest.raises(DCXError):
        engine.build_sa_batch([b"abc"], X=5)
    assert engine.build_sa_batch([], X=39) == []


def test_same_calls_as_host_abi(engine):
    """dcx_build_sa and dcx_build_sa_batch against the oracle's dcx_host_* restatement of the same C ABI, with
    identical arguments: the same status codes and the same output bytes in every SA width."""
    import ctypes

    from helpers import abi_batch, abi_build, host_abi
    H = host_abi()
    hh = ctypes.c_void_p()
    assert H.dcx_host_create(ctypes.byref(hh), 0, 0) == 0
    try:
        cases = [("three_brothers", read("the_three_brothers.txt")), ("mississippi", b"mississippi"),
                 ("copies", copies_text(50000, 31)), ("periodic", periodic_text(20000, 17, 4, 32)),
                 ("bytes", random_text(30000, 256, 33))]
        for name, t in cases:
            for X, width in ((3, 4), (21, 5), (39, 8), (133, 5)):
                say(f"host-abi case {name} X={X} width={width}")
                o = engine.make_options(X=X, base_case_chars=0)
                g = abi_build(engine._L, "dcx_", engine._h, t, o, width)
                c = abi_build(H, "dcx_host_", hh, t, o, width)
                assert g[0] == c[0] == 0
                assert g[1] == c[1], (name, X, width)
        for kw, width in (({"X": 40}, 8), ({}, 6), ({"discarding_threshold": 1.5}, 8),
                          ({"base_case_chars": (1 << 31) + 1}, 8)):
            o = engine.make_options(**kw)
            say(f"host-abi error case {kw} width={width}")
            assert abi_build(engine._L, "dcx_", engine._h, b"abracadabra", o, width)[0] == \
                abi_build(H, "dcx_host_", hh, b"abracadabra", o, width)[0] != 0
        texts = [random_text(n, 4, 40 + n) for n in (0, 1, 7, 300, 2000, 70000)]
        joined = np.concatenate([np.frombuffer(b"head", dtype=np.uint8)] + texts)
        offsets = np.cumsum([4] + [x.size for x in texts]).astype(np.uint64)
        for X in (7, 39):
            say(f"host-abi batch X={X} offsets[0]=4")
            o = engine.make_options(X=X)
            g = abi_batch(engine._L, "dcx_", engine._h, joined, offsets, o)
            c = abi_batch(H, "dcx_host_", hh, joined, offsets, o)
            assert g[0] == c[0] == 0
            assert np.array_equal(g[1], c[1])
    finally:
        H.dcx_host_destroy(hh)


def test_staged_host_download_all_widths(engine):
    """Host texts and outputs large enough for the pinned-ring upload and download (64 MiB chunks, several host
    threads per chunk, a partial last chunk) give outputs equal to the device output of the same build, byte
    for byte, in widths 4, 5 and 8, twice in a row; a batch whose sub-batch output crosses chunks equals its
    per-text builds."""
    import torch
    n = 20_000_003  # 80, 100 and 160 MB of output: partial last chunks, slices of uneven length
    t = copies_text(n, 21, sigma=4, copy_frac=0.05, max_len=5000)
    dt = torch.from_numpy(t).cuda()
    for width in (4, 5, 8):
        say(f"staged download n={n} width={width}")
        buf = torch.zeros(width * n, dtype=torch.uint8, device="cuda")
        engine.build_sa_device(dt.data_ptr(), n, buf.data_ptr(), width, X=39)
        torch.cuda.synchronize()
        want = buf.cpu().numpy().tobytes()
        got = engine.build_sa(t, sa_width=width, X=39)
        assert got.tobytes() == want, width
        again = engine.build_sa(t, sa_width=width, X=39)
        assert again.tobytes() == want, width
    assert engine.check_sa(t, np.frombuffer(want, dtype="<u8"), sa_width=8)
    texts = [copies_text(m, 50 + m % 97, sigma=4, copy_frac=0.05) for m in (6_000_001, 4_500_000, 7_000_017)]
    say("staged download batch of 3 texts, 17.5 M chars")
    sas = engine.build_sa_batch(texts, batch_max_chars=1 << 31)
    for k, x in enumerate(texts):
        assert np.array_equal(np.asarray(sas[k]).astype(np.uint64), engine.build_sa(x, sa_width=8, X=39)), k


def test_staged_ring_wraps_many_times():
    """With 64 KiB stages (DCX_STAGE_BYTES, read when a handle is created) a 1 MB text goes up in 16 chunks
    and its SA comes down in 62-123 chunks: both rings of four stages wrap many times, every stage is reused,
    and the host threads copy slices of uneven length (n * width is not a multiple of the stage).  The
    outputs, written twice into the same host buffer, equal the device output of the same build."""
    import ctypes
    import os

    import torch
    from distributed_suffix_sorting_evaluation_amd import Engine
    from distributed_suffix_sorting_evaluation_amd import engine as eng_mod
    os.environ["DCX_STAGE_BYTES"] = "65536"
    try:
        eng = Engine(0)
    finally:
        del os.environ["DCX_STAGE_BYTES"]
    try:
        n = 1_000_003
        t = copies_text(n, 7, sigma=4, copy_frac=0.05, max_len=3000)
        dt = torch.from_numpy(t).cuda()
        for width in (4, 5, 8):
            say(f"staged ring 64 KiB n={n} width={width}")
            buf = torch.zeros(width * n, dtype=torch.uint8, device="cuda")
            eng.build_sa_device(dt.data_ptr(), n, buf.data_ptr(), width, X=39)
            torch.cuda.synchronize()
            want = buf.cpu().numpy().tobytes()
            out = np.full(width * n, 0xAB, dtype=np.uint8)
            for _ in range(2):
                o = eng.make_options(X=39)
                rc = eng._L.dcx_build_sa(eng._h, t.ctypes.data, n, ctypes.byref(o), out.ctypes.data, width,
                                         ctypes.byref(eng_mod.Stats()))
                assert rc == 0
                assert out.tobytes() == want, width
                out[::4097] ^= 0x5A  # disturb the buffer before the second call
        say("staged ring 64 KiB batch")
        texts = [copies_text(m, 60 + m % 13, sigma=4, copy_frac=0.05, max_len=2000) for m in (300_001, 250_000, 410_007)]
        sas = eng.build_sa_batch(texts, batch_max_chars=1 << 31)
        for k, x in enumerate(texts):
            assert np.array_equal(np.asarray(sas[k]).astype(np.uint64), oracle.libsais_sa(x)), k
    finally:
        eng.close()
