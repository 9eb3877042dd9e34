/*
 * dcx.h -- C ABI of the MI355X DCX suffix-array engine.
 *
 * The engine replaces the `lightweight_dcx` cli path
 *     cli main -> compress_alphabet -> select_packed_dcx_variant -> PDCX<...>::compute_sa
 * (lightweight_dcx/src/executables/cli.cpp:417-689, lightweight_dcx/src/pdcx/pdcx.hpp:529-941)
 * with a single-GPU, in-HBM pipeline of hand-written gfx950 kernels.  Plain pointers and sizes
 * only; no exceptions or exit() cross this boundary -- every entry point returns a status code.
 */
#ifndef DCX_H
#define DCX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference prints a message and calls exit(1) on these conditions
 * (cli.cpp:394-399 missing file, cli.cpp:440-446 alphabet, merge_samples.hpp:892-900). */
#define DCX_OK 0
#define DCX_E_INVALID_X (-1)   /* X not in {3,7,13,21,31,39,57,73,91,95,133}              */
#define DCX_E_ALPHABET (-2)    /* alphabet cannot be represented                            */
#define DCX_E_NOMEM (-3)       /* device allocation failed                                  */
#define DCX_E_HIP (-4)         /* HIP runtime error                                         */
#define DCX_E_IO (-5)          /* file I/O error (CLI)                                      */
#define DCX_E_ARG (-6)         /* invalid argument                                          */
#define DCX_E_WIDTH (-7)       /* sa_width_bytes cannot hold the text length                */
#define DCX_E_ADMISSION (-8)   /* request exceeds the handle's HBM budget                   */
#define DCX_E_UNSUPPORTED (-9) /* size outside what this build supports                     */
#define DCX_E_CANCELLED (-10)  /* request cancelled                                         */
#define DCX_E_NO_GPU (-11)     /* no HIP device available                                   */

#define DCX_MAX_LEVELS 32

typedef struct dcx_handle dcx_handle;

/* Options of one SA build.  Mirrors PDCXConfig (lightweight_dcx/src/pdcx/config.hpp:27-56) and the
 * CLI flags of cli.cpp:149-305 that can influence the computation.  MPI-only flags have no
 * counterpart: the SA is unique, so they cannot change the output. */
typedef struct dcx_options {
    uint32_t X;                   /* --dcx dcX                       (cli.cpp:180-182)       */
    uint32_t pack_samples;        /* --use-char-packing-samples      (cli.cpp:231-234)       */
    uint32_t pack_merging;        /* --use-char-packing-merging      (cli.cpp:235-238)       */
    uint32_t text_on_device;      /* `text` is a device pointer                              */
    uint32_t sa_on_device;        /* `sa_out` is a device pointer                            */
    uint32_t profile_kernels;     /* time every radix pass with hipEvents (fills sort_* stats) */
    double discarding_threshold;  /* --discarding-threshold (0.7, config.hpp:42)             */
    uint64_t base_case_chars;     /* base case: n <= max(2X, this), at most 2^31 (pdcx.hpp:623) */
    uint64_t seed;                /* --seed (accepted; the SA does not depend on it)         */
    uint64_t batch_max_chars;     /* dcx_build_sa_batch: characters per sub-batch (0 = 2^24)  */
    uint64_t reserved[5];
} dcx_options;

typedef struct dcx_level_stats {
    uint64_t n;             /* characters at this recursion level                       */
    uint64_t samples;       /* difference-cover samples incl. the dummy (pdcx.hpp:473)   */
    uint64_t names;         /* highest lexicographic name = last_rank (pdcx.hpp:743)      */
    uint64_t reduced;       /* ranks kept after discarding, 0 if not used (pdcx.hpp:215)  */
    uint32_t base_case;     /* level solved by the sequential base case                  */
    uint32_t recursed;      /* names were not distinct -> recursive call                 */
    uint32_t discarding;    /* the recursive call used discarding                         */
    uint32_t key_chars;     /* characters in a sample key                                */
    uint32_t merge_chars;   /* characters compared before ranks in the merge             */
    uint32_t bits_per_char; /* packing width (level 0 with packing), else 0              */
} dcx_level_stats;

typedef struct dcx_stats {
    uint32_t levels;
    uint32_t sigma;               /* input alphabet size (compress_alphabet, cli.cpp:417)  */
    dcx_level_stats level[DCX_MAX_LEVELS];
    double ms_total;              /* host wall time of the call                             */
    double ms_device;             /* device time text-in-HBM -> SA-in-HBM (hipEvents)       */
    double ms_phase[8];           /* level-0 phases, see DCX_PHASE_*                        */
    uint64_t alg_bytes;           /* algorithmic bytes over all levels                      */
    uint64_t peak_device_bytes;   /* peak device memory held by the engine during the call   */
    /* filled when profile_kernels is set: LSD radix passes (one kernel launch each) */
    uint64_t sort_passes;         /* number of pass launches                                  */
    uint64_t sort_pass_items;     /* sum over launches of the (key, value) pairs moved        */
    double sort_pass_ms;          /* sum over launches of the kernel time                     */
    uint64_t sort_pass_bytes;     /* sum over launches of the algorithmic bytes: keys and values
                                     read and written (a first pass keyed by the character stream
                                     reads the stream; identity values are not read)          */
} dcx_stats;

/* ms_phase indices; the names follow the reference's timer keys. */
#define DCX_PHASE_PREPARATION 0 /* phase_00_preparation: histogram, alphabet, H2D          */
#define DCX_PHASE_SAMPLES 1     /* phase_01_02_samples_ranks: sample sort and names          */
#define DCX_PHASE_RECURSION 2   /* phase_03_recursion                                       */
#define DCX_PHASE_MERGE 3       /* phase_04_merge                                           */
#define DCX_PHASE_OUTPUT 4      /* SA width conversion and D2H                              */

/* Fills the reference defaults: X = 39, packing on, discarding_threshold = 0.7,
 * base_case_chars = 10000, host pointers. */
void dcx_default_options(dcx_options* opt);

/* Creates an engine bound to HIP device `device` with its own stream.  hbm_budget_bytes = 0 means
 * no limit; otherwise requests whose footprint estimate exceeds it fail with DCX_E_ADMISSION.
 * Replaces the PDCX constructor (pdcx.hpp:67-80). */
int dcx_create(dcx_handle** out, int device, uint64_t hbm_budget_bytes);

/* Releases the handle and all device memory it caches. */
int dcx_destroy(dcx_handle* h);

/* Suffix array of text[0..n) in standard order (a proper prefix sorts first), 0-based positions,
 * written as sa_width_bytes in {4, 5, 8} little-endian integers (5 = uint40, the reference's
 * index type, cli.cpp:107, uint_types.hpp:262).  Replaces PDCX::compute_sa (pdcx.hpp:529) together
 * with compress_alphabet (cli.cpp:417-463) and the variant dispatch (cli.cpp:475-653).  The text is
 * not modified.  `stats` may be NULL.  A handle is not reentrant; different handles may be used
 * from different threads.  n may be at most 2^32 (32-bit positions inside the engine, 64-bit counts);
 * longer texts return DCX_E_UNSUPPORTED. */
int dcx_build_sa(dcx_handle* h, const uint8_t* text, uint64_t n, const dcx_options* opt, void* sa_out,
                 uint32_t sa_width_bytes, dcx_stats* stats);

/* Suffix arrays of `count` independent texts in one call (batched mode).  Text k is
 * texts[offsets[k] .. offsets[k+1]) (offsets has count + 1 entries, host memory, non-decreasing, each text
 * at most 2^32 - 1 characters: with its separator it must fit 32-bit positions); its SA is written to sa_out[offsets[k] - offsets[0] .. offsets[k+1] - offsets[0])
 * as 32-bit positions local to the text (sa_out holds offsets[count] - offsets[0] entries).  The texts are joined with a separator symbol smaller than every character and run through one
 * DCX pipeline per sub-batch (at most batch_max_chars characters); a stable radix partition by text then
 * yields the per-text suffix arrays.  With several sub-batches, two are built at once: a second stream of
 * the handle (created on first use, kept until dcx_destroy) takes every other one in a host thread of the
 * library's own; the call returns when both are done.  No counterpart in the reference: it builds one SA
 * per cli run (cli.cpp:747-820); this entry point serves callers with many texts at once. */
int dcx_build_sa_batch(dcx_handle* h, const uint8_t* texts, const uint64_t* offsets, uint32_t count,
                       const dcx_options* opt, uint32_t* sa_out, dcx_stats* stats);

/* Limits of dcx_build_sa_noalloc: the longest text, and the device array per GPU it builds in. */
#define DCX_NOALLOC_MAX_CHARS (1ull << 20)
#define DCX_NOALLOC_SCRATCH_BYTES (128ull << 20)

/* dcx_build_sa for a device-resident text and a device output buffer (opt's pointer flags are ignored:
 * both pointers are taken as device pointers) that allocates nothing: every device buffer of the build is
 * carved from one array of DCX_NOALLOC_SCRATCH_BYTES that the library holds per GPU (from its first use on
 * that GPU until the process ends), the lookback state was sized by dcx_create, and no event, stream or
 * pinned buffer is created.  Texts of at most DCX_NOALLOC_MAX_CHARS characters; longer ones return
 * DCX_E_UNSUPPORTED.  opt->profile_kernels is ignored.  Calls on one GPU share that array, so they are
 * serialized: a call waits until the one before it on that GPU has finished on the device.  Output and
 * status codes are those of dcx_build_sa; a build that would not fit the array returns DCX_E_NOMEM.
 * No counterpart in the reference: it is for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc). */
int dcx_build_sa_noalloc(dcx_handle* h, const uint8_t* text, uint64_t n, const dcx_options* opt, void* sa_out,
                         uint32_t sa_width_bytes, dcx_stats* stats);

/* Validity check of a suffix array on the device, after check_suffixarray
 * (lightweight_dcx/src/sa_check.hpp:56-184): permutation test plus non-decreasing
 * (T[SA[k]], rank[SA[k]+1]).  *ok = 1 if sa is the SA of text.  `on_device` flags both pointers as
 * device pointers. */
int dcx_check_sa(dcx_handle* h, const uint8_t* text, uint64_t n, const void* sa, uint32_t sa_width_bytes,
                 int on_device, int* ok);

/* Human-readable message for a status code (the reference prints these on rank 0). */
const char* dcx_strerror(int code);

/* Detail of the last error on this handle (HIP error text, offending size ...). */
const char* dcx_last_error(dcx_handle* h);

#ifdef __cplusplus
}
#endif

#endif /* DCX_H */
