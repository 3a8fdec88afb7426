/*
 * ggks.h -- C ABI of the MI355X (gfx950) selection library libggks.so.
 *
 * Order statistics on device-resident vectors: bucket multiselect, MSD radix
 * multiselect and top-K, for float32, float64, uint32 and int32.
 *
 * Conventions shared by every entry point (they follow the reference):
 *  - k values are 1-based ranks of the k-th LARGEST element: k = 1 is the
 *    maximum, k = n the minimum (bucketMultiselect.cu:845-846,
 *    multiselectTimingFunctions.cu:41-48).  k lists may be unsorted and may
 *    repeat; outputs follow the caller's k order (bucketMultiselect.cu:615-628,
 *    :313-325).
 *  - Top-K returns the K largest values in descending order
 *    (multiselectTimingFunctions.cu:155-177).
 *  - Order is the order of the order-preserving unsigned key conversion
 *    (keyConversion.cu:8-42), i.e. the reference's radix order: IEEE order
 *    except that -0.0 sorts before +0.0.  NaN inputs are outside the contract.
 *  - The input vector is never modified (the reference's radix paths overwrite
 *    it, merrillSelect.cu:210-259; opts->destructive is accepted and ignored).
 *  - Return 0 on success or a negative GGKS_E* code; outputs are untouched on
 *    error.  ggks_last_error() gives the message of the calling thread's last
 *    failure.
 *  - n must satisfy 1 <= n < 2^32 (the reference uses 32-bit lengths).
 *
 * The scratch the kernels need comes from a per-context arena that grows on
 * demand and is reused across calls; ggks_scratch_bytes() reports the size
 * for admission control.  Calls on one context are serialised by a mutex.
 */
#ifndef GGKS_H
#define GGKS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ggks_dtype { GGKS_F32 = 0, GGKS_F64 = 1, GGKS_U32 = 2, GGKS_I32 = 3 };

enum ggks_status {
  GGKS_OK = 0,
  GGKS_EINVAL = -1,      /* bad argument: n < 1, k outside [1, n], null ptr */
  GGKS_ENOMEM = -2,      /* device or pinned memory could not be allocated */
  GGKS_EDEVICE = -3,     /* a HIP call failed (see ggks_last_error) */
  GGKS_EUNSUPPORTED = -4 /* argument combination not supported */
};

enum ggks_kind {
  GGKS_KIND_BUCKET = 0, /* bucket multiselect */
  GGKS_KIND_RADIX = 1,  /* MSD radix multiselect */
  GGKS_KIND_TOPK = 2    /* top-K (descending) */
};

/* Options; a NULL pointer means all defaults (zeros). */
typedef struct ggks_opts {
  int32_t blocks;       /* chunk count hint for the streaming passes; 0 = tuned */
  int32_t threads;      /* accepted for interface parity (reference passes the
                           device's maxThreadsPerBlock); kernels use 1024 */
  int32_t buckets_log2; /* bucket budget as log2 (8..15); 0 = tuned */
  int32_t destructive;  /* accepted; the input is never modified */
  uint64_t seed;        /* seed of the splitter sample; 0 = fixed default */
  int32_t radix_bits;   /* radix path: digit width per level (4..14); 0 = 14 */
  int32_t level2;       /* second level for children of more than 2^16 keys
                           (re-bucketed by the whole GPU): 0 = radix plans
                           only, 1 = every plan, -1 = never */
  int32_t reserved[4];
} ggks_opts;

typedef struct ggks_ctx ggks_ctx;

/* ---- contexts ------------------------------------------------------ */
/* stream: a hipStream_t (NULL = the device's null stream). */
int ggks_ctx_create(ggks_ctx **ctx, int device, void *stream);
int ggks_ctx_destroy(ggks_ctx *ctx);
/* Bytes of device scratch one call of `kind` needs (admission control). */
size_t ggks_scratch_bytes(int kind, int dtype, int64_t n, int64_t num_k);
const char *ggks_last_error(void);
const char *ggks_version(void);

/* ---- single vector, host k list / host outputs ----------------------
 * Replaces BucketMultiselect::bucketMultiselectWrapper
 * (bucketMultiselect.cu:837-871): device vector, length, host k array,
 * host output array.  ctx may be NULL (per-device default context, current
 * device, null stream). */
int ggks_bucket_multiselect(ggks_ctx *ctx, int dtype, const void *d_vec,
                            int64_t n, const uint32_t *h_k, int64_t num_k,
                            void *h_out, const ggks_opts *opt);
int ggks_bucket_multiselect_f32(const float *d_vec, int64_t n,
                                const uint32_t *h_k, int32_t num_k,
                                float *h_out, const ggks_opts *opt);
int ggks_bucket_multiselect_f64(const double *d_vec, int64_t n,
                                const uint32_t *h_k, int32_t num_k,
                                double *h_out, const ggks_opts *opt);
int ggks_bucket_multiselect_u32(const uint32_t *d_vec, int64_t n,
                                const uint32_t *h_k, int32_t num_k,
                                uint32_t *h_out, const ggks_opts *opt);
int ggks_bucket_multiselect_i32(const int32_t *d_vec, int64_t n,
                                const uint32_t *h_k, int32_t num_k,
                                int32_t *h_out, const ggks_opts *opt);

/* Replaces RadixSelect::RadixSelectWrapper (merrillSelect.cu:210-259) and
 * InplaceRadix::inplaceRadixSelectWrapper (inplaceRadixSelect.cu:83-108),
 * generalised to many k: MSD digits of the converted key, no sampling. */
int ggks_radix_multiselect(ggks_ctx *ctx, int dtype, const void *d_vec,
                           int64_t n, const uint32_t *h_k, int64_t num_k,
                           void *h_out, const ggks_opts *opt);
/* Typed forms on the default context (the reference's per-type overloads,
 * merrillSelect.cu:210-259 (uint, float), inplaceRadixSelect.cu:83-108 (uint,
 * float, double)); i32 through the same key conversion. */
int ggks_radix_multiselect_f32(const float *d_vec, int64_t n,
                               const uint32_t *h_k, int32_t num_k,
                               float *h_out, const ggks_opts *opt);
int ggks_radix_multiselect_f64(const double *d_vec, int64_t n,
                               const uint32_t *h_k, int32_t num_k,
                               double *h_out, const ggks_opts *opt);
int ggks_radix_multiselect_u32(const uint32_t *d_vec, int64_t n,
                               const uint32_t *h_k, int32_t num_k,
                               uint32_t *h_out, const ggks_opts *opt);
int ggks_radix_multiselect_i32(const int32_t *d_vec, int64_t n,
                               const uint32_t *h_k, int32_t num_k,
                               int32_t *h_out, const ggks_opts *opt);

/* Replaces randomizedTopkSelectWrapper (randomizedTopkSelect.cu:210-213) and
 * timeBucketTopkselect (multiselectTimingFunctions.cu:204-231): the K largest
 * values, descending, into the caller's host buffer h_out[K]. */
int ggks_topk(ggks_ctx *ctx, int dtype, const void *d_vec, int64_t n,
              int64_t K, void *h_out, const ggks_opts *opt);
/* Typed forms on the default context (randomizedTopkSelectWrapper<T>,
 * randomizedTopkSelect.cu:210-213, for T = float, double, uint; plus int). */
int ggks_topk_f32(const float *d_vec, int64_t n, int32_t K, float *h_out,
                  const ggks_opts *opt);
int ggks_topk_f64(const double *d_vec, int64_t n, int32_t K, double *h_out,
                  const ggks_opts *opt);
int ggks_topk_u32(const uint32_t *d_vec, int64_t n, int32_t K, uint32_t *h_out,
                  const ggks_opts *opt);
int ggks_topk_i32(const int32_t *d_vec, int64_t n, int32_t K, int32_t *h_out,
                  const ggks_opts *opt);

/* The reference's sort-and-choose (multiselectTimingFunctions.cu:41-95,
 * timeSortAndChooseMultiselect) on the device: a full radix sort of a copy of
 * the vector (rocPRIM), then h_out[i] = sorted[n - h_k[i]].  The baseline
 * bucket multiselect is raced against (analyzeMultiselect.cu:209-274); the
 * input is left intact. */
int ggks_sort_choose(ggks_ctx *ctx, int dtype, const void *d_vec, int64_t n,
                     const uint32_t *h_k, int64_t num_k, void *h_out);

/* ---- stream-ordered variants: device k list and device outputs ------
 * Nothing is synchronised and no host memory is touched, so these can be
 * captured in a hipGraph once the context's arena is large enough (call
 * ggks_ctx_reserve first).  An invalid k is reported by
 * ggks_ctx_check() after the stream has drained. */
int ggks_bucket_multiselect_async(ggks_ctx *ctx, int dtype, const void *d_vec,
                                  int64_t n, const uint32_t *d_k,
                                  int64_t num_k, void *d_out,
                                  const ggks_opts *opt);
int ggks_radix_multiselect_async(ggks_ctx *ctx, int dtype, const void *d_vec,
                                 int64_t n, const uint32_t *d_k, int64_t num_k,
                                 void *d_out, const ggks_opts *opt);
int ggks_topk_async(ggks_ctx *ctx, int dtype, const void *d_vec, int64_t n,
                    int64_t K, void *d_out, const ggks_opts *opt);
/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): bucket multiselect as ggks_bucket_multiselect_async (device k
 * list, device outputs, stream-ordered on `stream`, a hipStream_t or NULL),
 * but with no context and no allocation of any kind on any call: its scratch
 * is a static device array of GGKS_NOALLOC_SCRATCH_BYTES that every GPU the
 * library runs on holds until the process ends, and its plan tables are
 * written there by a kernel (no host staging buffers).  A select whose
 * scratch exceeds that size is refused with GGKS_ENOMEM before anything is
 * launched (with 1024 k: float32 / uint32 / int32 up to about 1.6e7 keys,
 * float64 up to about 8e6).  Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (other streams, other host threads) race: serialise
 * them.  k values must lie in [1, n]; an out-of-range device k is clamped
 * to 1 and not reported (there is no context to report it to). */
#define GGKS_NOALLOC_SCRATCH_BYTES (64ull << 20)
int ggks_bucket_multiselect_noalloc(int dtype, const void *d_vec, int64_t n,
                                    const uint32_t *d_k, int64_t num_k,
                                    void *d_out, const ggks_opts *opt,
                                    void *stream);
/* Grow the arena so that a call of `kind` with these sizes allocates
 * nothing. */
int ggks_ctx_reserve(ggks_ctx *ctx, int kind, int dtype, int64_t n,
                     int64_t num_k);
/* After the stream drained: GGKS_OK, or GGKS_EINVAL if a device k list held
 * a k outside [1, n] since the last check (the flag is then cleared). */
int ggks_ctx_check(ggks_ctx *ctx);
void *ggks_ctx_stream(ggks_ctx *ctx);

/* ---- inspection (tests and diagnostics) -----------------------------
 * Device addresses of the intermediate tables of the last engine select on
 * a context (valid after its stream drained, until the next call on the
 * context): the segment descriptor and state (SegState of the bucket
 * function), its LUT, the chunk table, the per-chunk rows (after the reduce: each chunk's
 * exclusive offset inside every bucket), the bucket totals (prefix inside
 * each group of 64 buckets) and group sums, the active-bucket bitmap, the
 * child records and their queries, and the reduced buffer.  Layouts are
 * csrc/ggks_types.hpp's; direct = 1 means the last select had no engine
 * segment (no tables). */
typedef struct ggks_debug_regions {
  const void *segs, *states, *lut, *chunks, *rows, *totals, *gsums, *bits, *children, *cq, *red;
  uint64_t n;
  uint32_t nchunks, bmax, nq, pbits;
  int32_t dtype, direct;
} ggks_debug_regions;
int ggks_ctx_debug_regions(ggks_ctx *ctx, ggks_debug_regions *out);

/* ---- stage timing ---------------------------------------------------
 * When armed with nslots > 0, every select on the context records HIP events
 * on its stream around each kernel stage into slot (call index % nslots).
 * After the stream drained, ggks_ctx_profile_read() fills
 * ms[slot * (GGKS_NSTAGES + 1) + stage] (stage durations) and
 * ms[slot * (GGKS_NSTAGES + 1) + GGKS_NSTAGES] (whole call, first event to
 * last).  Stages: 0 queries, 1 sample, 2 lut, 3 count (pass 1), 4 reduce,
 * 5 findk, 6 compact (pass 2), 7 finish.  nslots = 0 disarms. */
#define GGKS_NSTAGES 8
int ggks_ctx_profile_arm(ggks_ctx *ctx, int nslots);
int ggks_ctx_profile_read(ggks_ctx *ctx, float *ms, int nslots);
/* Record only the events in `mask` (bit m: the event before stage m; bit
 * GGKS_NSTAGES: the call's last event; default all).  Every event costs
 * device time between the kernels it separates, so a timed run records only
 * the two around the kernel it times.  ggks_ctx_profile_read() then reports
 * for a recorded event m the time to the next recorded one in ms[.. + m],
 * 0 for the others, and the first to the last recorded event as the whole
 * call.  Takes effect at the next ggks_ctx_profile_arm(), which latches the
 * mask for the selects it times and for their read.
 * With GGKS_PROFILE_KERNEL_EVENTS set, the events of the single-kernel
 * stages 3 (count) and 6 (compact) are not recorded between kernels but
 * passed to that kernel's launch as its start and stop events
 * (hipExtLaunchKernel), so they add no marker between the kernels and time
 * the kernel alone; events of the other stages are then not recorded. */
#define GGKS_PROFILE_KERNEL_EVENTS 0x80000000u
int ggks_ctx_profile_marks(ggks_ctx *ctx, uint32_t mask);

/* ---- device plumbing for hosts without a GPU framework --------------
 * Thin wrappers over the HIP runtime this library is linked with. */
int ggks_dev_alloc(void **ptr, size_t bytes);
int ggks_dev_free(void *ptr);
int ggks_h2d(void *dst, const void *src, size_t bytes);
int ggks_d2h(void *dst, const void *src, size_t bytes);
int ggks_d2d(void *dst, const void *src, size_t bytes);
int ggks_sync(void);
int ggks_device_count(int *count);
int ggks_set_device(int device);
/* Fill a device buffer with a byte pattern (L2/MALL flush helper). */
int ggks_dev_fill(void *ptr, int value, size_t bytes);

/* ---- batched selection over a segment table --------------------------
 * No reference counterpart (the reference selects one vector per call).
 * Many independent vectors of uneven length, each with its own k list or
 * top-K, selected together: one launch per engine kernel for all segments
 * (plus one finisher launch for every small segment), optionally replayed
 * from a captured hipGraph.  Every output equals what the single-vector
 * entry point returns for that segment alone.
 *
 * A batch vector holds all segments (one dtype); segment i is
 * vec[offset, offset + length).  Multiselect segments (kind BUCKET or RADIX)
 * take k values h_k[k_offset, k_offset + num_k) (1 <= k <= length) and write
 * out[out_offset + j] = the k_j-th largest; top-K segments (kind TOPK,
 * num_k = K <= length) write the K largest, descending, from out_offset.
 * Segments with num_k = 0 are skipped.  Output ranges must not overlap. */
typedef struct ggks_segment {
  uint64_t offset;     /* first element of the segment in the batch vector */
  uint64_t length;     /* elements, 1 <= length < 2^32 */
  uint64_t k_offset;   /* multiselect: first of its k values in h_k */
  uint64_t out_offset; /* first output slot; out_offset + num_k <= 2^32 */
  uint32_t num_k;      /* multiselect: number of k values; top-K: K */
  uint32_t kind;       /* GGKS_KIND_BUCKET, GGKS_KIND_RADIX or GGKS_KIND_TOPK */
} ggks_segment;

typedef struct ggks_batch ggks_batch;

typedef struct ggks_batch_stats {
  uint64_t elements;        /* sum of the lengths of segments with queries */
  uint64_t queries;         /* outputs written per run */
  uint64_t scratch_bytes;   /* device bytes of tables + scratch */
  uint32_t engine_segments; /* segments bucketed by the level kernels */
  uint32_t direct_segments; /* small segments finished whole in LDS */
  uint32_t chunks;          /* work-groups of each streaming pass */
  uint32_t graph;           /* 1 once a hipGraph has been captured */
} ggks_batch_stats;

#define GGKS_BATCH_GRAPH 1 /* ggks_batch_run flag: capture once, then replay */

/* Plan a batch and upload its tables (the k lists are copied; h_k may be
 * freed afterwards).  d_scratch: device memory of at least
 * ggks_batch_scratch_bytes() bytes, or NULL to let the batch allocate its
 * own.  stream: hipStream_t to run on, or NULL for a stream of its own (a
 * graph cannot be captured on the null stream).  The band tuning table
 * (ggks_tuning_set) in force at creation decides each segment's bucket
 * budget, chunk size and direct / engine route. */
int ggks_batch_create(ggks_batch **batch, int device, int dtype, const ggks_segment *segs,
                      int64_t num_segments, const uint32_t *h_k, int64_t total_k,
                      const ggks_opts *opt, void *d_scratch, size_t scratch_bytes,
                      void *stream);
/* 0 on a bad table (ggks_last_error tells why). */
size_t ggks_batch_scratch_bytes(int dtype, const ggks_segment *segs, int64_t num_segments,
                                const uint32_t *h_k, int64_t total_k, const ggks_opts *opt);
/* Enqueue one run on the batch's stream (asynchronous).  d_vec: the batch
 * vector; d_out: device outputs.  flags: GGKS_BATCH_GRAPH replays a hipGraph
 * captured on the first run with these two pointers (recaptured when they
 * change).  Runs of one batch share its scratch: they are stream-ordered. */
int ggks_batch_run(ggks_batch *batch, const void *d_vec, void *d_out, int flags);
/* Wait for the batch's stream; GGKS_EINVAL if a kernel flagged bad input. */
int ggks_batch_sync(ggks_batch *batch);
void *ggks_batch_stream(ggks_batch *batch);
int ggks_batch_info(ggks_batch *batch, ggks_batch_stats *info);
int ggks_batch_destroy(ggks_batch *batch);

/* Band tuning table used by batch planning, per dtype and floor(log2 length):
 * bucket budget 2^buckets_log2 (8..15), chunk of 2^chunk_log2 elements
 * (10..26) per work-group of the streaming passes, and direct = 1 to finish
 * segments of the band whole in LDS when they fit.  The autotuner of the
 * Python package measures and sets these; ggks_tuning_reset restores the
 * built-in defaults.  Not synchronised with concurrent batch creation. */
int ggks_tuning_set(int dtype, int log2_len, int buckets_log2, int chunk_log2, int direct);
int ggks_tuning_get(int dtype, int log2_len, int *buckets_log2, int *chunk_log2, int *direct);
/* Single-vector table, per dtype and floor(log2 n): bucket budget 2^buckets_log2
 * (8..15) and the number of chunks (work-groups of the streaming passes,
 * 1..512) the single-vector entry points plan with when ggks_opts leaves
 * them 0.  The autotuner measures and sets these.  ggks_tuning_reset
 * restores the built-in defaults of both tables. */
int ggks_tuning_single_set(int dtype, int log2_len, int buckets_log2, int chunks);
int ggks_tuning_single_get(int dtype, int log2_len, int *buckets_log2, int *chunks);
int ggks_tuning_reset(void);

/* ---- serving loop ----------------------------------------------------
 * No reference counterpart.  A server keeps select requests over vectors
 * resident in HBM and runs them by continuous batching: whenever one of its
 * max_inflight streams is free, the oldest pending request and the later
 * pending requests of the same dtype (up to max_batch) are planned as one
 * batch and admitted if the batch's scratch fits in the unused part of the
 * scratch budget (the group is halved until it does; otherwise it waits for
 * running batches to release memory).  Batches on different streams overlap
 * on the device.  Every output equals the single-vector entry point's for
 * that request.  One host thread drives a server; scheduling happens inside
 * submit, poll, drain and run_trace. */
typedef struct ggks_request {
  uint64_t offset;     /* first element of the vector in its dtype's resident buffer */
  uint64_t length;     /* elements, 1 <= length < 2^32 */
  const uint32_t *h_k; /* multiselect: num_k k values, copied at submission; top-K: unused */
  void *h_out;         /* host outputs (num_k values of the dtype), written on completion */
  uint64_t tag;        /* reported by ggks_server_poll */
  uint32_t num_k;      /* multiselect: number of k values; top-K: K */
  int32_t kind;        /* GGKS_KIND_BUCKET, GGKS_KIND_RADIX or GGKS_KIND_TOPK */
  int32_t dtype;       /* selects the resident buffer */
  int32_t reserved;
} ggks_request;

typedef struct ggks_server_stats {
  uint64_t submitted, completed, rejected, batches;
  uint64_t elements;           /* sum of lengths of completed requests */
  uint64_t max_pending;        /* deepest admission queue seen */
  uint64_t pending;            /* queue depth now */
  uint64_t scratch_budget;     /* bytes */
  uint64_t scratch_in_use;     /* bytes now */
  uint64_t peak_scratch_bytes; /* most bytes in use after an admission */
  uint64_t failed;             /* requests of batches whose kernels reported an
                                  internal error (latency -1, output untouched) */
} ggks_server_stats;

typedef struct ggks_server ggks_server;

/* resident[d] / resident_len[d]: device buffer and element count holding the
 * vectors of dtype d (NULL for a dtype not served).  scratch_budget: device
 * bytes the server allocates once and admits batches into.  max_inflight:
 * concurrent batches (streams), 1..16. */
int ggks_server_create(ggks_server **server, int device, const void *const *resident,
                       const uint64_t *resident_len, size_t scratch_budget, int max_inflight,
                       int max_batch);
/* Queue a request (validated now; GGKS_ENOMEM if it alone exceeds the budget). */
int ggks_server_submit(ggks_server *server, const ggks_request *req);
/* Schedule, then report up to max_done completed requests (tag and latency in
 * seconds from submission, or -1 for a request whose batch reported an
 * internal error; its output is untouched); returns how many were written.
 * Request slots are recycled once reported, so a long-running server's host
 * memory stays bounded by the requests in flight. */
int ggks_server_poll(ggks_server *server, uint64_t *tags, double *latency_s, int max_done);
/* Run until nothing is pending or in flight. */
int ggks_server_drain(ggks_server *server);
/* Replay an open-loop arrival trace (arrival_s ascending, seconds from the
 * call): request i is submitted at arrival_s[i]; latency_s[i] = completion
 * seen - arrival_s[i], or -1 if the request was refused or failed. */
int ggks_server_run_trace(ggks_server *server, const ggks_request *reqs, const double *arrival_s,
                          int64_t n, double *latency_s);
int ggks_server_stats_get(ggks_server *server, ggks_server_stats *stats);
int ggks_server_destroy(ggks_server *server);

/* ---- seeded synthetic inputs on the device ------------------------------
 * Replaces the reference's cuRAND problem generators (generateProblems.cu:
 * 25-428, clock-seeded there).  Element i is drawn from output start + i of
 * the SplitMix64 stream of `seed`.  Families (integer dtypes take 0, 2, 5,
 * 10 and 11; float dtypes all):
 *   0 uniform: f32 on the 2^-24 grid of [0,1), f64 on the 2^-53 grid,
 *     u32/i32 uniform in [lo, hi] (bit-identical to the Python package's host
 *     generators)
 *   1 normal(0,1) by Box-Muller in double (the device's libm may differ from
 *     the host's in the last bit)
 *   2 levels: floats holding the integers lo..hi (0 <= lo <= hi < 2^32),
 *     heavy duplicates (integers: as family 0)
 *   3 Cauchy(0,1), tan(pi (u - 1/2)) clipped to finite values
 *   4 plateau: 0.5 with probability 1/2, else uniform [0,1)
 *   5 sorted: position j (= start + i) of a non-decreasing sequence of
 *     length hi, floor((j 2^32 + r_j) / hi) with r_j < 2^32 from the stream
 *     (i32: minus 2^31; floats: times 2^-32); needs start + n <= hi
 *     (generateProblems.cu:34-45)
 *   6 bucket killer: masked uniform bits (f32 & 0x80EFFFFF, f64 &
 *     0x800FFFFFFFFFFFFF) with positions 0..252 (f64: 0..2044) holding the
 *     powers of two 2^-126..2^126 (2^-1022..2^1022) (:344-367, :385-423)
 *   7 noisy: 1 with probability 1/5 else 0, plus N(0, 0.01) (:249-277)
 *   8 half-normal |N(0,1)| (:322-341)
 *   9 huge: uniform * 10^6 (f32 product in float) (:287-298)
 *  10 ones: every value 1 (:180-197)
 *  11 ones-twos: 1 or 2 with probability 1/2 (:154-177)
 *  12 normal(0, 100^2) (:300-306)
 * Synchronous. */
int ggks_generate(void *d_ptr, int dtype, int family, uint64_t seed, uint64_t start, uint64_t n,
                  int64_t lo, int64_t hi);

#ifdef __cplusplus
}
#endif
#endif /* GGKS_H */
