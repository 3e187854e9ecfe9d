/*
 * stntt.h -- C ABI of the MI355X Stockham NTT library (libstntt.so).
 *
 * Forward and inverse number-theoretic transforms, natural order in and out,
 * over
 *   - the reference's 32-bit prime field: runtime (root, mod) as produced by
 *     findParams(n, min_mod), mod < 2^31 prime (odd when n >= 2: twiddles
 *     are Montgomery words), root of order exactly n;
 *   - Goldilocks, P = 2^64 - 2^32 + 1, root 7^((P-1)/n).
 * Buffers are contiguous [batch][n] rows, transformed in place from the
 * caller's point of view (src/ntt/ntt_cpu.cpp:87-90,
 * src/cuda/ntt/ntt_kernel.cuh:159,195 of the reference).
 *
 * Error model: every entry point returns a stntt_status; nothing throws and
 * nothing exits (the reference exits on device errors,
 * src/cuda/cu_util.cuh:15-28, and throws std::invalid_argument /
 * std::out_of_range on bad sizes, src/ntt/ntt_cpu.cpp:48-50,
 * src/cuda/cu_util.cuh:104-117).  stntt_last_error() gives the message of the
 * last failure on the calling thread.
 *
 * Threading: a context is externally synchronised (one thread at a time);
 * distinct contexts are independent.  Device entry points are asynchronous on
 * the given stream (0 = the context's own stream).  Transforms above the
 * single-pass sizes run through the context's scratch buffer: one issued on
 * a different stream than the previous such transform first waits for it
 * (an event), so they are correct on any mix of streams but do not overlap;
 * use one context per stream for concurrent large transforms.
 */
#ifndef STNTT_H
#define STNTT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum stntt_status {
    STNTT_OK = 0,
    STNTT_EINVAL = 1,  /* bad argument; the reference's std::invalid_argument */
    STNTT_ERANGE = 2,  /* host buffer too small; the reference's std::out_of_range */
    STNTT_ENOMEM = 3,  /* device allocation failed */
    STNTT_EDEVICE = 4, /* HIP runtime error; the reference's CCErr exit */
    STNTT_EAGAIN = 5,  /* serving: no slot / over the HBM budget right now */
    STNTT_ENOTFOUND = 6 /* serving: unknown request id */
} stntt_status;

/* STNTT_FIELD_U32: the reference's 32-bit prime field (runtime root and mod);
   STNTT_FIELD_GOLDILOCKS: P = 2^64 - 2^32 + 1;
   STNTT_FIELD_U64: any odd prime below 2^64 with a runtime root (Montgomery
   arithmetic); direct transforms, orders and polynomial multiply only. */
typedef enum stntt_field { STNTT_FIELD_U32 = 0, STNTT_FIELD_GOLDILOCKS = 1, STNTT_FIELD_U64 = 2 } stntt_field;
typedef enum stntt_direction { STNTT_FORWARD = 0, STNTT_INVERSE = 1 } stntt_direction;

/* transform flags */
#define STNTT_SIGNED_INPUT 1u /* 32-bit field: inputs are int32 (negative values allowed, as in the reference) */
/* Element order.  Default: natural in, natural out (nttStockham).
   STNTT_BITREV_OUT: natural in, bit-reversed out -- output word i holds
   X[bitrev(i)], the order of nttCt (src/ntt/ntt_cpu.cpp:11-42).
   STNTT_BITREV_IN: bit-reversed in, natural out -- the order inttGs takes
   (src/ntt/ntt_cpu.cpp:130-164).  Either direction accepts either flag (or
   both). */
#define STNTT_BITREV_IN 2u
#define STNTT_BITREV_OUT 4u

typedef struct stntt_ctx stntt_ctx;
typedef void *stntt_stream; /* a hipStream_t */

/* ---- library / context ------------------------------------------------ */
const char *stntt_version(void);
const char *stntt_last_error(void);
/* Context on one device: owns a stream, the resident twiddle plans and the
   scratch pool.  Replaces the reference's process-global twiddle symbols
   (src/cuda/ntt/implementations/common.cuh:8-9). */
int stntt_ctx_create(int device, stntt_ctx **out);
/* Waits for the context's stream, frees its plans and scratch.  Batches made
   on the context must be destroyed first; device buffers from
   stntt_dev_alloc are the caller's and are not freed here. */
int stntt_ctx_destroy(stntt_ctx *ctx);
stntt_stream stntt_ctx_stream(stntt_ctx *ctx);

/* ---- parameters (host only) ------------------------------------------- */
/* Smallest prime p = k n + 1 >= min_mod and root = g^((p-1)/n), g the
   smallest generator.  Replaces findParams (src/ntt/ntt_util.cpp:10-14);
   the same answers, over the whole 64-bit range (deterministic Miller-Rabin
   and Pollard rho in place of the reference's trial division, whose pow
   overflows above 2^32, ntt_util.h:72-96). */
int stntt_find_params(uint64_t n, uint64_t min_mod, uint64_t *root, uint64_t *mod);

/* ---- twiddle plans ------------------------------------------------------ */
/* Builds (or finds) the resident twiddle plan and writes the logical table
   T[i] = root^i mod p, i < n/2, natural order, to host memory -- the layout
   of the reference's initTwiddles (src/cuda/ntt/implementations/common.cuh:
   37-41).  The device keeps T factored as T[i] = HI[i >> b] * LO[i & (2^b-1)]
   and the values written here are expanded from those factors on the GPU. */
int stntt_twiddles_u32(stntt_ctx *ctx, uint64_t n, uint32_t root, uint32_t mod, int direction,
                       uint32_t *host_out);
int stntt_twiddles_goldilocks(stntt_ctx *ctx, uint32_t log2n, int direction, uint64_t *host_out);

/* ---- transforms on device buffers (asynchronous) ----------------------- */
/* In place on d_data[batches][2^log2n].  The GPU forms of nttStockham
   (src/ntt/ntt_cpu.cpp:87-90) and of the kernel chain launched by
   autoNtt<n> (src/cuda/ntt/ntt_kernel.cuh:169-184), for both directions. */
int stntt_ntt_u32(stntt_ctx *ctx, uint32_t *d_data, uint32_t log2n, uint64_t batches, uint32_t root,
                  uint32_t mod, int direction, uint32_t flags, stntt_stream stream);
int stntt_ntt_goldilocks(stntt_ctx *ctx, uint64_t *d_data, uint32_t log2n, uint64_t batches, int direction,
                         stntt_stream stream);
/* stntt_ntt_goldilocks with transform flags (STNTT_BITREV_IN / _OUT) */
int stntt_ntt_goldilocks_ex(stntt_ctx *ctx, uint64_t *d_data, uint32_t log2n, uint64_t batches, int direction,
                            uint32_t flags, stntt_stream stream);

/* stntt_ntt_goldilocks_ex that allocates nothing, for callers that cannot
   call hipMalloc: code under HIP graph capture, tight serving loops,
   harnesses that forbid allocation.  No counterpart in the reference.  The
   twiddle tables and the multi-pass ping-pong live in one fixed device
   array of the library (256 KiB of tables + 64 MiB of ping-pong per device,
   held until the process ends); each call regenerates its tables there, so
   a call -- or a replay of a captured call -- stands alone.  Limit: log2n
   <= 28 and a schedule whose scratch fits 64 MiB (one transform up to
   2^23, or 32 of 2^18, in natural order or with bit-reversed input; with
   bit-reversed output over three passes half that; larger is
   STNTT_EINVAL).  Every call on a device
   shares that device's array and keeps the last call's contents: calls
   that overlap on one device (several streams or threads at once) race. */
int stntt_ntt_goldilocks_noalloc(stntt_ctx *ctx, uint64_t *d_data, uint32_t log2n, uint64_t batches, int direction,
                                 uint32_t flags, stntt_stream stream);

/* Any odd prime p < 2^64 (STNTT_FIELD_U64): in place on d_data[batches][2^log2n]
   with a root of order exactly n (root^(n/2) = -1 mod p), e.g. from
   stntt_find_params(n, min_mod) with a 64-bit min_mod.  Every 64-bit input
   word is reduced mod p exactly; outputs are in [0, p).  flags: the order
   flags.  Twiddle plans for this field (and the others) are generated on the
   device.  No counterpart in the reference (its parameter search cannot
   produce 64-bit primes). */
int stntt_ntt_u64(stntt_ctx *ctx, uint64_t *d_data, uint32_t log2n, uint64_t batches, uint64_t root,
                  uint64_t mod, int direction, uint32_t flags, stntt_stream stream);
/* logical table T[i] = root^i mod p, i < n/2, as stntt_twiddles_u32 */
int stntt_twiddles_u64(stntt_ctx *ctx, uint32_t log2n, uint64_t root, uint64_t mod, int direction,
                       uint64_t *host_out);

/* ---- pointwise products and polynomial multiplication ------------------ */
/* a[i] = a[i] * b[i] mod p for i < count, a and b canonical (< p):
   elemWiseMul (src/ntt/ntt_util.h:28-33) on device buffers. */
int stntt_pointwise_u32(stntt_ctx *ctx, uint32_t *d_a, const uint32_t *d_b, uint64_t count, uint32_t mod,
                        stntt_stream stream);
int stntt_pointwise_goldilocks(stntt_ctx *ctx, uint64_t *d_a, const uint64_t *d_b, uint64_t count,
                               stntt_stream stream);
/* Cyclic (positive-wrapped) convolution of batches of length-2^log2n
   vectors, as pwcNtt (src/ntt/ntt_cpu.cpp:167-174): afterwards a holds
   a (*) b and b holds nttCt(b), its forward transform in bit-reversed order.
   Three transform launch sequences: forward(a) and forward(b) with
   bit-reversed output, then the inverse with bit-reversed input whose first
   pass multiplies by b as it loads a (the pointwise product is fused).  Up
   to 2^22 words per operand, forward(b) runs on an auxiliary stream of the
   context that starts after what `stream` holds and that `stream` waits for
   before the inverse (events), so the two forward transforms overlap.
   flags: STNTT_SIGNED_INPUT (32-bit field) for int32 inputs. */
int stntt_polymul_u32(stntt_ctx *ctx, uint32_t *d_a, uint32_t *d_b, uint32_t log2n, uint64_t batches,
                      uint32_t root, uint32_t mod, uint32_t flags, stntt_stream stream);
int stntt_polymul_goldilocks(stntt_ctx *ctx, uint64_t *d_a, uint64_t *d_b, uint32_t log2n, uint64_t batches,
                             stntt_stream stream);
int stntt_pointwise_u64(stntt_ctx *ctx, uint64_t *d_a, const uint64_t *d_b, uint64_t count, uint64_t mod,
                        stntt_stream stream);
int stntt_polymul_u64(stntt_ctx *ctx, uint64_t *d_a, uint64_t *d_b, uint32_t log2n, uint64_t batches,
                      uint64_t root, uint64_t mod, stntt_stream stream);

/* ---- transforms on host buffers (synchronous, autoNtt-shaped) ---------- */
/* Copies host_vec[0 .. n*batches) to the device, transforms it, copies it
   back over host_vec and returns the kernel-only time in *kernel_ms (HIP
   events around the launches, as the reference's cudaEvent pair,
   src/cuda/ntt/ntt_kernel.cuh:166-191).  host_len < n*batches is
   STNTT_ERANGE (src/cuda/cu_util.cuh:115-117).  n must be a power of two
   (STNTT_EINVAL otherwise).  The i32 form takes the reference's int storage:
   any int input, reduced with the reference's modulo() semantics. */
int stntt_auto_ntt_i32(stntt_ctx *ctx, int32_t *host_vec, size_t host_len, uint32_t n, uint32_t batches,
                       uint32_t root, uint32_t mod, int direction, float *kernel_ms);
int stntt_auto_ntt_goldilocks(stntt_ctx *ctx, uint64_t *host_vec, size_t host_len, uint32_t n,
                              uint32_t batches, int direction, float *kernel_ms);
/* the same with transform flags: STNTT_BITREV_OUT gives nttCt
   (src/ntt/ntt_cpu.cpp:11-42), STNTT_INVERSE with STNTT_BITREV_IN gives
   inttGs (src/ntt/ntt_cpu.cpp:130-164) */
int stntt_auto_ntt_i32_ex(stntt_ctx *ctx, int32_t *host_vec, size_t host_len, uint32_t n, uint32_t batches,
                          uint32_t root, uint32_t mod, int direction, uint32_t flags, float *kernel_ms);
/* pwcNtt (src/ntt/ntt_cpu.cpp:167-174) on host int vectors of length n:
   host_a becomes a (*) b, host_b becomes nttCt(b); *kernel_ms as above */
int stntt_auto_polymul_i32(stntt_ctx *ctx, int32_t *host_a, int32_t *host_b, uint32_t n, uint32_t root,
                           uint32_t mod, float *kernel_ms);

/* ---- the reference's stage loop, truncated or index-only ----------------
   nttStockham with nsteps and nttStockhamIdx (src/ntt/ntt_cpu.cpp:46-127):
   the reference's radix-2 Stockham loop run stage by stage on the device (a
   launch per stage over all rows), for the first nsteps stages (nsteps < 0
   or 0: none, the data untouched; at most 30).  Stage i uses the
   reference's stride J = pow(2, i, mod) (:68, :105) -- 2^i below the
   transform size -- and its zero-initialised ping-pong buffer, so calls
   with nsteps past log2 n give the reference's words too (the reference
   divides by J, so a zero stride is STNTT_EINVAL).  Arithmetic is exact in
   Z_mod for any root and any 2 <= mod < 2^31 (the reference's int products
   wrap from mod = 46341 on); with nsteps = log2 n the words equal the full
   transform's.  This is a diagnostic path: one HBM pass per stage, where
   stntt_ntt_u32 runs 2-3.  The Idx form moves 4-byte words only (mod
   selects the strides): with nsteps = log2 n it is the bit-reversal
   permutation.  flags: STNTT_SIGNED_INPUT (int32 data, reduced with the
   reference's modulo()). */
int stntt_stockham_steps_u32(stntt_ctx *ctx, uint32_t *d_data, uint32_t log2n, uint64_t batches, uint32_t root,
                             uint32_t mod, int nsteps, uint32_t flags, stntt_stream stream);
int stntt_stockham_idx_u32(stntt_ctx *ctx, uint32_t *d_data, uint32_t log2n, uint64_t batches, uint32_t mod,
                           int nsteps, stntt_stream stream);
/* the same on host int vectors of n * batches words, autoNtt-shaped (copy
   in, run, copy back; kernel-only milliseconds in *kernel_ms) */
int stntt_auto_stockham_steps_i32(stntt_ctx *ctx, int32_t *host_vec, size_t host_len, uint32_t n, uint32_t batches,
                                  uint32_t root, uint32_t mod, int nsteps, float *kernel_ms);
int stntt_auto_stockham_idx_i32(stntt_ctx *ctx, int32_t *host_vec, size_t host_len, uint32_t n, uint32_t batches,
                                uint32_t mod, int nsteps, float *kernel_ms);

/* ---- batched mixed transforms -------------------------------------------
   Thousands of independent transforms of mixed size (log2n <= 24), field and
   direction.  Transforms that fit a workgroup's LDS (32 KiB: 2^12
   Goldilocks, 2^13 32-bit words) run in one persistent launch: resident
   workgroups pull descriptors from a device work queue (largest first) and
   each runs its transform to completion.  Larger ones are grouped by
   (field, size, direction, parameters) and each group runs the multi-pass
   tile kernels over tables of its descriptors' buffers, full-chip launches
   with a scratch owned by the batch.  No counterpart in the reference, whose
   autoNtt<n> runs one transform (or one same-size batch) per call
   (src/cuda/ntt/ntt_kernel.cuh:152-201). */
typedef struct stntt_desc {
    uint32_t field;     /* stntt_field */
    uint32_t direction; /* stntt_direction */
    uint32_t log2n;
    uint32_t flags;     /* STNTT_SIGNED_INPUT for int32 data */
    uint32_t root, mod; /* 32-bit field only */
    const void *src;    /* device buffer of 2^log2n elements */
    void *dst;          /* may equal src (in place) */
} stntt_desc;

typedef struct stntt_batch stntt_batch;
/* Validates every descriptor, makes the twiddle plans resident, orders the
   queue and uploads it.  The descriptors' buffers must stay valid. */
int stntt_batch_create(stntt_ctx *ctx, const stntt_desc *descs, size_t count, stntt_batch **out);
/* the persistent launch (preceded by a one-thread queue reset kernel) and
   the groups' pass launches, on up to four streams at once: `stream` and
   three auxiliary streams of the context (kept for its life, shared with
   its other batches and its polynomial multiplies), which start after what
   `stream` holds and which `stream` waits for before anything enqueued on
   it later (events; under capture they become parallel branches of the
   graph).  Runs of one batch
   share its queue and scratch: a run (or graph replay) on another stream
   than the previous one waits for it with an event.  destroy waits for the
   last run. */
int stntt_batch_run(stntt_batch *b, stntt_stream stream);
/* capture the launch into a HIP graph once; later runs replay it */
int stntt_batch_capture(stntt_batch *b, stntt_stream stream);
int stntt_batch_run_graph(stntt_batch *b, stntt_stream stream);
/* the same transforms as one plain per-transform launch sequence each (the
   reference's calling pattern), for comparison */
int stntt_batch_run_plain(stntt_batch *b, stntt_stream stream);
int stntt_batch_destroy(stntt_batch *b);
/* Tuning: radix split of the batched kernel's sub-steps for (field, log2n),
   4 bits per sub-step (first in the low bits), as chosen by the autotuner.
   Every context starts with the splits tuned on MI355X (compiled into the
   library from stntt_amd/tuning/gfx950.json); this replaces one of them.
   nsub = 0 removes the entry: the default split formula then applies. */
int stntt_set_batch_split(stntt_ctx *ctx, int field, uint32_t log2n, uint32_t nsub, uint32_t radix_logs);
/* Tuning: stages per tile pass of the multi-pass path for (field, log2n),
   one log2 radix per pass (first pass first), as chosen by the autotuner.
   Every context starts with the splits tuned on MI355X (compiled into the
   library from stntt_amd/tuning/gfx950.json); this replaces one of them,
   and npasses = 0 removes it (the fewest-pass default then applies).  Each
   pass covers 4..12 stages (10..12: the large-radix kernel).  Applies to
   single transforms (batches == 1, what the autotuner times) of sizes that
   run in several passes; batches keep the fewest-pass default, and sizes
   up to 2^10 (8-byte words) or 2^12 (32-bit words) always run in one pass. */
int stntt_set_pass_split(stntt_ctx *ctx, int field, uint32_t log2n, uint32_t npasses, const uint32_t *logr);
/* stntt_set_pass_split with the tile shape too: passes of radix 2^10..2^12
   run the large-radix kernel on tiles of R x 2^logc[i] words (the full
   128 KiB tile down to a quarter of it; logc[i] = -1 picks the widest tile
   that gives every CU two tiles); passes up to 2^9 take -1 (their tiles have
   128-byte rows).  logc may be null (all -1).  Used by the autotuner. */
int stntt_set_pass_shape(stntt_ctx *ctx, int field, uint32_t log2n, uint32_t npasses, const uint32_t *logr,
                         const int32_t *logc);

/* ---- serving: continuous batching under an HBM budget -------------------
   No counterpart in the reference, whose drivers run one synchronous
   transform at a time (src/bench.cu:26-55).  Requests are queued from any
   thread; a server thread admits them in arrival order while the device
   bytes they hold stay within the budget, launches each admission round on a
   free worker stream (small transforms of the round in one persistent batch
   launch, larger ones on the multi-pass path) and retires finished rounds,
   so requests that arrive while others run are admitted as memory frees up
   and their work overlaps on the device.  Twiddles stay resident per worker.
   Host buffers must be pinned (stntt_server_host_alloc). */
typedef struct stntt_request {
    uint32_t field, direction, log2n, flags;
    uint32_t root, mod;    /* 32-bit field only */
    const void *host_in;   /* pinned, 2^log2n elements */
    void *host_out;        /* pinned, may equal host_in */
} stntt_request;

typedef struct stntt_server_stats {
    uint64_t completed;
    double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> result in host_out */
    double span_s;                          /* first submit -> last completion */
    uint64_t peak_bytes;                    /* peak request bytes held: admitted + cached, <= budget */
    uint64_t rounds;                        /* admission rounds launched */
    uint64_t batched;                       /* requests run through the batch kernel */
    uint64_t held_bytes;                    /* request bytes held now (admitted + cached) */
    uint64_t grouped;                       /* requests run in a grouped multi-pass run of one shape */
} stntt_server_stats;

typedef struct stntt_server stntt_server;
/* workers: concurrent streams (1..16).  The budget covers a fixed
   per-worker reservation for scratch plus every request buffer the server
   holds: admitted ones and idle cached ones (released, other size classes
   first, when a new allocation needs the room).  A request the device
   cannot hold even with nothing else in flight fails with STNTT_ENOMEM. */
int stntt_server_create(int device, uint64_t hbm_budget_bytes, uint32_t workers, stntt_server **out);
int stntt_server_destroy(stntt_server *srv);
int stntt_server_host_alloc(stntt_server *srv, size_t bytes, void **host_ptr);
/* srv may be null (the memory outlives the server that allocated it) */
int stntt_server_host_free(stntt_server *srv, void *host_ptr);
/* queue a request; never blocks.  A request that cannot fit the budget even
   alone is STNTT_EINVAL. */
int stntt_server_submit(stntt_server *srv, const stntt_request *req, uint64_t *id);
/* block until the request is done; returns its latency and forgets it.
   Unknown ids are STNTT_ENOTFOUND. */
int stntt_server_wait(stntt_server *srv, uint64_t id, double *latency_ms);
int stntt_server_poll(stntt_server *srv, uint64_t id, int *done);
int stntt_server_stats_get(stntt_server *srv, stntt_server_stats *out);
int stntt_server_reset_stats(stntt_server *srv);

/* ---- device memory and stream plumbing --------------------------------- */
int stntt_dev_alloc(stntt_ctx *ctx, size_t bytes, void **d_ptr);
int stntt_dev_free(stntt_ctx *ctx, void *d_ptr);
int stntt_memcpy_htod(stntt_ctx *ctx, void *d_dst, const void *h_src, size_t bytes, stntt_stream stream);
int stntt_memcpy_dtoh(stntt_ctx *ctx, void *h_dst, const void *d_src, size_t bytes, stntt_stream stream);
/* device-to-device copy by a plain streaming kernel (asynchronous); bytes % 8 == 0 */
int stntt_memcpy_dtod(stntt_ctx *ctx, void *d_dst, const void *d_src, size_t bytes, stntt_stream stream);
int stntt_stream_sync(stntt_ctx *ctx, stntt_stream stream);
/* Event pair on a stream: start/stop record on `stream`, elapsed waits for
   stop and returns milliseconds. */
int stntt_timer_start(stntt_ctx *ctx, stntt_stream stream);
int stntt_timer_stop(stntt_ctx *ctx, stntt_stream stream, float *ms);

/* ---- per-kernel timing ---------------------------------------------------
   With profiling on, every kernel of a transform is bracketed by HIP events
   on the stream it runs on (the reference brackets its launches with a
   cudaEvent pair, src/cuda/ntt/ntt_kernel.cuh:166-191).  profile_read waits
   for the last profiled transform and writes one duration per kernel. */
int stntt_set_profiling(stntt_ctx *ctx, int on);
int stntt_profile_read(stntt_ctx *ctx, float *kernel_ms, int capacity, int *count);

#ifdef __cplusplus
}
#endif
#endif
