// stntt.hip -- C ABI, contexts, resident twiddle plans and pass scheduling
// for the MI355X Stockham NTT.  Declarations and reference citations are in
// include/stntt.h.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/stntt.h"
#include "launch.hpp"
#include "ntt_batch.hpp"
#include "ntt_pass.hpp"
#include "plan.hpp"
#include "server.hpp"
#include "tuned_gfx950.hpp"

namespace stntt {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(STNTT_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                                 \
    } while (0)

// ---------------------------------------------------------------------------
// host modular arithmetic (exact, 128-bit products)
// ---------------------------------------------------------------------------
using u128 = unsigned __int128;

static uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    b %= m;
    while (e) {
        if (e & 1) r = (uint64_t)((u128)r * b % m);
        b = (uint64_t)((u128)b * b % m);
        e >>= 1;
    }
    return r;
}

// inverse by extended Euclid; 0 when a is not invertible
static uint64_t invmod(uint64_t a, uint64_t m) {
    __int128 t = 0, nt = 1, r = m, nr = a % m;
    while (nr) {
        __int128 q = r / nr, tmp;
        tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    if (r != 1) return 0;
    if (t < 0) t += m;
    return (uint64_t)t;
}


// ---------------------------------------------------------------------------
// twiddle plans: resident factor tables
//   hi[e] = w^((n/Rmax) e), e < Rmax;   lo[e] = w^e, e < n/Rmax
// so that w^e = hi[e >> b] * lo[e & (2^b - 1)] with 2^b = n / Rmax.
// ---------------------------------------------------------------------------
struct PlanKey {
    int field, dir;
    uint32_t log_n;
    uint64_t root, mod;
    bool operator<(const PlanKey &o) const {
        return std::tie(field, dir, log_n, root, mod) < std::tie(o.field, o.dir, o.log_n, o.root, o.mod);
    }
};

struct Plan {
    PlanKey key{};
    void *d_tables = nullptr;  // hi then lo
    size_t bytes = 0;
    uint32_t log_rmax = 0;
    FieldGL::Rt rt_gl{};
    FieldU32::Rt rt_u32{};
    FieldM64::Rt rt_m64{};
    const void *hi() const { return d_tables; }
    const void *lo() const { return (const char *)d_tables + (entry_bytes(key.field) << log_rmax); }
    // a 32-bit Montgomery word, a Goldilocks word, a 64-bit Montgomery word
    static size_t entry_bytes(int field) { return field == STNTT_FIELD_U32 ? 4 : 8; }
};

// Batched transforms above the LDS-resident size, grouped by everything that
// selects a plan and flags: each group runs the multi-pass tile kernels over
// tables of its descriptors' buffers (a full-chip launch per pass), instead
// of one workgroup streaming each transform through global scratch.
struct BatchGroup {
    int field = 0, dir = 0;
    uint32_t log2n = 0, flags = 0;
    uint64_t root = 0, mod = 0;
    const Plan *plan = nullptr;
    uint32_t count = 0;
    void **d_src = nullptr;  // device table of source pointers
    void **d_dst = nullptr;  // device table of destination pointers
};

}  // namespace stntt

struct stntt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::map<stntt::PlanKey, std::unique_ptr<stntt::Plan>> plans;
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // Multi-pass transforms on different streams share `scratch`: each one
    // records scratch_ev on its stream when its passes are enqueued, and one
    // on another stream waits for it first.
    hipEvent_t scratch_ev = nullptr;
    hipStream_t scratch_last = nullptr;  // stream of that record (compared, never used)
    void *staging = nullptr;  // device buffer of the host-buffer entry points
    size_t staging_bytes = 0;
    // Auxiliary streams, created on first use and kept for the context's
    // life (HIP maps a process's streams onto a few hardware queues, and
    // streams created and destroyed per call were measured to land on busy
    // ones): a polynomial multiply's side transform runs on aux[0], a
    // batch's lanes 1..3 on aux[0..2].  Work on them is forked from and
    // joined back into the caller's stream by events.
    static constexpr int MAX_AUX = 3;
    hipStream_t aux[MAX_AUX] = {};
    hipEvent_t side_fork = nullptr, side_join = nullptr;
    void *side_scratch = nullptr;
    size_t side_scratch_bytes = 0;
    // per-pass timing: when on, events bracket every kernel of a transform
    bool profile = false;
    static constexpr int MAX_PROFILE_EVENTS = 9;
    hipEvent_t prof_ev[MAX_PROFILE_EVENTS] = {};
    int prof_count = 0;  // events recorded by the last profiled transform
    // autotuned sub-step splits of the batched kernel: (field, log2n) -> (nsub, radix_logs)
    std::map<std::pair<int, uint32_t>, std::pair<uint32_t, uint32_t>> batch_split;
    // autotuned tile-pass splits: (field, log2n) -> (stages, tile columns log2 or -1) per pass
    std::map<std::pair<int, uint32_t>, std::vector<std::pair<int, int>>> pass_split;
    int num_cus = 0;
};

struct stntt_batch {
    stntt_ctx *ctx = nullptr;
    std::vector<stntt_desc> descs;        // as given (for the plain path)
    stntt::BatchTask *d_tasks = nullptr;  // persistent-kernel queue, largest first
    uint32_t count = 0;                   // tasks in the queue
    std::vector<stntt::BatchGroup> groups;
    void *d_group_scratch = nullptr;  // ping-pong of the largest group
    uint32_t *d_counter = nullptr;
    uint64_t *d_scratch = nullptr;
    uint64_t scratch_stride = 0;  // 64-bit words per workgroup
    uint32_t grid = 0;
    // The groups and the queue launch are independent of each other: they
    // run on up to MAX_LANES streams at once, lane 0 being the run's own
    // stream and lanes 1.. the context's auxiliary streams, forked from it
    // and joined back into it by events (captured as parallel branches of
    // the graph).  Work is dealt to lanes largest first, each to the least
    // loaded lane; each lane has its own ping-pong scratch, at
    // lane_off[lane] in d_group_scratch.
    static constexpr int MAX_LANES = 4;
    static_assert(MAX_LANES - 1 <= stntt_ctx::MAX_AUX, "a lane per auxiliary stream");
    int lanes = 1;
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[MAX_LANES] = {};
    int queue_lane = 0;                  // lane of the queue launch
    std::vector<int> group_lane;         // lane of each group
    std::vector<size_t> launch_order;    // groups, largest first
    size_t lane_off[MAX_LANES] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // runs share the queue counter and the group scratch: a run on another
    // stream than the previous one waits for it (run_ev, recorded after each)
    hipEvent_t run_ev = nullptr;
    hipStream_t run_last = nullptr;  // stream of that record (compared, never used)
};

namespace stntt {

static int set_device(stntt_ctx *ctx) {
    HIP_TRY(hipSetDevice(ctx->device));
    return STNTT_OK;
}

// the context's first `count` auxiliary streams, created on first use
static int ctx_aux_streams(stntt_ctx *ctx, int count) {
    for (int i = 0; i < count && i < stntt_ctx::MAX_AUX; ++i)
        if (!ctx->aux[i]) HIP_TRY(hipStreamCreateWithFlags(&ctx->aux[i], hipStreamNonBlocking));
    return STNTT_OK;
}

// true when s is recording a graph: the cross-stream orderings below (event
// record / wait between eager launches) are skipped for captured work
static bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

static int ensure_buffer(void **buf, size_t *cap, size_t bytes) {
    if (*cap >= bytes) return STNTT_OK;
    if (*buf) {
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipFree(*buf));
        *buf = nullptr;
        *cap = 0;
    }
    if (hipMalloc(buf, bytes) != hipSuccess) {
        *buf = nullptr;
        return fail(STNTT_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
    *cap = bytes;
    return STNTT_OK;
}

// Twiddle factor tables generated on the device (the reference builds its
// table on the host, src/cuda/ntt/implementations/common.cuh:37-41): entry e
// of hi is wstep^e, entry e of lo is w^e, each by square-and-multiply in the
// field's twiddle arithmetic (exact, so the words equal a host power loop).
template <class FP>
__global__ void gen_tables_kernel(typename FP::Tw *tab, uint32_t rmax, uint32_t nlo, typename FP::Tw w,
                                  typename FP::Tw wstep, typename FP::Rt rt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rmax + nlo) return;
    uint32_t e = i < rmax ? i : i - rmax;
    typename FP::Tw b = i < rmax ? wstep : w;
    typename FP::Tw r = FP::tw_one(rt);
    while (e) {
        if (e & 1) r = FP::twiddle(r, b, rt);
        b = FP::twiddle(b, b, rt);
        e >>= 1;
    }
    tab[i] = r;
}

template <class FP>
static int gen_tables(stntt_ctx *ctx, Plan *plan, typename FP::Tw w, typename FP::Tw wstep, uint64_t rmax,
                      uint64_t nlo, const typename FP::Rt &rt) {
    const uint64_t count = rmax + nlo;  // <= 2^14 + 2^(L-14)
    hipLaunchKernelGGL(gen_tables_kernel<FP>, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0, ctx->stream,
                       (typename FP::Tw *)plan->d_tables, (uint32_t)rmax, (uint32_t)nlo, w, wstep, rt);
    HIP_TRY(hipGetLastError());
    // transforms may run on other streams: the tables are complete on return
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return STNTT_OK;
}

static int get_plan(stntt_ctx *ctx, int field, uint32_t log_n, uint64_t root, uint64_t mod, int dir,
                    const Plan **out) {
    PlanKey key{field, dir, log_n, root, mod};
    auto it = ctx->plans.find(key);
    if (it != ctx->plans.end()) {
        *out = it->second.get();
        return STNTT_OK;
    }
    auto plan = std::make_unique<Plan>();
    plan->key = key;
    uint32_t log_rmax = log_n < MAX_LOGR ? log_n : MAX_LOGR;
    plan->log_rmax = log_rmax;
    const uint64_t n = 1ull << log_n, rmax = 1ull << log_rmax, nlo = n >> log_rmax;
    const uint64_t p = field == STNTT_FIELD_GOLDILOCKS ? GL_P : mod;
    // the root for this direction; Goldilocks uses 7^((P-1)/n), 7 being the
    // smallest generator (what findParams' generator search returns)
    uint64_t w = field == STNTT_FIELD_GOLDILOCKS ? (log_n ? powmod(7, (GL_P - 1) >> log_n, GL_P) : 1) : root % p;
    if (dir == STNTT_INVERSE) w = invmod(w, p);
    const uint64_t wstep = powmod(w, nlo, p);  // w^(n / Rmax): the hi table's ratio
    plan->bytes = (size_t)(rmax + nlo) * Plan::entry_bytes(field);
    HIP_TRY(hipMalloc(&plan->d_tables, plan->bytes));
    int rc;
    if (field == STNTT_FIELD_GOLDILOCKS) {
        plan->rt_gl.ninv = invmod(n % GL_P, GL_P);
        rc = gen_tables<FieldGL>(ctx, plan.get(), w, wstep, rmax, nlo, plan->rt_gl);
    } else if (field == STNTT_FIELD_U32) {
        const uint32_t p32 = (uint32_t)p;
        plan->rt_u32.p = p32;
        plan->rt_u32.pinv = u32_neg_inverse(p32);
        plan->rt_u32.bar = (uint32_t)((1ull << 32) / p32);
        plan->rt_u32.c32 = (uint32_t)((1ull << 32) % p32);
        plan->rt_u32.ninv = u32_to_mont(invmod(n % p32, p32), p32);
        plan->rt_u32.mu = ~0ull / p32;
        rc = gen_tables<FieldU32>(ctx, plan.get(), u32_to_mont(w, p32), u32_to_mont(wstep, p32), rmax, nlo,
                                  plan->rt_u32);
    } else {
        M64Rt &rt = plan->rt_m64;
        rt.p = p;
        rt.pinv = m64_neg_inverse(p);
        rt.r1 = (uint64_t)(((u128)1 << 64) % p);
        rt.r2 = (uint64_t)((u128)rt.r1 * rt.r1 % p);
        rt.ninv = invmod(n % p, p);
        auto mont = [&](uint64_t x) { return (uint64_t)(((u128)x << 64) % p); };
        rc = gen_tables<FieldM64>(ctx, plan.get(), mont(w), mont(wstep), rmax, nlo, rt);
    }
    if (rc) {
        (void)hipFree(plan->d_tables);
        return rc;
    }
    *out = plan.get();
    ctx->plans.emplace(key, std::move(plan));
    return STNTT_OK;
}

// ---------------------------------------------------------------------------
// allocation-free Goldilocks transforms (stntt_ntt_goldilocks_noalloc): the
// twiddle tables and the multi-pass ping-pong live in one fixed device
// array, so a call allocates nothing.  Every call on a device shares that
// device's copy and keeps the last call's contents; calls that overlap on
// one device race.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_TABLE_BYTES = 256 << 10;   // hi (2^14 words) + lo (up to 2^14 words)
constexpr size_t NOALLOC_DATA_BYTES = 64ull << 20;  // ping-pong of the pass schedule
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_TABLE_BYTES + NOALLOC_DATA_BYTES];
// The device address of `scratch` on each device (each holds its own copy);
// in host code the array's name is not a device address.
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// once per device, with that device current (context creation)
static int noalloc_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return STNTT_OK;  // the entry point reports it
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return STNTT_OK;
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
    return STNTT_OK;
}

static void *noalloc_ptr(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return g_scratch_ptr[device];
}

template <class FP>
__global__ void copy_kernel(const typename FP::T *in, typename FP::T *out, uint64_t count) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = in[i];
}

// The context's scratch buffer, grown to at least `bytes`, for work on
// stream s.  Work on different streams shares it: the previous user's
// completion (scratch_ev) is waited for on s first.  release_scratch marks
// the end of this use.  Graph captures skip the cross-stream ordering.
static int acquire_scratch(stntt_ctx *ctx, size_t bytes, hipStream_t s, void **out) {
    int rc = ensure_buffer(&ctx->scratch, &ctx->scratch_bytes, bytes);
    if (rc) return rc;
    if (ctx->scratch_last && ctx->scratch_last != s && !stream_capturing(s))
        HIP_TRY(hipStreamWaitEvent(s, ctx->scratch_ev, 0));
    *out = ctx->scratch;
    return STNTT_OK;
}

static int release_scratch(stntt_ctx *ctx, hipStream_t s) {
    if (stream_capturing(s)) return STNTT_OK;
    HIP_TRY(hipEventRecord(ctx->scratch_ev, s));
    ctx->scratch_last = s;
    return STNTT_OK;
}

// ---------------------------------------------------------------------------
// the reference's radix-2 stage loop, stage by stage (nsteps, nttStockhamIdx)
// ---------------------------------------------------------------------------
// One stage of src/ntt/ntt_cpu.cpp:66-79 (ARITH) or of its index-only twin
// :103-118 (!ARITH) with stride J, from `in` to `out`, one thread per pair
// (row b, s, j) with s < n / (2J), j < J:
//     out[2sJ + j] = l + r,  out[(2s + 1)J + j] = w^(Js) (l - r)
//     l = in[sJ + j], r = in[sJ + j + n/2]
// (or the plain moves).  J is the reference's pow(2, i, mod), which equals
// 2^i below the transform size but may be anything past it, so the strides
// are general; positions of `out` the stage does not write keep their
// contents, as in the reference's ping-pong.  Arithmetic is exact in the
// 32-bit field: inputs are reduced with the reference's modulo() semantics
// as they are read, twiddles w^e by square-and-multiply (a diagnostic path,
// so no table is needed and any root and any modulus 2 <= p < 2^31 work).
template <bool ARITH>
__global__ void stockham_stage_kernel(const uint32_t *in, uint32_t *out, uint32_t log_n, uint32_t J, uint32_t pairs,
                                      uint64_t total, uint32_t root, uint32_t p, uint32_t bar, uint32_t c32,
                                      uint64_t mu, int is_signed) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint64_t b = i / pairs;
    const uint32_t q = (uint32_t)(i - b * pairs);
    const uint32_t s = q / J, j = q - s * J;
    const uint64_t base = b << log_n;
    const uint32_t half = (1u << log_n) >> 1;
    uint32_t l = in[base + (uint64_t)s * J + j];
    uint32_t r = in[base + (uint64_t)s * J + j + half];
    if constexpr (ARITH) {
        l = u32_reduce_input(l, p, bar, c32, is_signed != 0);
        r = u32_reduce_input(r, p, bar, c32, is_signed != 0);
        uint32_t w = 1 % p, bw = root;  // root < p
        for (uint32_t e = J * s; e; e >>= 1) {  // J s < n / 2
            if (e & 1) w = u32_mul_any(w, bw, p, mu);
            bw = u32_mul_any(bw, bw, p, mu);
        }
        out[base + 2ull * s * J + j] = u32_add(l, r, p);
        out[base + (2ull * s + 1) * J + j] = u32_mul_any(u32_sub(l, r, p), w, p, mu);
    } else {
        out[base + 2ull * s * J + j] = l;
        out[base + (2ull * s + 1) * J + j] = r;
    }
}

__global__ void fill_u32_kernel(uint32_t *out, uint64_t count, uint32_t value) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = value;
}

// nsteps stages of the reference's loop over d_data[batches][2^log_n], in
// place from the caller's view: a zeroed ping-pong buffer, a launch per
// stage, then a copy back when the result ended in the other buffer (the
// reference's odd-swap copy, src/ntt/ntt_cpu.cpp:81-84).
// stages a nsteps call may ask for: pow(2, i, mod) is exact in the
// reference's int arithmetic up to i = 30
constexpr int MAX_STAGE_STEPS = 30;

template <bool ARITH>
static int stockham_steps(stntt_ctx *ctx, uint32_t *d_data, uint32_t log_n, uint64_t batches, uint32_t root,
                          uint32_t mod, int nsteps, bool is_signed, hipStream_t s) {
    const uint64_t n = 1ull << log_n, total_words = batches << log_n;
    if (batches == 0 || nsteps <= 0) return STNTT_OK;  // no stage: the input as it is
    if (total_words >= (1ull << 40)) return fail(STNTT_EINVAL, "buffer too large");
    for (int i = 0; i < nsteps; ++i)  // the reference divides by J (src/ntt/ntt_cpu.cpp:69)
        if (powmod(2, (uint64_t)i, mod) == 0)
            return fail(STNTT_EINVAL, "stage %d has stride pow(2, %d, %u) = 0", i, i, mod);
    void *scratch = nullptr;
    int rc = acquire_scratch(ctx, total_words * 4, s, &scratch);
    if (rc) return rc;
    auto blocks = [](uint64_t count) { return dim3((uint32_t)((count + 255) / 256)); };
    hipLaunchKernelGGL(fill_u32_kernel, blocks(total_words), dim3(256), 0, s, (uint32_t *)scratch, total_words, 0u);
    HIP_TRY(hipGetLastError());
    uint32_t *cur = d_data, *other = (uint32_t *)scratch;
    const uint32_t p = mod;
    const uint32_t bar = (uint32_t)((1ull << 32) / p), c32 = (uint32_t)((1ull << 32) % p);
    const uint64_t mu = ~0ull / p;
    for (int i = 0; i < nsteps; ++i) {
        const uint64_t J = powmod(2, (uint64_t)i, mod);  // pow(2, i, mod), src/ntt/ntt_cpu.cpp:68
        const uint64_t pairs = J * (n / (2 * J));
        if (pairs) {
            const uint64_t count = batches * pairs;
            hipLaunchKernelGGL(stockham_stage_kernel<ARITH>, blocks(count), dim3(256), 0, s, cur, other, log_n,
                               (uint32_t)J, (uint32_t)pairs, count, root % p, p, bar, c32, mu, is_signed ? 1 : 0);
            HIP_TRY(hipGetLastError());
        }
        std::swap(cur, other);
    }
    if (cur != d_data) {
        hipLaunchKernelGGL(copy_kernel<FieldU32>, blocks(total_words), dim3(256), 0, s, cur, d_data, total_words);
        HIP_TRY(hipGetLastError());
    }
    return release_scratch(ctx, s);
}

// The pass schedule of `batches` transforms of 2^L words: the fewest-pass
// default, or the context's tuned split for a single transform (tuned
// splits were measured on one transform, i.e. latency; a batch is
// throughput-bound and keeps the default).
// Bit-reversed orders stage a tile through LDS on the first or last pass,
// which only the 128-byte-row tile kernel does: schedules for them use no
// large-radix pass.
template <class FP>
static Schedule transform_schedule(const stntt_ctx *ctx, int field, uint32_t L, uint64_t batches,
                                   uint32_t flags = 0) {
    const bool allow_big = !(flags & (STNTT_BITREV_IN | STNTT_BITREV_OUT));
    Schedule sch = make_schedule<FP>((int)L, batches, allow_big);
    if (!sch.single && batches == 1) {
        auto it = ctx->pass_split.find({field, L});
        if (it != ctx->pass_split.end()) {
            bool big = false;
            for (const auto &pr : it->second) big = big || pr.first >= MIN_LOGR_BIG;
            if (allow_big || !big) {
                sch.npasses = (int)it->second.size();
                for (int i = 0; i < sch.npasses; ++i) {
                    sch.logr[i] = it->second[i].first;
                    sch.logc[i] = it->second[i].second;
                }
            }
        }
    }
    return sch;
}

// Whether a multi-pass schedule runs on two buffers (the caller's and one
// scratch; an odd pass count ends with the last pass in place): all but
// bit-reversed output over an odd pass count, whose last pass writes other
// positions than it reads.
static bool two_buffer_schedule(const Schedule &sch, uint32_t flags) {
    return !((flags & STNTT_BITREV_OUT) && (sch.npasses & 1));
}

// scratch bytes a schedule needs besides the data (no copy-back: the last
// pass writes the caller's buffer): one buffer of the data's size, two for
// a three-buffer schedule
template <class FP>
static uint64_t schedule_scratch_bytes(const Schedule &sch, uint64_t total_words, uint32_t flags) {
    if (sch.npasses < 2) return 0;
    return (sch.npasses > 2 && !two_buffer_schedule(sch, flags) ? 2 : 1) * total_words * sizeof(typename FP::T);
}

// Runs the pass schedule of `plan` over d_data[batches][n] in place.
// flags: STNTT_SIGNED_INPUT, STNTT_BITREV_IN, STNTT_BITREV_OUT.  mul_in, when
// set, is multiplied into the input word by word as the first pass loads it
// (with bit-reversed input only on multi-pass schedules).  Gathered form:
// in_tab / out_tab (device arrays of `batches` pointers) replace d_data as
// the source of the first pass and the destination of the last, and
// ext_scratch (2 n batches words) replaces the context's scratch, so that
// nothing is allocated (graph capture).
template <class FP>
static int run_transform(stntt_ctx *ctx, const Plan &plan, typename FP::T *d_data, uint64_t batches,
                         uint32_t flags, hipStream_t s, const typename FP::T *mul_in = nullptr,
                         const typename FP::T *const *in_tab = nullptr, typename FP::T *const *out_tab = nullptr,
                         typename FP::T *ext_scratch = nullptr) {
    using T = typename FP::T;
    const uint32_t L = plan.key.log_n;
    if (batches == 0) return STNTT_OK;
    // kernel arguments and grids are 32-bit: one launch covers at most 2^31
    // transforms and 2^31 workgroups
    if (batches >= (1ull << 31) || (batches << L) >> L != batches || (batches << L) >= (1ull << 40))
        return fail(STNTT_EINVAL, "%llu transforms of 2^%u words is more than one call handles",
                    (unsigned long long)batches, L);
    if (L == 0) return STNTT_OK;  // zero stages: identity (src/ntt/ntt_cpu.cpp:64-84)
    const Schedule sch = transform_schedule<FP>(ctx, plan.key.field, L, batches, flags);
    // tile columns (log2) of each pass
    int logc[4] = {0, 0, 0, 0};
    if (!sch.single) {
        uint32_t lj = 0;
        for (int p = 0; p < sch.npasses; ++p) {
            if (sch.logr[p] >= MIN_LOGR_BIG) {
                logc[p] = big_logc_choice<FP>((int)L, sch.logr[p], (int)lj, p == 0, batches, ctx->num_cus, sch.logc[p]);
                if (logc[p] < 0)
                    return fail(STNTT_EINVAL, "no tile shape for a radix-2^%d pass of 2^%u words", sch.logr[p], L);
            } else {
                logc[p] = tile_logc<FP>();
            }
            const uint64_t tiles = batches * ((((1ull << L) >> sch.logr[p])) >> logc[p]);
            if (tiles >= (1ull << 31))
                return fail(STNTT_EINVAL, "%llu tiles exceed one launch", (unsigned long long)tiles);
            lj += sch.logr[p];
        }
    }
    const bool inv = plan.key.dir == STNTT_INVERSE;
    const bool rev_in = flags & STNTT_BITREV_IN, rev_out = flags & STNTT_BITREV_OUT;
    if (mul_in && !sch.single && !rev_in)
        return fail(STNTT_EINVAL, "a fused pointwise operand needs bit-reversed input on multi-pass sizes");
    // Where the passes write.  The passes alternate between the caller's
    // buffers (d_data, or the buffers of out_tab) and one scratch buffer,
    // and an odd count ends with the last pass in place: it reads
    // x[j + m J] and writes y[k J + j] (J = n / R), the same positions, each
    // tile its own columns, every load of a tile before any of its stores.
    // Two buffers of n w bytes per transform instead of three, so a 2^24
    // Goldilocks transform's 256 MiB stays in the 256 MB MALL between
    // passes.  Bit-reversed output over an odd pass count (that last pass
    // writes other positions than it reads) keeps two scratch halves that
    // the passes alternate between, the last pass landing in the caller's
    // buffers.  No copy-back either way.
    const bool two_buf = two_buffer_schedule(sch, flags);
    T *scratch[2] = {nullptr, nullptr};
    const uint64_t total = batches << L;
    if (sch.npasses > 1) {
        if (ext_scratch) {
            scratch[0] = ext_scratch;
        } else {
            void *p = nullptr;
            int rc = acquire_scratch(ctx, schedule_scratch_bytes<FP>(sch, total, flags), s, &p);
            if (rc) return rc;
            scratch[0] = (T *)p;
        }
        scratch[1] = two_buf ? nullptr : scratch[0] + total;
    }
    PassArgs<FP> a{};
    a.hi = (const typename FP::Tw *)plan.hi();
    a.lo = (const typename FP::Tw *)plan.lo();
    a.log_n = L;
    a.log_rmax = plan.log_rmax;
    a.batches = (uint32_t)batches;
    if constexpr (std::is_same_v<FP, FieldGL>) a.rt = plan.rt_gl;
    else if constexpr (std::is_same_v<FP, FieldU32>) a.rt = plan.rt_u32;
    else a.rt = plan.rt_m64;

    bool src_caller = true;  // the pass reads the caller's buffers (else src)
    T *src = nullptr;
    uint32_t log_j = 0;
    // per-pass events for a single caller's transform on the context's
    // scratch only (grouped runs of a batch and a polynomial multiply's side
    // transform go out on several streams at once)
    const bool prof = ctx->profile && !in_tab && !ext_scratch && sch.npasses + 2 <= stntt_ctx::MAX_PROFILE_EVENTS;
    if (prof) {
        ctx->prof_count = 0;
        HIP_TRY(hipEventRecord(ctx->prof_ev[ctx->prof_count++], s));
    }
    for (int p = 0; p < sch.npasses; ++p) {
        const int logr = sch.logr[p];
        const bool first = p == 0, last = p == sch.npasses - 1;
        const bool dst_caller = two_buf ? (last || (p & 1)) : last;
        T *dst = dst_caller ? nullptr : scratch[two_buf ? 0 : (p & 1)];
        // the caller's input on the first pass (in_tab or d_data), its
        // output buffers on later ones (out_tab or d_data)
        a.in = src_caller ? d_data : src;
        a.in_tab = src_caller ? (first ? in_tab : out_tab) : nullptr;
        a.out = dst_caller ? d_data : dst;
        a.out_tab = dst_caller ? out_tab : nullptr;
        a.log_j = log_j;
        // the inverse's n^-1: through the first pass's twiddles where the
        // field allows it (every J = 1 output takes a twiddle product),
        // otherwise on the last pass's store
        const bool scale_in = inv && !sch.single && FP::SCALE_IN;
        a.flags = (first ? (PASS_FIRST | (flags & STNTT_SIGNED_INPUT ? PASS_SIGNED : 0) |
                            (scale_in ? PASS_SCALE_IN : 0))
                         : 0) |
                  (last ? (PASS_LAST | (inv && !scale_in ? PASS_SCALE : 0)) : 0);
        a.mul_in = first ? mul_in : nullptr;
        bool perm = false;
        if (sch.single) {
            a.flags |= (rev_in ? PASS_BITREV_IN : 0) | (rev_out ? PASS_BITREV_OUT : 0);
        } else {
            perm = (first && rev_in) || (last && rev_out);
        }
        uint32_t grid;
        bool narrow = false;
        if (sch.single) {
            uint64_t c = 1ull << pass_logc<FP>(logr, true);
            grid = (uint32_t)((batches + c - 1) / c);
            // too few wide tiles to fill the chip: one transform per
            // workgroup instead (a wide tile computes all C columns)
            if (logr <= narrow_max_logr<FP>() && grid < 2 * (uint64_t)ctx->num_cus && c > 1) {
                narrow = true;
                grid = (uint32_t)batches;
            }
        } else {
            a.tiles_per_transform = (uint32_t)(((1ull << L) >> logr) >> logc[p]);
            grid = (uint32_t)(batches * a.tiles_per_transform);  // checked before the first launch
        }
        if (!sch.single && logr >= MIN_LOGR_BIG)
            launch_big<FP>(plan.key.dir, first, logr, logc[p], a, grid, ctx->num_cus, s);
        else
            launch_pass<FP>(plan.key.dir, sch.single, first, perm, narrow, logr, a, grid, s);
        HIP_TRY(hipGetLastError());
        if (prof) HIP_TRY(hipEventRecord(ctx->prof_ev[ctx->prof_count++], s));
        src_caller = dst_caller;
        src = dst;
        log_j += logr;
    }
    if (sch.npasses > 1 && !ext_scratch) return release_scratch(ctx, s);
    return STNTT_OK;
}

template <class FP>
__global__ void expand_twiddles_kernel(typename FP::T *out, const typename FP::Tw *hi, const typename FP::Tw *lo,
                                       uint32_t log_b, uint64_t count, typename FP::Rt rt) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    typename FP::T x;
    if constexpr (std::is_same_v<FP, FieldGL>) {
        x = gl_canon(gl_mul(hi[i >> log_b], lo[i & ((1ull << log_b) - 1)]));
    } else if constexpr (std::is_same_v<FP, FieldU32>) {
        x = u32_mont(u32_mont(hi[i >> log_b], lo[i & ((1ull << log_b) - 1)], rt.p, rt.pinv), 1, rt.p, rt.pinv);
    } else {  // Montgomery form product, then back to normal form
        x = m64_mul(m64_mul(hi[i >> log_b], lo[i & ((1ull << log_b) - 1)], rt.p, rt.pinv), 1, rt.p, rt.pinv);
    }
    out[i] = x;
}

// ---------------------------------------------------------------------------
// argument checks shared by the entry points
// ---------------------------------------------------------------------------
static int check_u32_params(uint32_t log2n, uint64_t root, uint64_t mod, int dir) {
    if (log2n > 30) return fail(STNTT_EINVAL, "log2n %u too large", log2n);
    if (mod < 2 || mod >= (1ull << 31))
        return fail(STNTT_EINVAL, "mod %llu outside [2, 2^31)", (unsigned long long)mod);
    if (dir != STNTT_FORWARD && dir != STNTT_INVERSE) return fail(STNTT_EINVAL, "bad direction %d", dir);
    if (log2n >= 1) {
        // twiddles are Montgomery words (an odd modulus); an even one would
        // admit the forward transform only, n = 2^k having no inverse
        if (!(mod & 1)) return fail(STNTT_EINVAL, "mod %llu must be odd", (unsigned long long)mod);
        uint64_t half = powmod(root, 1ull << (log2n - 1), mod);
        if (half != mod - 1)
            return fail(STNTT_EINVAL, "root %llu is not of order n = 2^%u mod %llu (root^(n/2) != -1)",
                        (unsigned long long)root, log2n, (unsigned long long)mod);
        if (dir == STNTT_INVERSE && invmod((1ull << log2n) % mod, mod) == 0)
            return fail(STNTT_EINVAL, "n is not invertible mod %llu", (unsigned long long)mod);
    }
    return STNTT_OK;
}

static int check_u64_params(uint32_t log2n, uint64_t root, uint64_t mod, int dir) {
    if (log2n > 32) return fail(STNTT_EINVAL, "log2n %u too large (at most 32)", log2n);
    if (mod < 3 || !(mod & 1)) return fail(STNTT_EINVAL, "mod %llu must be odd and at least 3", (unsigned long long)mod);
    if (dir != STNTT_FORWARD && dir != STNTT_INVERSE) return fail(STNTT_EINVAL, "bad direction %d", dir);
    if (log2n >= 1) {
        if (powmod(root % mod, 1ull << (log2n - 1), mod) != mod - 1)
            return fail(STNTT_EINVAL, "root %llu is not of order n = 2^%u mod %llu (root^(n/2) != -1)",
                        (unsigned long long)root, log2n, (unsigned long long)mod);
        if (dir == STNTT_INVERSE && invmod((1ull << log2n) % mod, mod) == 0)
            return fail(STNTT_EINVAL, "n is not invertible mod %llu", (unsigned long long)mod);
    }
    return STNTT_OK;
}

static int log2_exact(uint64_t n, uint32_t *out) {
    if (n == 0 || (n & (n - 1))) return fail(STNTT_EINVAL, "size (%llu) must be a power of 2", (unsigned long long)n);
    uint32_t l = 0;
    while ((1ull << l) < n) ++l;
    *out = l;
    return STNTT_OK;
}

template <class T, class Run>
static int auto_ntt_common(stntt_ctx *ctx, T *host_vec, size_t host_len, uint32_t n, uint32_t batches,
                           float *kernel_ms, Run run) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    uint32_t L;
    int rc = log2_exact(n, &L);
    if (rc) return rc;
    if (batches == 0) return fail(STNTT_EINVAL, "batches must be positive");
    size_t count = (size_t)n * batches;
    if (host_len < count || !host_vec)
        return fail(STNTT_ERANGE, "host buffer (%zu elements) is smaller than device buffer (%zu)", host_len, count);
    if ((rc = set_device(ctx))) return rc;
    if ((rc = ensure_buffer(&ctx->staging, &ctx->staging_bytes, count * sizeof(T)))) return rc;
    HIP_TRY(hipMemcpyAsync(ctx->staging, host_vec, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev_start, ctx->stream));
    if ((rc = run(L, (T *)ctx->staging))) return rc;
    HIP_TRY(hipEventRecord(ctx->ev_stop, ctx->stream));
    HIP_TRY(hipMemcpyAsync(host_vec, ctx->staging, count * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (kernel_ms) *kernel_ms = ms;
    return STNTT_OK;
}

// ---------------------------------------------------------------------------
// batched mixed transforms
// ---------------------------------------------------------------------------
// Default sub-step split for the batched kernel: radix 16 where possible,
// the remainder spread evenly; the count is made even for transforms that
// stream through global memory so the ping-pong ends in dst.
static void default_batch_split(uint32_t L, bool global_path, uint32_t *nsub, uint32_t *radix_logs) {
    uint32_t ns = L == 0 ? 0 : (L + 3) / 4;
    if (global_path && (ns & 1)) ns += 1;
    uint32_t logs = 0;
    for (uint32_t i = 0; i < ns; ++i) logs |= (L / ns + (i < L % ns ? 1u : 0u)) << (4 * i);
    *nsub = ns;
    *radix_logs = logs;
}

static bool batch_global_path(int field, uint32_t L) {
    uint64_t w = field == STNTT_FIELD_GOLDILOCKS ? 8 : 4;
    return (w << L) > (uint64_t)BATCH_LDS_BYTES;
}

static int batch_validate_split(uint32_t L, bool global_path, uint32_t nsub, uint32_t logs) {
    if (nsub == 0 || nsub > 8) return fail(STNTT_EINVAL, "bad sub-step count %u", nsub);
    uint32_t sum = 0;
    for (uint32_t i = 0; i < nsub; ++i) {
        uint32_t lr = (logs >> (4 * i)) & 15;
        if (lr < 1 || lr > 4) return fail(STNTT_EINVAL, "sub-step radix 2^%u outside [2, 16]", lr);
        sum += lr;
    }
    if (sum != L) return fail(STNTT_EINVAL, "sub-step radices cover 2^%u, not 2^%u", sum, L);
    if (global_path && (nsub & 1)) return fail(STNTT_EINVAL, "streamed transforms need an even sub-step count");
    return STNTT_OK;
}

static uint64_t elem_bytes(uint32_t field) { return field == STNTT_FIELD_U32 ? 4 : 8; }

// largest transform a batch descriptor may hold
constexpr uint32_t BATCH_MAX_LOG = 24;

static int batch_build(stntt_ctx *ctx, const stntt_desc *descs, size_t count, stntt_batch *b) {
    std::vector<BatchTask> tasks;
    tasks.reserve(count);
    std::map<std::tuple<int, int, uint32_t, uint32_t, uint64_t, uint64_t>, size_t> gidx;
    std::vector<std::vector<void *>> gsrc, gdst;
    for (size_t i = 0; i < count; ++i) {
        const stntt_desc &d = descs[i];
        if (d.field != STNTT_FIELD_U32 && d.field != STNTT_FIELD_GOLDILOCKS)
            return fail(STNTT_EINVAL, "descriptor %zu: bad field %u", i, d.field);
        if (d.direction > 1) return fail(STNTT_EINVAL, "descriptor %zu: bad direction %u", i, d.direction);
        if (d.log2n > BATCH_MAX_LOG)
            return fail(STNTT_EINVAL, "descriptor %zu: log2n %u above the batched limit %u", i, d.log2n, BATCH_MAX_LOG);
        if (d.flags & ~STNTT_SIGNED_INPUT)
            return fail(STNTT_EINVAL, "descriptor %zu: flags 0x%x (batches take natural order only)", i, d.flags);
        if (!d.src || !d.dst) return fail(STNTT_EINVAL, "descriptor %zu: null buffer", i);
        const Plan *plan;
        int rc;
        const uint64_t root = d.field == STNTT_FIELD_U32 ? d.root % d.mod : 0;
        const uint64_t mod = d.field == STNTT_FIELD_U32 ? d.mod : GL_P;
        if (d.field == STNTT_FIELD_U32 && (rc = check_u32_params(d.log2n, d.root, d.mod, d.direction))) return rc;
        if ((rc = get_plan(ctx, (int)d.field, d.log2n, root, mod, (int)d.direction, &plan))) return rc;
        const uint32_t sflag = (d.field == STNTT_FIELD_U32 && (d.flags & STNTT_SIGNED_INPUT)) ? STNTT_SIGNED_INPUT : 0;
        if (batch_global_path(d.field, d.log2n)) {  // a group of the multi-pass path
            auto key = std::make_tuple((int)d.field, (int)d.direction, d.log2n, sflag, root, mod);
            auto it = gidx.find(key);
            if (it == gidx.end()) {
                BatchGroup g;
                g.field = (int)d.field;
                g.dir = (int)d.direction;
                g.log2n = d.log2n;
                g.flags = sflag;
                g.root = root;
                g.mod = mod;
                g.plan = plan;
                it = gidx.emplace(key, b->groups.size()).first;
                b->groups.push_back(g);
                gsrc.emplace_back();
                gdst.emplace_back();
            }
            gsrc[it->second].push_back(const_cast<void *>(d.src));
            gdst[it->second].push_back(d.dst);
            continue;
        }
        BatchTask t{};
        if (d.field == STNTT_FIELD_U32) t.rt32 = plan->rt_u32;
        else t.rtgl = plan->rt_gl;
        t.src = d.src;
        t.dst = d.dst;
        t.hi = plan->hi();
        t.lo = plan->lo();
        t.field = d.field;
        t.dir = d.direction;
        t.log_n = d.log2n;
        t.log_rmax = plan->log_rmax;
        t.flags = PASS_FIRST | PASS_LAST | (d.direction == STNTT_INVERSE ? PASS_SCALE : 0) | (sflag ? PASS_SIGNED : 0);
        auto it = ctx->batch_split.find({(int)d.field, d.log2n});
        if (it != ctx->batch_split.end()) {
            t.nsub = it->second.first;
            t.radix_logs = it->second.second;
        } else {
            default_batch_split(d.log2n, false, &t.nsub, &t.radix_logs);
        }
        if (d.log2n > 0 && (rc = batch_validate_split(d.log2n, false, t.nsub, t.radix_logs))) return rc;
        tasks.push_back(t);
    }
    // largest first (longest-processing-time order) so the tail is short
    std::stable_sort(tasks.begin(), tasks.end(), [](const BatchTask &x, const BatchTask &y) {
        uint64_t cx = ((uint64_t)(x.field ? 2 : 1) << x.log_n) * (x.log_n + 1);
        uint64_t cy = ((uint64_t)(y.field ? 2 : 1) << y.log_n) * (y.log_n + 1);
        return cx > cy;
    });
    int per_cu = 0;
    HIP_TRY(batch_occupancy(&per_cu));
    if (per_cu < 1) per_cu = 1;
    b->count = (uint32_t)tasks.size();
    b->grid = (uint32_t)std::min<uint64_t>((uint64_t)ctx->num_cus * per_cu, std::max<size_t>(tasks.size(), 1));
    b->scratch_stride = 0;  // no queue task streams through global memory
    HIP_TRY(hipMalloc(&b->d_tasks, std::max<size_t>(tasks.size(), 1) * sizeof(BatchTask)));
    if (!tasks.empty())
        HIP_TRY(hipMemcpy(b->d_tasks, tasks.data(), tasks.size() * sizeof(BatchTask), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&b->d_counter, sizeof(uint32_t)));
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        BatchGroup &g = b->groups[gi];
        g.count = (uint32_t)gsrc[gi].size();
        const size_t tab = g.count * sizeof(void *);
        HIP_TRY(hipMalloc((void **)&g.d_src, tab));
        HIP_TRY(hipMalloc((void **)&g.d_dst, tab));
        HIP_TRY(hipMemcpy(g.d_src, gsrc[gi].data(), tab, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(g.d_dst, gdst[gi].data(), tab, hipMemcpyHostToDevice));
    }
    // Lanes.  Work estimate of a transform: n log2 n butterflies' worth,
    // Goldilocks words 2.5 times a 32-bit word's.
    auto work = [](uint32_t field, uint32_t L) -> uint64_t {
        return ((uint64_t)(field == STNTT_FIELD_GOLDILOCKS ? 5 : 2) << L) * std::max<uint32_t>(L, 1);
    };
    uint64_t queue_work = 0;
    for (const BatchTask &t : tasks) queue_work += work(t.field, t.log_n);
    std::vector<uint64_t> gwork(b->groups.size());
    for (size_t gi = 0; gi < b->groups.size(); ++gi)
        gwork[gi] = b->groups[gi].count * work(b->groups[gi].field, b->groups[gi].log2n);
    b->launch_order.resize(b->groups.size());
    for (size_t gi = 0; gi < b->groups.size(); ++gi) b->launch_order[gi] = gi;
    std::stable_sort(b->launch_order.begin(), b->launch_order.end(),
                     [&](size_t x, size_t y) { return gwork[x] > gwork[y]; });
    const size_t units = b->groups.size() + (b->count ? 1 : 0);
    b->lanes = (int)std::max<size_t>(1, std::min<size_t>(stntt_batch::MAX_LANES, units));
    uint64_t load[stntt_batch::MAX_LANES] = {};
    auto least = [&]() {
        int l = 0;
        for (int i = 1; i < b->lanes; ++i)
            if (load[i] < load[l]) l = i;
        return l;
    };
    size_t lane_need[stntt_batch::MAX_LANES] = {};
    b->group_lane.assign(b->groups.size(), 0);
    bool queue_placed = !b->count;
    for (size_t k = 0; k < b->launch_order.size() || !queue_placed; ++k) {
        const size_t gi = k < b->launch_order.size() ? b->launch_order[k] : 0;
        if (!queue_placed && (k >= b->launch_order.size() || queue_work >= gwork[gi])) {
            b->queue_lane = least();
            load[b->queue_lane] += queue_work;
            queue_placed = true;
        }
        if (k >= b->launch_order.size()) break;
        const int l = least();
        b->group_lane[gi] = l;
        load[l] += gwork[gi];
        // a ping-pong of the group's words: the most any pass schedule of
        // the group needs (splits may be re-tuned after the batch is built)
        const BatchGroup &g = b->groups[gi];
        lane_need[l] = std::max(lane_need[l], (size_t)2 * g.count * (elem_bytes(g.field) << g.log2n));
    }
    size_t scratch = 0;
    for (int l = 0; l < b->lanes; ++l) {
        b->lane_off[l] = scratch;
        scratch += (lane_need[l] + 255) & ~(size_t)255;
    }
    if (scratch) HIP_TRY(hipMalloc(&b->d_group_scratch, scratch));
    if (b->lanes > 1) {
        if (int rc = ctx_aux_streams(ctx, b->lanes - 1)) return rc;
        HIP_TRY(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming));
        for (int l = 1; l < b->lanes; ++l) HIP_TRY(hipEventCreateWithFlags(&b->join_ev[l], hipEventDisableTiming));
    }
    b->descs.assign(descs, descs + count);
    return STNTT_OK;
}

static int batch_launch(stntt_batch *b, hipStream_t s) {
    hipStream_t ls[stntt_batch::MAX_LANES] = {s};
    for (int l = 1; l < b->lanes; ++l) ls[l] = b->ctx->aux[l - 1];
    if (b->lanes > 1) {  // fork: every lane starts after what s holds
        HIP_TRY(hipEventRecord(b->fork_ev, s));
        for (int l = 1; l < b->lanes; ++l) HIP_TRY(hipStreamWaitEvent(ls[l], b->fork_ev, 0));
    }
    if (b->count || b->groups.empty()) {
        // (an empty batch still enqueues the reset, so a capture of it holds
        // a node)
        hipStream_t q = ls[b->queue_lane];
        launch_batch_reset(b->d_counter, q);
        HIP_TRY(hipGetLastError());
        if (b->count) {
            launch_batch(b->grid, b->d_tasks, b->count, b->d_counter, b->d_scratch, b->scratch_stride, q);
            HIP_TRY(hipGetLastError());
        }
    }
    for (size_t gi : b->launch_order) {
        const BatchGroup &g = b->groups[gi];
        const int l = b->group_lane[gi];
        void *scratch = (char *)b->d_group_scratch + b->lane_off[l];
        int rc;
        if (g.field == STNTT_FIELD_GOLDILOCKS)
            rc = run_transform<FieldGL>(b->ctx, *g.plan, nullptr, g.count, g.flags, ls[l], nullptr,
                                        (const uint64_t *const *)g.d_src, (uint64_t *const *)g.d_dst,
                                        (uint64_t *)scratch);
        else
            rc = run_transform<FieldU32>(b->ctx, *g.plan, nullptr, g.count, g.flags, ls[l], nullptr,
                                         (const uint32_t *const *)g.d_src, (uint32_t *const *)g.d_dst,
                                         (uint32_t *)scratch);
        if (rc) return rc;
    }
    for (int l = 1; l < b->lanes; ++l) {  // join: s continues after every lane
        HIP_TRY(hipEventRecord(b->join_ev[l], ls[l]));
        HIP_TRY(hipStreamWaitEvent(s, b->join_ev[l], 0));
    }
    return STNTT_OK;
}

static void batch_free(stntt_batch *b) {
    if (b->run_ev) (void)hipEventDestroy(b->run_ev);
    if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
    for (int l = 1; l < stntt_batch::MAX_LANES; ++l)
        if (b->join_ev[l]) (void)hipEventDestroy(b->join_ev[l]);
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_tasks) (void)hipFree(b->d_tasks);
    if (b->d_counter) (void)hipFree(b->d_counter);
    if (b->d_scratch) (void)hipFree(b->d_scratch);
    for (BatchGroup &g : b->groups) {
        if (g.d_src) (void)hipFree(g.d_src);
        if (g.d_dst) (void)hipFree(g.d_dst);
    }
    if (b->d_group_scratch) (void)hipFree(b->d_group_scratch);
}

// ---------------------------------------------------------------------------
// serving
// ---------------------------------------------------------------------------

static int server_check_request(const stntt_request &r) {
    if (r.field != STNTT_FIELD_U32 && r.field != STNTT_FIELD_GOLDILOCKS) return fail(STNTT_EINVAL, "bad field");
    if (r.direction > 1) return fail(STNTT_EINVAL, "bad direction");
    if (!r.host_in || !r.host_out) return fail(STNTT_EINVAL, "null host buffer");
    if (r.flags & ~STNTT_SIGNED_INPUT) return fail(STNTT_EINVAL, "flags 0x%x (served in natural order only)", r.flags);
    if (r.field == STNTT_FIELD_GOLDILOCKS) {
        if (r.log2n > 28) return fail(STNTT_EINVAL, "log2n %u too large for serving", r.log2n);
        return STNTT_OK;
    }
    if (r.log2n > 28) return fail(STNTT_EINVAL, "log2n %u too large for serving", r.log2n);
    return check_u32_params(r.log2n, r.root, r.mod, r.direction);
}

// per-worker scratch reserved up front: multi-pass ping-pong for the largest
// served size plus the batch kernel's streaming slices
static uint64_t server_worker_reserve(uint32_t max_log) {
    return 2 * (8ull << max_log) + (uint64_t)SERVER_BATCH_GRID * (8ull << SERVER_BATCH_MAX_LOG);
}
constexpr uint32_t SERVER_MAX_LOG = 22;

// Frees cached (idle) device buffers, other size classes than `keep`
// first, until the bytes held by the server (admitted requests plus the
// cache) are at most `target`.  A cached buffer belongs to no round in
// flight, so freeing it races with nothing.
static void server_trim_cache(stntt_server *S, uint64_t target, size_t keep) {
    for (int pass = 0; pass < 2 && S->used + S->cached > target; ++pass) {
        for (auto it = S->cache.begin(); it != S->cache.end() && S->used + S->cached > target; ++it) {
            if (pass == 0 && it->first == keep) continue;
            auto &v = it->second;
            while (!v.empty() && S->used + S->cached > target) {
                (void)hipFree(v.back());
                v.pop_back();
                S->cached -= it->first;
            }
        }
    }
}

// A device buffer of `bytes` for a request being admitted (the caller has
// checked used + bytes <= budget).  A cached buffer of the size class is
// reused; otherwise cached buffers of other classes are released until the
// new one fits the budget beside everything held, and on an allocation
// failure the whole cache is released and the allocation tried once more.
// Null when the device cannot take it now.
static void *server_get_buffer(stntt_server *S, size_t bytes) {
    auto &v = S->cache[bytes];
    if (!v.empty()) {
        void *p = v.back();
        v.pop_back();
        S->cached -= bytes;
        return p;
    }
    server_trim_cache(S, S->budget - bytes, bytes);
    void *p = nullptr;
    if (hipMalloc(&p, bytes) == hipSuccess) return p;
    server_trim_cache(S, 0, 0);
    if (hipMalloc(&p, bytes) == hipSuccess) return p;
    return nullptr;
}

// a retired request's buffer goes back to the cache (still held)
static void server_cache_put(stntt_server *S, ServerRequest &r) {
    S->used -= r.cost;
    S->cache[r.d_bytes].push_back(r.d_buf);
    S->cached += r.d_bytes;
    r.d_buf = nullptr;
}

// Launch one admission round on worker w.  Runs on the server thread with
// the lock released; the requests it touches are owned by the round.
static hipEvent_t server_event(stntt_server *S) {
    if (!S->free_events.empty()) {
        hipEvent_t e = S->free_events.back();
        S->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Requests of one shape (2^L words each) that one grouped run can take out
// of `want`: as many as the worker's ping-pong scratch holds for the
// multi-pass schedule of that many transforms (at least one).
template <class FP>
static uint64_t server_chunk(const stntt_ctx *ctx, int field, uint32_t L, uint64_t want) {
    uint64_t c = std::min<uint64_t>(want, SERVER_MAX_ROUND);
    while (c > 1) {
        const uint64_t need = schedule_scratch_bytes<FP>(transform_schedule<FP>(ctx, field, L, c), c << L, 0);
        if (need <= ctx->scratch_bytes) break;
        c = std::max<uint64_t>(1, std::min<uint64_t>(c - 1, ctx->scratch_bytes / (need / c)));
    }
    return c;
}

// Called by the server thread only (it alone touches the event pool).
static int server_launch_round(stntt_server *S, int w, std::vector<ServerRequest *> &round, ServerRound &rd) {
    ServerWorker &W = S->workers[w];
    hipStream_t s = W.stream;
    BatchTask *const h_tasks = W.h_tasks + (size_t)rd.slot * SERVER_MAX_ROUND;
    void **const h_tab = W.h_tab + (size_t)rd.slot * SERVER_MAX_ROUND;
    auto ev = [&](std::vector<hipEvent_t> &into) -> hipEvent_t {
        hipEvent_t e = server_event(S);
        if (e) into.push_back(e);
        return e;
    };
    // 1. uploads in admission order on the upload stream; the worker waits
    for (ServerRequest *r : round) {
        const stntt_request &q = r->req;
        HIP_TRY(hipMemcpyAsync(r->d_buf, q.host_in, (size_t)elem_bytes(q.field) << q.log2n, hipMemcpyHostToDevice,
                               S->up));
    }
    hipEvent_t ev_in = ev(rd.aux);
    if (!ev_in) return fail(STNTT_EDEVICE, "hipEventCreate failed");
    HIP_TRY(hipEventRecord(ev_in, S->up));
    HIP_TRY(hipStreamWaitEvent(s, ev_in, 0));
    // 2. the small transforms in one persistent launch, then their downloads
    uint32_t nb = 0;
    ServerGroup small;
    for (ServerRequest *r : round) {
        const stntt_request &q = r->req;
        if (q.log2n > SERVER_BATCH_MAX_LOG) continue;
        BatchTask &t = h_tasks[nb++];
        t = BatchTask{};
        const Plan *plan;
        int rc;
        if (q.field == STNTT_FIELD_U32) {
            if ((rc = get_plan(W.ctx, STNTT_FIELD_U32, q.log2n, q.root % q.mod, q.mod, q.direction, &plan))) return rc;
            t.rt32 = plan->rt_u32;
        } else {
            if ((rc = get_plan(W.ctx, STNTT_FIELD_GOLDILOCKS, q.log2n, 0, GL_P, q.direction, &plan))) return rc;
            t.rtgl = plan->rt_gl;
        }
        t.src = r->d_buf;
        t.dst = r->d_buf;
        t.hi = plan->hi();
        t.lo = plan->lo();
        t.field = q.field;
        t.dir = q.direction;
        t.log_n = q.log2n;
        t.log_rmax = plan->log_rmax;
        t.flags = PASS_FIRST | PASS_LAST | (q.direction == STNTT_INVERSE ? PASS_SCALE : 0) |
                  ((q.field == STNTT_FIELD_U32 && (q.flags & STNTT_SIGNED_INPUT)) ? PASS_SIGNED : 0);
        default_batch_split(q.log2n, batch_global_path(q.field, q.log2n), &t.nsub, &t.radix_logs);
        small.ids.push_back(r->id);
    }
    auto download = [&](ServerGroup &g, const std::vector<ServerRequest *> &reqs) -> int {
        hipEvent_t ev_c = ev(rd.aux);
        if (!ev_c) return fail(STNTT_EDEVICE, "hipEventCreate failed");
        HIP_TRY(hipEventRecord(ev_c, s));
        HIP_TRY(hipStreamWaitEvent(S->down, ev_c, 0));
        for (ServerRequest *r : reqs)
            HIP_TRY(hipMemcpyAsync(r->req.host_out, r->d_buf, (size_t)elem_bytes(r->req.field) << r->req.log2n,
                                   hipMemcpyDeviceToHost, S->down));
        g.ev = server_event(S);
        if (!g.ev) return fail(STNTT_EDEVICE, "hipEventCreate failed");
        HIP_TRY(hipEventRecord(g.ev, S->down));
        rd.groups.push_back(std::move(g));
        rd.pending++;
        return STNTT_OK;
    };
    if (nb) {
        HIP_TRY(hipMemcpyAsync(W.d_tasks, h_tasks, nb * sizeof(BatchTask), hipMemcpyHostToDevice, s));
        launch_batch_reset(W.d_counter, s);
        launch_batch(std::min<uint32_t>(nb, SERVER_BATCH_GRID), W.d_tasks, nb, W.d_counter, W.d_scratch,
                     (uint64_t)1 << SERVER_BATCH_MAX_LOG, s);
        HIP_TRY(hipGetLastError());
        std::vector<ServerRequest *> reqs;
        for (ServerRequest *r : round)
            if (r->req.log2n <= SERVER_BATCH_MAX_LOG) reqs.push_back(r);
        int rc = download(small, reqs);
        if (rc) return rc;
    }
    // 3. the larger transforms, grouped by shape (field, size, direction,
    //    parameters, signed input) in order of first arrival.  A group runs
    //    the multi-pass path once over pointer tables (in place), in chunks
    //    that the worker's ping-pong scratch holds; a lone request runs the
    //    tuned single-transform schedule.  Each chunk's download follows its
    //    last pass.
    using ShapeKey = std::tuple<int, int, uint32_t, uint32_t, uint64_t, uint64_t>;
    std::map<ShapeKey, std::vector<ServerRequest *>> shapes;
    std::vector<ShapeKey> order;
    for (ServerRequest *r : round) {
        const stntt_request &q = r->req;
        if (q.log2n <= SERVER_BATCH_MAX_LOG) continue;
        const bool u32 = q.field == STNTT_FIELD_U32;
        ShapeKey k{(int)q.field, (int)q.direction, q.log2n, u32 ? (q.flags & STNTT_SIGNED_INPUT) : 0u,
                   u32 ? q.root % q.mod : 0, u32 ? q.mod : GL_P};
        auto &v = shapes[k];
        if (v.empty()) order.push_back(k);
        v.push_back(r);
    }
    uint32_t ntab = 0;  // pointer-table entries this round has used
    for (const ShapeKey &k : order) {
        const std::vector<ServerRequest *> &v = shapes[k];
        const stntt_request &q0 = v.front()->req;
        const bool gl = q0.field == STNTT_FIELD_GOLDILOCKS;
        for (size_t i = 0; i < v.size();) {
            const uint64_t left = v.size() - i;
            const uint64_t chunk = gl ? server_chunk<FieldGL>(W.ctx, STNTT_FIELD_GOLDILOCKS, q0.log2n, left)
                                      : server_chunk<FieldU32>(W.ctx, STNTT_FIELD_U32, q0.log2n, left);
            std::vector<ServerRequest *> part(v.begin() + i, v.begin() + i + chunk);
            i += chunk;
            int rc;
            if (chunk == 1) {
                ServerRequest *r = part[0];
                const stntt_request &q = r->req;
                rc = gl ? stntt_ntt_goldilocks(W.ctx, (uint64_t *)r->d_buf, q.log2n, 1, q.direction, s)
                        : stntt_ntt_u32(W.ctx, (uint32_t *)r->d_buf, q.log2n, 1, q.root, q.mod, q.direction, q.flags,
                                        s);
            } else {
                void **tab = W.d_tab + ntab;
                for (uint64_t j = 0; j < chunk; ++j) h_tab[ntab + j] = part[j]->d_buf;
                HIP_TRY(hipMemcpyAsync(tab, h_tab + ntab, chunk * sizeof(void *), hipMemcpyHostToDevice, s));
                ntab += (uint32_t)chunk;
                const Plan *plan;
                if ((rc = get_plan(W.ctx, std::get<0>(k), q0.log2n, std::get<4>(k), std::get<5>(k), q0.direction,
                                   &plan)))
                    return rc;
                if (gl)
                    rc = run_transform<FieldGL>(W.ctx, *plan, nullptr, chunk, std::get<3>(k), s, nullptr,
                                                (const uint64_t *const *)tab, (uint64_t *const *)tab,
                                                (uint64_t *)W.ctx->scratch);
                else
                    rc = run_transform<FieldU32>(W.ctx, *plan, nullptr, chunk, std::get<3>(k), s, nullptr,
                                                 (const uint32_t *const *)tab, (uint32_t *const *)tab,
                                                 (uint32_t *)W.ctx->scratch);
                S->grouped_requests += chunk;
            }
            if (rc) return rc;
            ServerGroup g;
            for (ServerRequest *r : part) g.ids.push_back(r->id);
            if ((rc = download(g, part))) return rc;
        }
    }
    return STNTT_OK;
}

static void server_loop(stntt_server *S) {
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        // 1. retire finished groups, and rounds whose groups all finished
        bool progressed = false;
        for (auto it = S->rounds.begin(); it != S->rounds.end();) {
            for (ServerGroup &g : it->groups) {
                if (g.done || hipEventQuery(g.ev) != hipSuccess) continue;
                const auto now = Clock::now();
                for (uint64_t id : g.ids) {
                    ServerRequest &r = S->reqs[id];
                    r.done = true;
                    r.latency_ms = std::chrono::duration<double, std::milli>(now - r.t_submit).count();
                    S->latencies.push_back(r.latency_ms);
                    server_cache_put(S, r);
                    S->completed++;
                }
                S->t_last = now;
                g.done = true;
                it->pending--;
                S->free_events.push_back(g.ev);
                progressed = true;
            }
            if (it->pending == 0) {
                S->workers[it->worker].inflight--;
                S->workers[it->worker].slot_busy[it->slot] = false;
                for (hipEvent_t e : it->aux) S->free_events.push_back(e);
                it = S->rounds.erase(it);
            } else {
                ++it;
            }
        }
        if (progressed) S->cv_done.notify_all();
        if (S->stop && S->pending.empty() && S->rounds.empty()) break;

        // 2. admit a round onto the least busy worker with a round slot free
        int w = -1;
        for (size_t i = 0; i < S->workers.size(); ++i)
            if (S->workers[i].inflight < SERVER_WORKER_ROUNDS &&
                (w < 0 || S->workers[i].inflight < S->workers[w].inflight))
                w = (int)i;
        std::vector<ServerRequest *> round;
        if (w >= 0) {
            while (!S->pending.empty() && round.size() < SERVER_MAX_ROUND) {
                ServerRequest &r = S->reqs[S->pending.front()];
                if (S->used + r.cost > S->budget) break;  // wait for memory to free up
                void *buf = server_get_buffer(S, r.d_bytes);
                if (!buf) {
                    if (!round.empty() || !S->rounds.empty()) break;  // memory frees up as rounds retire
                    // nothing in flight and the whole cache released: the
                    // device cannot hold this request now, fail it
                    r.done = true;
                    r.status = STNTT_ENOMEM;
                    S->pending.pop_front();
                    S->cv_done.notify_all();
                    continue;
                }
                r.d_buf = buf;
                S->used += r.cost;
                S->peak_used = std::max(S->peak_used, S->used + S->cached);
                round.push_back(&r);
                S->pending.pop_front();
            }
        }
        if (!round.empty()) {
            ServerWorker &W = S->workers[w];
            int slot = 0;  // inflight < SERVER_WORKER_ROUNDS, so one slot is free
            while (slot + 1 < SERVER_WORKER_ROUNDS && W.slot_busy[slot]) ++slot;
            W.slot_busy[slot] = true;
            W.inflight++;
            S->rounds_launched++;
            for (ServerRequest *r : round) S->batched_requests += r->req.log2n <= SERVER_BATCH_MAX_LOG ? 1 : 0;
            ServerRound rd{w, slot, {}, {}, 0};
            lk.unlock();
            int rc = server_launch_round(S, w, round, rd);
            lk.lock();
            if (rc) {  // fail the round's requests once nothing of it is in flight
                (void)hipStreamSynchronize(S->up);
                (void)hipStreamSynchronize(S->workers[w].stream);
                (void)hipStreamSynchronize(S->down);
                for (ServerRequest *r : round) {
                    r->done = true;
                    r->status = rc;
                    server_cache_put(S, *r);
                }
                S->workers[w].inflight--;
                S->workers[w].slot_busy[slot] = false;
                for (ServerGroup &g : rd.groups) S->free_events.push_back(g.ev);
                for (hipEvent_t e : rd.aux) S->free_events.push_back(e);
                S->cv_done.notify_all();
            } else {
                S->rounds.push_back(std::move(rd));
            }
            continue;
        }
        if (!progressed) S->cv_work.wait_for(lk, std::chrono::microseconds(S->rounds.empty() ? 2000 : 20));
    }
}

}  // namespace stntt

using namespace stntt;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// 64-bit primality and factoring for the parameter search.  The reference
// uses trial division (isPrime, uniquePrimeFactors: ntt_util.cpp:51-84),
// which is exact but takes ~2^32 steps for 64-bit candidates; these give the
// same answers: Miller-Rabin with the first 12 prime bases is deterministic
// below 3.3e24 (Sorenson and Webster, Math. Comp. 86, 2017), and factors
// come from trial division by small primes, then Pollard's rho with Brent's
// cycle finding, each factor confirmed prime.
// ---------------------------------------------------------------------------
static uint64_t mulmod_u64(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

static bool is_prime_u64(uint64_t n) {
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint64_t q : bases)
        if (n % q == 0) return n == q;
    uint64_t d = n - 1;
    int r = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++r;
    }
    for (uint64_t a : bases) {
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool composite = true;
        for (int i = 1; i < r && composite; ++i) {
            x = mulmod_u64(x, x, n);
            if (x == n - 1) composite = false;
        }
        if (composite) return false;
    }
    return true;
}

static uint64_t gcd_u64(uint64_t a, uint64_t b) {
    while (b) {
        uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// a nontrivial factor of an odd composite n
static uint64_t rho_factor(uint64_t n) {
    for (uint64_t c = 1;; ++c) {
        auto f = [&](uint64_t x) { return (uint64_t)(((u128)x * x + c) % n); };
        uint64_t y = 2, x = 2, d = 1, q = 1, ys = 2;
        const uint64_t m = 128;
        for (uint64_t len = 1; d == 1; len <<= 1) {
            x = y;
            for (uint64_t i = 0; i < len; ++i) y = f(y);
            for (uint64_t k = 0; k < len && d == 1; k += m) {
                ys = y;
                for (uint64_t i = 0; i < m && i < len - k; ++i) {
                    y = f(y);
                    q = mulmod_u64(q, x > y ? x - y : y - x, n);
                }
                d = gcd_u64(q, n);
            }
        }
        if (d == n) {  // the batched product overshot: step one at a time
            do {
                ys = f(ys);
                d = gcd_u64(x > ys ? x - ys : ys - x, n);
            } while (d == 1);
        }
        if (d != n) return d;
    }
}

static void prime_factors(uint64_t n, std::vector<uint64_t> &out) {
    if (n == 1) return;
    if (is_prime_u64(n)) {
        out.push_back(n);
        return;
    }
    uint64_t d = rho_factor(n);
    prime_factors(d, out);
    prime_factors(n / d, out);
}

// distinct prime factors of m, ascending
static void distinct_prime_factors(uint64_t m, std::vector<uint64_t> &f) {
    for (uint64_t d = 2; d < 1000 && d * d <= m; ++d)
        if (m % d == 0) {
            f.push_back(d);
            while (m % d == 0) m /= d;
        }
    if (m > 1) prime_factors(m, f);
    std::sort(f.begin(), f.end());
    f.erase(std::unique(f.begin(), f.end()), f.end());
}

// ---------------------------------------------------------------------------
// pointwise products and polynomial multiplication
// ---------------------------------------------------------------------------
template <class FP>
__global__ void pointwise_kernel(typename FP::T *a, const typename FP::T *b, uint64_t count, typename FP::Rt rt) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = FP::mulg(a[i], b[i], rt);
}

template <class FP>
static int launch_pointwise(typename FP::T *a, const typename FP::T *b, uint64_t count, typename FP::Rt rt,
                            hipStream_t s) {
    const uint64_t blocks = (count + 255) / 256;
    if (blocks > 0x7FFFFFFFull) return fail(STNTT_EINVAL, "count %llu too large", (unsigned long long)count);
    if (blocks) hipLaunchKernelGGL(pointwise_kernel<FP>, dim3((uint32_t)blocks), dim3(256), 0, s, a, b, count, rt);
    HIP_TRY(hipGetLastError());
    return STNTT_OK;
}

// Polynomial multiplies of at most this many words per operand run their
// two forward transforms on two streams: below it a transform leaves most
// of the chip idle (Goldilocks 2^22: 132 -> 111 us for the pair), above it
// the two compete for the same units (32-bit 2^24: 198 -> 213 us).
constexpr uint64_t POLYMUL_SIDE_MAX_WORDS = 1ull << 22;

// forward(a), forward(b) to bit-reversed order, then the inverse from
// bit-reversed order with b multiplied in as the first pass loads a
template <class FP>
static int polymul_run(stntt_ctx *ctx, const Plan &fwd, const Plan &inv, typename FP::T *a, typename FP::T *b,
                       uint64_t batches, uint32_t flags, hipStream_t s) {
    using T = typename FP::T;
    const uint32_t L = fwd.key.log_n;
    const uint32_t fflags = flags | STNTT_BITREV_OUT;
    int rc;
    if ((batches << L) <= POLYMUL_SIDE_MAX_WORDS) {
        // the two forward transforms at once: b's on the side stream
        const uint64_t need = schedule_scratch_bytes<FP>(
            transform_schedule<FP>(ctx, fwd.key.field, L, batches, fflags), batches << L, fflags);
        if ((rc = ctx_aux_streams(ctx, 1))) return rc;
        if (!ctx->side_fork) {
            HIP_TRY(hipEventCreateWithFlags(&ctx->side_fork, hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&ctx->side_join, hipEventDisableTiming));
        }
        hipStream_t side = ctx->aux[0];
        if (need && (rc = ensure_buffer(&ctx->side_scratch, &ctx->side_scratch_bytes, need))) return rc;
        HIP_TRY(hipEventRecord(ctx->side_fork, s));
        HIP_TRY(hipStreamWaitEvent(side, ctx->side_fork, 0));
        if ((rc = run_transform<FP>(ctx, fwd, a, batches, fflags, s))) return rc;
        if ((rc = run_transform<FP>(ctx, fwd, b, batches, fflags, side, nullptr, nullptr, nullptr,
                                    (T *)ctx->side_scratch)))
            return rc;
        HIP_TRY(hipEventRecord(ctx->side_join, side));
        HIP_TRY(hipStreamWaitEvent(s, ctx->side_join, 0));
    } else {
        if ((rc = run_transform<FP>(ctx, fwd, a, batches, fflags, s))) return rc;
        if ((rc = run_transform<FP>(ctx, fwd, b, batches, fflags, s))) return rc;
    }
    return run_transform<FP>(ctx, inv, a, batches, STNTT_BITREV_IN, s, b);
}

extern "C" {

const char *stntt_version(void) { return "stntt-mi355x 0.1 (gfx950)"; }

const char *stntt_last_error(void) { return g_err.c_str(); }

int stntt_ctx_create(int device, stntt_ctx **out) {
    if (!out) return fail(STNTT_EINVAL, "null out pointer");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(STNTT_EDEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(STNTT_EINVAL, "device %d out of range", device);
    auto ctx = std::make_unique<stntt_ctx>();
    ctx->device = device;
    auto setup = [&]() -> int {
        HIP_TRY(hipSetDevice(device));
        HIP_TRY(hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device));
        if (int rc = noalloc_setup(device)) return rc;
        HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&ctx->ev_start));
        HIP_TRY(hipEventCreate(&ctx->ev_stop));
        for (auto &e : ctx->prof_ev) HIP_TRY(hipEventCreate(&e));
        HIP_TRY(hipEventCreateWithFlags(&ctx->scratch_ev, hipEventDisableTiming));
        // the splits tuned on this architecture (tuned_gfx950.hpp); callers
        // may replace any of them with stntt_set_pass_split / _batch_split
        for (const TunedPassSplit &t : TUNED_PASS_SPLITS)
            if (int rc = stntt_set_pass_shape(ctx.get(), t.field, t.log2n, t.npasses, t.logr, t.logc)) return rc;
        for (const TunedBatchSplit &t : TUNED_BATCH_SPLITS)
            if (int rc = stntt_set_batch_split(ctx.get(), t.field, t.log2n, t.nsub, t.radix_logs)) return rc;
        return STNTT_OK;
    };
    if (int rc = setup()) {
        stntt_ctx_destroy(ctx.release());  // releases whatever was created
        return rc;
    }
    *out = ctx.release();
    return STNTT_OK;
}

int stntt_ctx_destroy(stntt_ctx *ctx) {
    if (!ctx) return STNTT_OK;
    (void)hipSetDevice(ctx->device);
    // transforms on callers' streams may still read the tables and scratch
    (void)hipDeviceSynchronize();
    for (auto &kv : ctx->plans) (void)hipFree(kv.second->d_tables);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->staging) (void)hipFree(ctx->staging);
    if (ctx->side_scratch) (void)hipFree(ctx->side_scratch);
    if (ctx->side_fork) (void)hipEventDestroy(ctx->side_fork);
    if (ctx->side_join) (void)hipEventDestroy(ctx->side_join);
    for (hipStream_t a : ctx->aux)
        if (a) (void)hipStreamDestroy(a);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    for (auto &e : ctx->prof_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->scratch_ev) (void)hipEventDestroy(ctx->scratch_ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return STNTT_OK;
}

stntt_stream stntt_ctx_stream(stntt_ctx *ctx) { return ctx ? (stntt_stream)ctx->stream : nullptr; }

int stntt_find_params(uint64_t n, uint64_t min_mod, uint64_t *root, uint64_t *mod) {
    if (!root || !mod) return fail(STNTT_EINVAL, "null output pointer");
    if (n == 0) return fail(STNTT_EINVAL, "n must be positive");
    // smallest prime p = k n + 1 >= min_mod (findModulus, ntt_util.cpp:16-24)
    uint64_t k = (min_mod > 1 ? (min_mod - 1 + n - 1) / n : 1);
    if (k < 1) k = 1;
    uint64_t p;
    for (;; ++k) {
        if (k > (UINT64_MAX - 1) / n) return fail(STNTT_EINVAL, "no prime k n + 1 >= %llu below 2^64",
                                                  (unsigned long long)min_mod);
        p = k * n + 1;
        if (is_prime_u64(p)) break;
    }
    // the smallest generator: c with c^((p-1)/q) != 1 for every prime q | p - 1
    // (findGenerator / isPrimitiveRoot, ntt_util.cpp:26-49)
    std::vector<uint64_t> f;
    distinct_prime_factors(p - 1, f);
    uint64_t g = 0;
    for (uint64_t c = 1; c < p && !g; ++c) {
        if (powmod(c, p - 1, p) != 1) continue;
        bool ok = true;
        for (uint64_t q : f)
            if (powmod(c, (p - 1) / q, p) == 1) {
                ok = false;
                break;
            }
        if (ok) g = c;
    }
    if (!g) return fail(STNTT_EINVAL, "no generator mod %llu", (unsigned long long)p);
    *mod = p;
    *root = powmod(g, (p - 1) / n, p);
    return STNTT_OK;
}

static constexpr uint32_t TRANSFORM_FLAGS = STNTT_SIGNED_INPUT | STNTT_BITREV_IN | STNTT_BITREV_OUT;

int stntt_ntt_u32(stntt_ctx *ctx, uint32_t *d_data, uint32_t log2n, uint64_t batches, uint32_t root, uint32_t mod,
                  int direction, uint32_t flags, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (flags & ~TRANSFORM_FLAGS) return fail(STNTT_EINVAL, "unknown flags 0x%x", flags);
    int rc = check_u32_params(log2n, root, mod, direction);
    if (rc) return rc;
    if (batches && !d_data) return fail(STNTT_EINVAL, "null data pointer");
    if ((rc = set_device(ctx))) return rc;
    const Plan *plan;
    if ((rc = get_plan(ctx, STNTT_FIELD_U32, log2n, root % mod, mod, direction, &plan))) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    return run_transform<FieldU32>(ctx, *plan, d_data, batches, flags, s);
}

int stntt_ntt_goldilocks(stntt_ctx *ctx, uint64_t *d_data, uint32_t log2n, uint64_t batches, int direction,
                         stntt_stream stream) {
    return stntt_ntt_goldilocks_ex(ctx, d_data, log2n, batches, direction, 0, stream);
}

int stntt_ntt_goldilocks_ex(stntt_ctx *ctx, uint64_t *d_data, uint32_t log2n, uint64_t batches, int direction,
                            uint32_t flags, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (flags & ~(STNTT_BITREV_IN | STNTT_BITREV_OUT))
        return fail(STNTT_EINVAL, "flags 0x%x (Goldilocks takes the order flags only)", flags);
    if (log2n > 32) return fail(STNTT_EINVAL, "log2n %u exceeds the 2-adicity 32 of P - 1", log2n);
    if (direction != STNTT_FORWARD && direction != STNTT_INVERSE)
        return fail(STNTT_EINVAL, "bad direction %d", direction);
    if (batches && !d_data) return fail(STNTT_EINVAL, "null data pointer");
    int rc;
    if ((rc = set_device(ctx))) return rc;
    const Plan *plan;
    if ((rc = get_plan(ctx, STNTT_FIELD_GOLDILOCKS, log2n, 0, GL_P, direction, &plan))) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    return run_transform<FieldGL>(ctx, *plan, d_data, batches, flags, s);
}

int stntt_ntt_goldilocks_noalloc(stntt_ctx *ctx, uint64_t *d_data, uint32_t log2n, uint64_t batches, int direction,
                                 uint32_t flags, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (flags & ~(STNTT_BITREV_IN | STNTT_BITREV_OUT))
        return fail(STNTT_EINVAL, "flags 0x%x (Goldilocks takes the order flags only)", flags);
    if (log2n > 28) return fail(STNTT_EINVAL, "log2n %u above the allocation-free limit 28", log2n);
    if (direction != STNTT_FORWARD && direction != STNTT_INVERSE)
        return fail(STNTT_EINVAL, "bad direction %d", direction);
    if (batches && !d_data) return fail(STNTT_EINVAL, "null data pointer");
    if (batches == 0 || log2n == 0) return STNTT_OK;
    if (batches >= (1ull << 31) || (batches << log2n) >> log2n != batches)
        return fail(STNTT_EINVAL, "%llu transforms of 2^%u words is more than one call handles",
                    (unsigned long long)batches, log2n);
    // the schedule run_transform will run (order flags rule out large-radix passes)
    const Schedule sch = transform_schedule<FieldGL>(ctx, STNTT_FIELD_GOLDILOCKS, log2n, batches, flags);
    const uint64_t need = schedule_scratch_bytes<FieldGL>(sch, batches << log2n, flags);
    if (need > NOALLOC_DATA_BYTES)
        return fail(STNTT_EINVAL,
                    "%llu transforms of 2^%u words need %llu bytes of scratch; the allocation-free path holds %llu",
                    (unsigned long long)batches, log2n, (unsigned long long)need,
                    (unsigned long long)NOALLOC_DATA_BYTES);
    int rc;
    if ((rc = set_device(ctx))) return rc;
    void *base = noalloc_ptr(ctx->device);
    if (!base) return fail(STNTT_EDEVICE, "no allocation-free scratch on device %d", ctx->device);
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    // the plan's factor tables, regenerated into the array by every call so
    // that each call (and each replay of a captured one) stands alone
    Plan plan{};
    plan.key = PlanKey{STNTT_FIELD_GOLDILOCKS, direction, log2n, 0, GL_P};
    plan.log_rmax = log2n < MAX_LOGR ? log2n : MAX_LOGR;
    plan.d_tables = base;
    const uint64_t n = 1ull << log2n, rmax = 1ull << plan.log_rmax, nlo = n >> plan.log_rmax;
    uint64_t w = powmod(7, (GL_P - 1) >> log2n, GL_P);
    if (direction == STNTT_INVERSE) w = invmod(w, GL_P);
    plan.rt_gl.ninv = invmod(n % GL_P, GL_P);
    hipLaunchKernelGGL(gen_tables_kernel<FieldGL>, dim3((uint32_t)((rmax + nlo + 255) / 256)), dim3(256), 0, s,
                       (uint64_t *)base, (uint32_t)rmax, (uint32_t)nlo, w, powmod(w, nlo, GL_P), plan.rt_gl);
    HIP_TRY(hipGetLastError());
    return run_transform<FieldGL>(ctx, plan, d_data, batches, flags, s, nullptr, nullptr, nullptr,
                                  (uint64_t *)((char *)base + NOALLOC_TABLE_BYTES));
}

int stntt_pointwise_u32(stntt_ctx *ctx, uint32_t *d_a, const uint32_t *d_b, uint64_t count, uint32_t mod,
                        stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (mod < 2 || mod >= (1u << 31)) return fail(STNTT_EINVAL, "mod %u outside [2, 2^31)", mod);
    if (count && (!d_a || !d_b)) return fail(STNTT_EINVAL, "null data pointer");
    int rc;
    if ((rc = set_device(ctx))) return rc;
    FieldU32::Rt rt{};
    rt.p = mod;
    rt.mu = ~0ull / mod;
    return launch_pointwise<FieldU32>(d_a, d_b, count, rt, stream ? (hipStream_t)stream : ctx->stream);
}

int stntt_pointwise_goldilocks(stntt_ctx *ctx, uint64_t *d_a, const uint64_t *d_b, uint64_t count,
                               stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (count && (!d_a || !d_b)) return fail(STNTT_EINVAL, "null data pointer");
    int rc;
    if ((rc = set_device(ctx))) return rc;
    return launch_pointwise<FieldGL>(d_a, d_b, count, FieldGL::Rt{}, stream ? (hipStream_t)stream : ctx->stream);
}

int stntt_polymul_u32(stntt_ctx *ctx, uint32_t *d_a, uint32_t *d_b, uint32_t log2n, uint64_t batches,
                      uint32_t root, uint32_t mod, uint32_t flags, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (flags & ~STNTT_SIGNED_INPUT) return fail(STNTT_EINVAL, "flags 0x%x (only STNTT_SIGNED_INPUT)", flags);
    int rc = check_u32_params(log2n, root, mod, STNTT_INVERSE);
    if (rc) return rc;
    if (batches && (!d_a || !d_b)) return fail(STNTT_EINVAL, "null data pointer");
    if (d_a == d_b && batches) return fail(STNTT_EINVAL, "a and b must be distinct buffers");
    if ((rc = set_device(ctx))) return rc;
    const Plan *fwd, *inv;
    if ((rc = get_plan(ctx, STNTT_FIELD_U32, log2n, root % mod, mod, STNTT_FORWARD, &fwd))) return rc;
    if ((rc = get_plan(ctx, STNTT_FIELD_U32, log2n, root % mod, mod, STNTT_INVERSE, &inv))) return rc;
    return polymul_run<FieldU32>(ctx, *fwd, *inv, d_a, d_b, batches, flags,
                                 stream ? (hipStream_t)stream : ctx->stream);
}

int stntt_polymul_goldilocks(stntt_ctx *ctx, uint64_t *d_a, uint64_t *d_b, uint32_t log2n, uint64_t batches,
                             stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (log2n > 32) return fail(STNTT_EINVAL, "log2n %u exceeds the 2-adicity 32 of P - 1", log2n);
    if (batches && (!d_a || !d_b)) return fail(STNTT_EINVAL, "null data pointer");
    if (d_a == d_b && batches) return fail(STNTT_EINVAL, "a and b must be distinct buffers");
    int rc;
    if ((rc = set_device(ctx))) return rc;
    const Plan *fwd, *inv;
    if ((rc = get_plan(ctx, STNTT_FIELD_GOLDILOCKS, log2n, 0, GL_P, STNTT_FORWARD, &fwd))) return rc;
    if ((rc = get_plan(ctx, STNTT_FIELD_GOLDILOCKS, log2n, 0, GL_P, STNTT_INVERSE, &inv))) return rc;
    return polymul_run<FieldGL>(ctx, *fwd, *inv, d_a, d_b, batches, 0, stream ? (hipStream_t)stream : ctx->stream);
}

int stntt_ntt_u64(stntt_ctx *ctx, uint64_t *d_data, uint32_t log2n, uint64_t batches, uint64_t root, uint64_t mod,
                  int direction, uint32_t flags, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (flags & ~(STNTT_BITREV_IN | STNTT_BITREV_OUT))
        return fail(STNTT_EINVAL, "flags 0x%x (the 64-bit field takes the order flags only)", flags);
    int rc = check_u64_params(log2n, root, mod, direction);
    if (rc) return rc;
    if (batches && !d_data) return fail(STNTT_EINVAL, "null data pointer");
    if ((rc = set_device(ctx))) return rc;
    const Plan *plan;
    if ((rc = get_plan(ctx, STNTT_FIELD_U64, log2n, root % mod, mod, direction, &plan))) return rc;
    return run_transform<FieldM64>(ctx, *plan, d_data, batches, flags, stream ? (hipStream_t)stream : ctx->stream);
}

int stntt_pointwise_u64(stntt_ctx *ctx, uint64_t *d_a, const uint64_t *d_b, uint64_t count, uint64_t mod,
                        stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (mod < 3 || !(mod & 1)) return fail(STNTT_EINVAL, "mod %llu must be odd and at least 3", (unsigned long long)mod);
    if (count && (!d_a || !d_b)) return fail(STNTT_EINVAL, "null data pointer");
    int rc;
    if ((rc = set_device(ctx))) return rc;
    M64Rt rt{};
    rt.p = mod;
    rt.pinv = m64_neg_inverse(mod);
    rt.r1 = (uint64_t)(((u128)1 << 64) % mod);
    rt.r2 = (uint64_t)((u128)rt.r1 * rt.r1 % mod);
    // mulg(a, b) = mont(a, mont(b, R^2)) = a b mod p, exact for any 64-bit
    // words: each Montgomery product has one operand below p
    return launch_pointwise<FieldM64>(d_a, d_b, count, rt, stream ? (hipStream_t)stream : ctx->stream);
}

int stntt_polymul_u64(stntt_ctx *ctx, uint64_t *d_a, uint64_t *d_b, uint32_t log2n, uint64_t batches, uint64_t root,
                      uint64_t mod, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    int rc = check_u64_params(log2n, root, mod, STNTT_INVERSE);
    if (rc) return rc;
    if (batches && (!d_a || !d_b)) return fail(STNTT_EINVAL, "null data pointer");
    if (d_a == d_b && batches) return fail(STNTT_EINVAL, "a and b must be distinct buffers");
    if ((rc = set_device(ctx))) return rc;
    const Plan *fwd, *inv;
    if ((rc = get_plan(ctx, STNTT_FIELD_U64, log2n, root % mod, mod, STNTT_FORWARD, &fwd))) return rc;
    if ((rc = get_plan(ctx, STNTT_FIELD_U64, log2n, root % mod, mod, STNTT_INVERSE, &inv))) return rc;
    return polymul_run<FieldM64>(ctx, *fwd, *inv, d_a, d_b, batches, 0, stream ? (hipStream_t)stream : ctx->stream);
}

int stntt_twiddles_u64(stntt_ctx *ctx, uint32_t log2n, uint64_t root, uint64_t mod, int direction, uint64_t *host_out) {
    if (!ctx || !host_out) return fail(STNTT_EINVAL, "null argument");
    int rc = check_u64_params(log2n, root, mod, direction);
    if (rc) return rc;
    if (log2n == 0) return STNTT_OK;
    if ((rc = set_device(ctx))) return rc;
    const Plan *plan;
    if ((rc = get_plan(ctx, STNTT_FIELD_U64, log2n, root % mod, mod, direction, &plan))) return rc;
    uint64_t count = (1ull << log2n) / 2;
    if ((rc = ensure_buffer(&ctx->staging, &ctx->staging_bytes, count * 8))) return rc;
    hipLaunchKernelGGL(expand_twiddles_kernel<FieldM64>, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0,
                       ctx->stream, (uint64_t *)ctx->staging, (const uint64_t *)plan->hi(),
                       (const uint64_t *)plan->lo(), log2n - plan->log_rmax, count, plan->rt_m64);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(host_out, ctx->staging, count * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return STNTT_OK;
}

int stntt_twiddles_u32(stntt_ctx *ctx, uint64_t n, uint32_t root, uint32_t mod, int direction, uint32_t *host_out) {
    if (!ctx || !host_out) return fail(STNTT_EINVAL, "null argument");
    uint32_t L;
    int rc = log2_exact(n, &L);
    if (rc) return rc;
    if ((rc = check_u32_params(L, root, mod, direction))) return rc;
    if (n < 2) return STNTT_OK;
    if ((rc = set_device(ctx))) return rc;
    const Plan *plan;
    if ((rc = get_plan(ctx, STNTT_FIELD_U32, L, root % mod, mod, direction, &plan))) return rc;
    uint64_t count = n / 2;
    if ((rc = ensure_buffer(&ctx->staging, &ctx->staging_bytes, count * 4))) return rc;
    hipLaunchKernelGGL(expand_twiddles_kernel<FieldU32>, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0,
                       ctx->stream, (uint32_t *)ctx->staging, (const uint32_t *)plan->hi(), (const uint32_t *)plan->lo(),
                       L - plan->log_rmax, count, plan->rt_u32);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(host_out, ctx->staging, count * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return STNTT_OK;
}

int stntt_twiddles_goldilocks(stntt_ctx *ctx, uint32_t log2n, int direction, uint64_t *host_out) {
    if (!ctx || !host_out) return fail(STNTT_EINVAL, "null argument");
    if (log2n > 32) return fail(STNTT_EINVAL, "log2n %u too large", log2n);
    if (log2n == 0) return STNTT_OK;
    int rc;
    if ((rc = set_device(ctx))) return rc;
    const Plan *plan;
    if ((rc = get_plan(ctx, STNTT_FIELD_GOLDILOCKS, log2n, 0, GL_P, direction, &plan))) return rc;
    uint64_t count = (1ull << log2n) / 2;
    if ((rc = ensure_buffer(&ctx->staging, &ctx->staging_bytes, count * 8))) return rc;
    hipLaunchKernelGGL(expand_twiddles_kernel<FieldGL>, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0,
                       ctx->stream, (uint64_t *)ctx->staging, (const uint64_t *)plan->hi(),
                       (const uint64_t *)plan->lo(), log2n - plan->log_rmax, count, plan->rt_gl);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(host_out, ctx->staging, count * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return STNTT_OK;
}

int stntt_auto_ntt_i32(stntt_ctx *ctx, int32_t *host_vec, size_t host_len, uint32_t n, uint32_t batches, uint32_t root,
                       uint32_t mod, int direction, float *kernel_ms) {
    return auto_ntt_common<int32_t>(ctx, host_vec, host_len, n, batches, kernel_ms, [&](uint32_t L, int32_t *d) {
        return stntt_ntt_u32(ctx, (uint32_t *)d, L, batches, root, mod, direction, STNTT_SIGNED_INPUT, nullptr);
    });
}

int stntt_auto_ntt_i32_ex(stntt_ctx *ctx, int32_t *host_vec, size_t host_len, uint32_t n, uint32_t batches,
                          uint32_t root, uint32_t mod, int direction, uint32_t flags, float *kernel_ms) {
    return auto_ntt_common<int32_t>(ctx, host_vec, host_len, n, batches, kernel_ms, [&](uint32_t L, int32_t *d) {
        return stntt_ntt_u32(ctx, (uint32_t *)d, L, batches, root, mod, direction, flags | STNTT_SIGNED_INPUT,
                             nullptr);
    });
}

int stntt_auto_polymul_i32(stntt_ctx *ctx, int32_t *host_a, int32_t *host_b, uint32_t n, uint32_t root, uint32_t mod,
                           float *kernel_ms) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (!host_a || !host_b) return fail(STNTT_EINVAL, "null host vector");
    uint32_t L;
    int rc = log2_exact(n, &L);
    if (rc) return rc;
    if ((rc = set_device(ctx))) return rc;
    const size_t bytes = (size_t)n * 4;
    if ((rc = ensure_buffer(&ctx->staging, &ctx->staging_bytes, 2 * bytes))) return rc;
    int32_t *da = (int32_t *)ctx->staging, *db = da + n;
    HIP_TRY(hipMemcpyAsync(da, host_a, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(db, host_b, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev_start, ctx->stream));
    if ((rc = stntt_polymul_u32(ctx, (uint32_t *)da, (uint32_t *)db, L, 1, root, mod, STNTT_SIGNED_INPUT, nullptr)))
        return rc;
    HIP_TRY(hipEventRecord(ctx->ev_stop, ctx->stream));
    HIP_TRY(hipMemcpyAsync(host_a, da, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(host_b, db, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (kernel_ms) *kernel_ms = ms;
    return STNTT_OK;
}

int stntt_auto_ntt_goldilocks(stntt_ctx *ctx, uint64_t *host_vec, size_t host_len, uint32_t n, uint32_t batches,
                              int direction, float *kernel_ms) {
    return auto_ntt_common<uint64_t>(ctx, host_vec, host_len, n, batches, kernel_ms, [&](uint32_t L, uint64_t *d) {
        return stntt_ntt_goldilocks(ctx, d, L, batches, direction, nullptr);
    });
}

static int check_steps_args(uint32_t log2n, uint32_t mod, int nsteps) {
    if (log2n > 30) return fail(STNTT_EINVAL, "log2n %u too large", log2n);
    if (mod < 2 || mod >= (1u << 31)) return fail(STNTT_EINVAL, "mod %u outside [2, 2^31)", mod);
    if (nsteps > MAX_STAGE_STEPS) return fail(STNTT_EINVAL, "nsteps %d above %d", nsteps, MAX_STAGE_STEPS);
    return STNTT_OK;
}

int stntt_stockham_steps_u32(stntt_ctx *ctx, uint32_t *d_data, uint32_t log2n, uint64_t batches, uint32_t root,
                             uint32_t mod, int nsteps, uint32_t flags, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (flags & ~STNTT_SIGNED_INPUT) return fail(STNTT_EINVAL, "flags 0x%x (only STNTT_SIGNED_INPUT)", flags);
    int rc = check_steps_args(log2n, mod, nsteps);
    if (rc) return rc;
    if (batches && !d_data) return fail(STNTT_EINVAL, "null data pointer");
    if ((rc = set_device(ctx))) return rc;
    return stockham_steps<true>(ctx, d_data, log2n, batches, root, mod, nsteps, flags & STNTT_SIGNED_INPUT,
                                stream ? (hipStream_t)stream : ctx->stream);
}

int stntt_stockham_idx_u32(stntt_ctx *ctx, uint32_t *d_data, uint32_t log2n, uint64_t batches, uint32_t mod,
                           int nsteps, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    int rc = check_steps_args(log2n, mod, nsteps);
    if (rc) return rc;
    if (batches && !d_data) return fail(STNTT_EINVAL, "null data pointer");
    if ((rc = set_device(ctx))) return rc;
    return stockham_steps<false>(ctx, d_data, log2n, batches, 0, mod, nsteps, false,
                                 stream ? (hipStream_t)stream : ctx->stream);
}

int stntt_auto_stockham_steps_i32(stntt_ctx *ctx, int32_t *host_vec, size_t host_len, uint32_t n, uint32_t batches,
                                  uint32_t root, uint32_t mod, int nsteps, float *kernel_ms) {
    return auto_ntt_common<int32_t>(ctx, host_vec, host_len, n, batches, kernel_ms, [&](uint32_t L, int32_t *d) {
        return stntt_stockham_steps_u32(ctx, (uint32_t *)d, L, batches, root, mod, nsteps, STNTT_SIGNED_INPUT,
                                        nullptr);
    });
}

int stntt_auto_stockham_idx_i32(stntt_ctx *ctx, int32_t *host_vec, size_t host_len, uint32_t n, uint32_t batches,
                                uint32_t mod, int nsteps, float *kernel_ms) {
    return auto_ntt_common<int32_t>(ctx, host_vec, host_len, n, batches, kernel_ms, [&](uint32_t L, int32_t *d) {
        return stntt_stockham_idx_u32(ctx, (uint32_t *)d, L, batches, mod, nsteps, nullptr);
    });
}

int stntt_batch_create(stntt_ctx *ctx, const stntt_desc *descs, size_t count, stntt_batch **out) {
    if (!ctx || !out || (count && !descs)) return fail(STNTT_EINVAL, "null argument");
    if (count > 0xFFFFFFFFull) return fail(STNTT_EINVAL, "too many descriptors");
    int rc;
    if ((rc = set_device(ctx))) return rc;
    auto b = std::make_unique<stntt_batch>();
    b->ctx = ctx;
    if (hipEventCreateWithFlags(&b->run_ev, hipEventDisableTiming) != hipSuccess) {
        b->run_ev = nullptr;
        return fail(STNTT_EDEVICE, "hipEventCreateWithFlags failed");
    }
    if ((rc = batch_build(ctx, descs, count, b.get()))) {
        batch_free(b.get());
        return rc;
    }
    *out = b.release();
    return STNTT_OK;
}

// order a run of b on s after the previous one (another stream: an event
// wait), and mark this one's end
static int batch_order_begin(stntt_batch *b, hipStream_t s) {
    if (b->run_last && b->run_last != s && !stream_capturing(s)) HIP_TRY(hipStreamWaitEvent(s, b->run_ev, 0));
    return STNTT_OK;
}
static int batch_order_end(stntt_batch *b, hipStream_t s) {
    if (stream_capturing(s)) return STNTT_OK;
    HIP_TRY(hipEventRecord(b->run_ev, s));
    b->run_last = s;
    return STNTT_OK;
}

int stntt_batch_run(stntt_batch *b, stntt_stream stream) {
    if (!b) return fail(STNTT_EINVAL, "null batch");
    int rc;
    if ((rc = set_device(b->ctx))) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : b->ctx->stream;
    if ((rc = batch_order_begin(b, s))) return rc;
    if ((rc = batch_launch(b, s))) return rc;
    return batch_order_end(b, s);
}

int stntt_batch_capture(stntt_batch *b, stntt_stream stream) {
    if (!b) return fail(STNTT_EINVAL, "null batch");
    int rc;
    if ((rc = set_device(b->ctx))) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : b->ctx->stream;
    if (b->graph_exec) return STNTT_OK;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = batch_launch(b, s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(STNTT_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
    return STNTT_OK;
}

int stntt_batch_run_graph(stntt_batch *b, stntt_stream stream) {
    if (!b) return fail(STNTT_EINVAL, "null batch");
    if (!b->graph_exec) return fail(STNTT_EINVAL, "batch not captured");
    int rc;
    if ((rc = set_device(b->ctx))) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : b->ctx->stream;
    if ((rc = batch_order_begin(b, s))) return rc;
    HIP_TRY(hipGraphLaunch(b->graph_exec, s));
    return batch_order_end(b, s);
}

int stntt_batch_run_plain(stntt_batch *b, stntt_stream stream) {
    if (!b) return fail(STNTT_EINVAL, "null batch");
    int rc;
    if ((rc = set_device(b->ctx))) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : b->ctx->stream;
    for (const stntt_desc &d : b->descs) {
        size_t bytes = (size_t)(d.field == STNTT_FIELD_GOLDILOCKS ? 8 : 4) << d.log2n;
        if (d.src != d.dst) HIP_TRY(hipMemcpyAsync(d.dst, d.src, bytes, hipMemcpyDeviceToDevice, s));
        if (d.field == STNTT_FIELD_GOLDILOCKS) rc = stntt_ntt_goldilocks(b->ctx, (uint64_t *)d.dst, d.log2n, 1,
                                                                         d.direction, s);
        else rc = stntt_ntt_u32(b->ctx, (uint32_t *)d.dst, d.log2n, 1, d.root, d.mod, d.direction, d.flags, s);
        if (rc) return rc;
    }
    return STNTT_OK;
}

int stntt_batch_destroy(stntt_batch *b) {
    if (!b) return STNTT_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    if (b->run_last) (void)hipEventSynchronize(b->run_ev);  // the last run, on any stream
    batch_free(b);
    delete b;
    return STNTT_OK;
}

int stntt_set_batch_split(stntt_ctx *ctx, int field, uint32_t log2n, uint32_t nsub, uint32_t radix_logs) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (field != STNTT_FIELD_U32 && field != STNTT_FIELD_GOLDILOCKS) return fail(STNTT_EINVAL, "bad field");
    if (log2n > 18) return fail(STNTT_EINVAL, "log2n %u above the batched limit 18", log2n);
    if (nsub == 0) {
        ctx->batch_split.erase({field, log2n});
        return STNTT_OK;
    }
    int rc = batch_validate_split(log2n, batch_global_path(field, log2n), nsub, radix_logs);
    if (rc) return rc;
    ctx->batch_split[{field, log2n}] = {nsub, radix_logs};
    return STNTT_OK;
}

// Frees everything a server holds once its thread has stopped (or was never
// started): streams, worker contexts and buffers, cached device buffers,
// events.  Used by destroy and by create's failure paths.
static void server_release(stntt_server *srv) {
    (void)hipSetDevice(srv->device);
    if (srv->up) (void)hipStreamSynchronize(srv->up);
    if (srv->down) (void)hipStreamSynchronize(srv->down);
    for (auto &W : srv->workers) {
        if (W.stream) (void)hipStreamSynchronize(W.stream);
        if (W.h_tasks) (void)hipHostFree(W.h_tasks);
        if (W.d_tasks) (void)hipFree(W.d_tasks);
        if (W.d_counter) (void)hipFree(W.d_counter);
        if (W.d_scratch) (void)hipFree(W.d_scratch);
        if (W.h_tab) (void)hipHostFree(W.h_tab);
        if (W.d_tab) (void)hipFree(W.d_tab);
        if (W.ctx) stntt_ctx_destroy(W.ctx);
    }
    srv->workers.clear();
    for (auto &kv : srv->cache)
        for (void *p : kv.second) (void)hipFree(p);
    srv->cache.clear();
    for (auto &kv : srv->reqs)
        if (kv.second.d_buf) (void)hipFree(kv.second.d_buf);
    for (hipEvent_t e : srv->free_events) (void)hipEventDestroy(e);
    srv->free_events.clear();
    if (srv->up) (void)hipStreamDestroy(srv->up);
    if (srv->down) (void)hipStreamDestroy(srv->down);
    srv->up = srv->down = nullptr;
}

static int server_setup(stntt_server *S, int device, uint32_t workers) {
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&S->up, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&S->down, hipStreamNonBlocking));
    for (uint32_t i = 0; i < workers; ++i) {
        S->workers.emplace_back();
        ServerWorker &W = S->workers.back();
        int rc = stntt_ctx_create(device, &W.ctx);
        if (rc) return rc;
        W.stream = W.ctx->stream;
        // scratch for the largest served multi-pass transform, allocated now so
        // that no later growth synchronises the device
        if ((rc = ensure_buffer(&W.ctx->scratch, &W.ctx->scratch_bytes, 2 * (8ull << SERVER_MAX_LOG)))) return rc;
        HIP_TRY(hipHostMalloc((void **)&W.h_tasks, SERVER_WORKER_ROUNDS * SERVER_MAX_ROUND * sizeof(BatchTask)));
        HIP_TRY(hipMalloc(&W.d_tasks, SERVER_MAX_ROUND * sizeof(BatchTask)));
        HIP_TRY(hipMalloc(&W.d_counter, sizeof(uint32_t)));
        HIP_TRY(hipMalloc(&W.d_scratch, (size_t)SERVER_BATCH_GRID * (8ull << SERVER_BATCH_MAX_LOG)));
        HIP_TRY(hipHostMalloc((void **)&W.h_tab, SERVER_WORKER_ROUNDS * SERVER_MAX_ROUND * sizeof(void *)));
        HIP_TRY(hipMalloc((void **)&W.d_tab, SERVER_MAX_ROUND * sizeof(void *)));
    }
    return STNTT_OK;
}

int stntt_server_create(int device, uint64_t hbm_budget_bytes, uint32_t workers, stntt_server **out) {
    if (!out) return fail(STNTT_EINVAL, "null out pointer");
    if (workers < 1 || workers > 16) return fail(STNTT_EINVAL, "workers %u outside [1, 16]", workers);
    const uint64_t reserve = workers * server_worker_reserve(SERVER_MAX_LOG);
    if (hbm_budget_bytes <= reserve)
        return fail(STNTT_EINVAL, "budget %llu does not cover the worker reservation %llu",
                    (unsigned long long)hbm_budget_bytes, (unsigned long long)reserve);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return fail(STNTT_EINVAL, "device %d out of range", device);
    auto S = std::make_unique<stntt_server>();
    S->device = device;
    S->budget = hbm_budget_bytes - reserve;
    if (int rc = server_setup(S.get(), device, workers)) {
        server_release(S.get());
        return rc;
    }
    stntt_server *raw = S.get();
    S->thread = std::thread(server_loop, raw);
    *out = S.release();
    return STNTT_OK;
}

int stntt_server_destroy(stntt_server *srv) {
    if (!srv) return STNTT_OK;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    if (srv->thread.joinable()) srv->thread.join();
    server_release(srv);
    delete srv;
    return STNTT_OK;
}

int stntt_server_host_alloc(stntt_server *srv, size_t bytes, void **host_ptr) {
    if (!srv || !host_ptr) return fail(STNTT_EINVAL, "null argument");
    if (hipHostMalloc(host_ptr, bytes ? bytes : 1) != hipSuccess) {
        *host_ptr = nullptr;
        return fail(STNTT_ENOMEM, "hipHostMalloc of %zu bytes failed", bytes);
    }
    return STNTT_OK;
}

int stntt_server_host_free(stntt_server *srv, void *host_ptr) {
    (void)srv;  // pinned host memory does not depend on the server; it may be gone
    HIP_TRY(hipHostFree(host_ptr));
    return STNTT_OK;
}

int stntt_server_submit(stntt_server *srv, const stntt_request *req, uint64_t *id) {
    if (!srv || !req || !id) return fail(STNTT_EINVAL, "null argument");
    int rc = server_check_request(*req);
    if (rc) return rc;
    const uint64_t bytes = elem_bytes(req->field) << req->log2n;
    if (req->log2n > SERVER_MAX_LOG) return fail(STNTT_EINVAL, "log2n %u above the served maximum %u", req->log2n,
                                                 SERVER_MAX_LOG);
    if (bytes > srv->budget) return fail(STNTT_EINVAL, "request of %llu bytes exceeds the HBM budget",
                                         (unsigned long long)bytes);
    {
        std::lock_guard<std::mutex> g(srv->mu);
        ServerRequest r;
        r.id = srv->next_id++;
        r.req = *req;
        r.t_submit = Clock::now();
        r.d_bytes = bytes;
        r.cost = bytes;
        if (!srv->any) {
            srv->t_first = r.t_submit;
            srv->any = true;
        }
        *id = r.id;
        srv->reqs.emplace(r.id, r);
        srv->pending.push_back(r.id);
    }
    srv->cv_work.notify_one();
    return STNTT_OK;
}

int stntt_server_wait(stntt_server *srv, uint64_t id, double *latency_ms) {
    if (!srv) return fail(STNTT_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    if (srv->reqs.find(id) == srv->reqs.end())
        return fail(STNTT_ENOTFOUND, "unknown request %llu", (unsigned long long)id);
    // look the request up again after every wake-up: another thread waiting
    // on the same id may have collected (erased) it meanwhile
    auto it = srv->reqs.end();
    srv->cv_done.wait(lk, [&] {
        it = srv->reqs.find(id);
        return it == srv->reqs.end() || it->second.done;
    });
    if (it == srv->reqs.end())
        return fail(STNTT_ENOTFOUND, "request %llu was collected by another waiter", (unsigned long long)id);
    int status = it->second.status;
    if (latency_ms) *latency_ms = it->second.latency_ms;
    srv->reqs.erase(it);
    if (status) return fail(status, "request %llu failed on the device", (unsigned long long)id);
    return STNTT_OK;
}

int stntt_server_poll(stntt_server *srv, uint64_t id, int *done) {
    if (!srv || !done) return fail(STNTT_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(srv->mu);
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return fail(STNTT_ENOTFOUND, "unknown request %llu", (unsigned long long)id);
    *done = it->second.done ? 1 : 0;
    return STNTT_OK;
}

int stntt_server_stats_get(stntt_server *srv, stntt_server_stats *out) {
    if (!srv || !out) return fail(STNTT_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(srv->mu);
    *out = stntt_server_stats{};
    out->completed = srv->completed;
    out->peak_bytes = srv->peak_used;
    out->held_bytes = srv->used + srv->cached;
    out->rounds = srv->rounds_launched;
    out->batched = srv->batched_requests;
    out->grouped = srv->grouped_requests;
    if (!srv->latencies.empty()) {
        std::vector<double> v = srv->latencies;
        std::sort(v.begin(), v.end());
        auto pct = [&](double q) { return v[std::min(v.size() - 1, (size_t)(q * (v.size() - 1) + 0.5))]; };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        out->max_ms = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        out->mean_ms = sum / v.size();
        out->span_s = std::chrono::duration<double>(srv->t_last - srv->t_first).count();
    }
    return STNTT_OK;
}

int stntt_server_reset_stats(stntt_server *srv) {
    if (!srv) return fail(STNTT_EINVAL, "null server");
    std::lock_guard<std::mutex> g(srv->mu);
    srv->latencies.clear();
    srv->completed = 0;
    srv->peak_used = srv->used + srv->cached;
    srv->rounds_launched = 0;
    srv->batched_requests = 0;
    srv->grouped_requests = 0;
    srv->any = false;
    return STNTT_OK;
}

int stntt_set_pass_split(stntt_ctx *ctx, int field, uint32_t log2n, uint32_t npasses, const uint32_t *logr) {
    return stntt_set_pass_shape(ctx, field, log2n, npasses, logr, nullptr);
}

int stntt_set_pass_shape(stntt_ctx *ctx, int field, uint32_t log2n, uint32_t npasses, const uint32_t *logr,
                         const int32_t *logc) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (field != STNTT_FIELD_U32 && field != STNTT_FIELD_GOLDILOCKS && field != STNTT_FIELD_U64)
        return fail(STNTT_EINVAL, "bad field");
    if (npasses == 0) {
        ctx->pass_split.erase({field, log2n});
        return STNTT_OK;
    }
    if (npasses < 2 || npasses > 4 || !logr) return fail(STNTT_EINVAL, "pass count %u outside [2, 4]", npasses);
    const int tlogc = field == STNTT_FIELD_U32 ? tile_logc<FieldU32>() : tile_logc<FieldGL>();
    uint32_t sum = 0;
    std::vector<std::pair<int, int>> v;
    for (uint32_t i = 0; i < npasses; ++i) {
        const uint32_t r = logr[i];
        const int want = logc ? logc[i] : -1;
        if (r < (uint32_t)MIN_LOGR_TILE || r > (uint32_t)MAX_LOGR_BIG)
            return fail(STNTT_EINVAL, "pass radix 2^%u outside [2^%d, 2^%d]", r, MIN_LOGR_TILE, MAX_LOGR_BIG);
        if (r <= (uint32_t)MAX_LOGR_TILE) {
            // a tile pass needs n / R >= C columns, and a later pass J >= C
            if (want >= 0 && want != tlogc)
                return fail(STNTT_EINVAL, "pass %u: a radix-2^%u tile pass has 2^%d columns", i, r, tlogc);
            if (i > 0 && sum < (uint32_t)tlogc) return fail(STNTT_EINVAL, "pass %u starts with J < %d columns", i, 1 << tlogc);
            if (log2n < r + (uint32_t)tlogc) return fail(STNTT_EINVAL, "pass %u has fewer than %d columns", i, 1 << tlogc);
        } else {
            // a large-radix pass: some tile width from the full 128 KiB down to a quarter must fit
            int lc = -1;
            if (field == STNTT_FIELD_U32)
                lc = big_logc_choice<FieldU32>((int)log2n, (int)r, (int)sum, i == 0, 1, ctx->num_cus, want);
            else
                lc = big_logc_choice<FieldGL>((int)log2n, (int)r, (int)sum, i == 0, 1, ctx->num_cus, want);
            if (lc < 0) return fail(STNTT_EINVAL, "pass %u: no tile shape for radix 2^%u", i, r);
        }
        sum += r;
        v.push_back({(int)r, want});
    }
    if (sum != log2n) return fail(STNTT_EINVAL, "pass radices cover 2^%u, not 2^%u", sum, log2n);
    ctx->pass_split[{field, log2n}] = v;
    return STNTT_OK;
}

int stntt_dev_alloc(stntt_ctx *ctx, size_t bytes, void **d_ptr) {
    if (!ctx || !d_ptr) return fail(STNTT_EINVAL, "null argument");
    int rc;
    if ((rc = set_device(ctx))) return rc;
    if (hipMalloc(d_ptr, bytes ? bytes : 1) != hipSuccess) {
        *d_ptr = nullptr;
        return fail(STNTT_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
    return STNTT_OK;
}

int stntt_dev_free(stntt_ctx *ctx, void *d_ptr) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    int rc;
    if ((rc = set_device(ctx))) return rc;
    HIP_TRY(hipFree(d_ptr));
    return STNTT_OK;
}

int stntt_memcpy_htod(stntt_ctx *ctx, void *d_dst, const void *h_src, size_t bytes, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    HIP_TRY(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return STNTT_OK;
}

int stntt_memcpy_dtoh(stntt_ctx *ctx, void *h_dst, const void *d_src, size_t bytes, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    HIP_TRY(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return STNTT_OK;
}

int stntt_memcpy_dtod(stntt_ctx *ctx, void *d_dst, const void *d_src, size_t bytes, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    if (bytes % 8) return fail(STNTT_EINVAL, "byte count %zu is not a multiple of 8", bytes);
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    uint64_t count = bytes / 8;
    if (count == 0) return STNTT_OK;
    hipLaunchKernelGGL(copy_kernel<FieldGL>, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0, s,
                       (const uint64_t *)d_src, (uint64_t *)d_dst, count);
    HIP_TRY(hipGetLastError());
    return STNTT_OK;
}

int stntt_stream_sync(stntt_ctx *ctx, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    HIP_TRY(hipStreamSynchronize(stream ? (hipStream_t)stream : ctx->stream));
    return STNTT_OK;
}

int stntt_set_profiling(stntt_ctx *ctx, int on) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    ctx->profile = on != 0;
    ctx->prof_count = 0;
    return STNTT_OK;
}

int stntt_profile_read(stntt_ctx *ctx, float *kernel_ms, int capacity, int *count) {
    if (!ctx || !count) return fail(STNTT_EINVAL, "null argument");
    int k = ctx->prof_count > 0 ? ctx->prof_count - 1 : 0;
    *count = k;
    if (k == 0) return STNTT_OK;
    HIP_TRY(hipEventSynchronize(ctx->prof_ev[k]));
    for (int i = 0; i < k && i < capacity; ++i)
        HIP_TRY(hipEventElapsedTime(&kernel_ms[i], ctx->prof_ev[i], ctx->prof_ev[i + 1]));
    return STNTT_OK;
}

int stntt_timer_start(stntt_ctx *ctx, stntt_stream stream) {
    if (!ctx) return fail(STNTT_EINVAL, "null context");
    HIP_TRY(hipEventRecord(ctx->ev_start, stream ? (hipStream_t)stream : ctx->stream));
    return STNTT_OK;
}

int stntt_timer_stop(stntt_ctx *ctx, stntt_stream stream, float *ms) {
    if (!ctx || !ms) return fail(STNTT_EINVAL, "null argument");
    HIP_TRY(hipEventRecord(ctx->ev_stop, stream ? (hipStream_t)stream : ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev_stop));
    HIP_TRY(hipEventElapsedTime(ms, ctx->ev_start, ctx->ev_stop));
    return STNTT_OK;
}

}  // extern "C"
