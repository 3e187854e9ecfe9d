// server.hpp -- continuous-batching transform server (host side).
//
// No counterpart in the reference (its drivers call autoNtt synchronously,
// one transform at a time: src/bench.cu:26-55, src/main.cu:37-60).
//
// A server owns W worker contexts (one HIP stream, scratch pool and resident
// twiddle plans each), one upload and one download stream, and a
// device-buffer cache in power-of-two size classes.  Clients submit requests
// (field, direction, log2 n, pinned host in/out) from any thread; a server
// thread
//   1. admits pending requests in arrival order while the device bytes held
//      by admitted requests stay within the HBM budget;
//   2. launches each admission round: its uploads go on the upload stream in
//      admission order (one transfer at a time at full PCIe bandwidth, not
//      shared among workers); the least busy worker with fewer than
//      SERVER_WORKER_ROUNDS rounds in flight takes the round: its stream
//      waits for the uploads and runs
//      the round's small transforms (log2 n <= SERVER_BATCH_MAX_LOG) as one
//      persistent descriptor-queue launch, then the larger ones through the
//      multi-pass path, requests of one shape (field, size, direction,
//      parameters) together as one grouped run over pointer tables, as many
//      at a time as the worker's ping-pong scratch holds; the download
//      stream takes each group's results back as soon as its kernels are
//      done.  Uploads, kernels and downloads of different rounds overlap;
//   3. retires each group of a round (the batched small requests, or one
//      grouped run of large requests) when its download event has fired,
//      moves its buffers to the cache and wakes waiters.
// The budget covers every device byte held for requests: admitted buffers
// plus cached idle ones.  A request of a size class with no cached buffer
// first releases cached buffers of other classes until its allocation fits;
// one the device cannot hold with nothing else in flight fails with
// STNTT_ENOMEM rather than waiting for memory that will never free up.
// Requests arriving while others run are admitted as soon as budget frees
// up: continuous batching, not a queue drained one request at a time.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/stntt.h"
#include "ntt_batch.hpp"

namespace stntt {

constexpr uint32_t SERVER_BATCH_MAX_LOG = 16;   // largest size sent through the batch kernel
constexpr uint32_t SERVER_BATCH_GRID = 256;     // workgroups of one round's batch launch
constexpr uint32_t SERVER_MAX_ROUND = 256;      // requests per admission round
// rounds a worker holds in flight: the next round's uploads and launches are
// enqueued while the previous one still computes or downloads (each round
// in flight has its own host staging slot; device-side state is ordered by
// the worker's stream)
constexpr int SERVER_WORKER_ROUNDS = 3;

using Clock = std::chrono::steady_clock;

struct ServerRequest {
    uint64_t id;
    stntt_request req;
    Clock::time_point t_submit;
    double latency_ms = 0;
    int status = 0;
    bool done = false;
    void *d_buf = nullptr;
    size_t d_bytes = 0;  // size-class bytes of d_buf
    uint64_t cost = 0;   // bytes charged against the budget
};

struct ServerWorker {
    stntt_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    // pinned staging of a round's descriptors and grouped buffer pointers,
    // SERVER_WORKER_ROUNDS slots of SERVER_MAX_ROUND entries: the host
    // writes a slot before its copies are enqueued, so a slot stays with its
    // round until the round retires
    BatchTask *h_tasks = nullptr;
    void **h_tab = nullptr;
    bool slot_busy[SERVER_WORKER_ROUNDS] = {};
    BatchTask *d_tasks = nullptr;
    uint32_t *d_counter = nullptr;
    uint64_t *d_scratch = nullptr;  // SERVER_BATCH_GRID slices of 2^SERVER_BATCH_MAX_LOG words
    void **d_tab = nullptr;         // device copy of a round's pointer tables (in and out tables of grouped runs)
    int inflight = 0;
};

// requests of a round that complete together: their downloads end at ev
struct ServerGroup {
    hipEvent_t ev = nullptr;
    std::vector<uint64_t> ids;
    bool done = false;
};

struct ServerRound {
    int worker;
    int slot;  // the worker's staging slot
    std::vector<ServerGroup> groups;
    std::vector<hipEvent_t> aux;  // ordering events, recycled when the round retires
    size_t pending = 0;           // groups not yet retired
};

}  // namespace stntt

struct stntt_server {
    int device = 0;
    hipStream_t up = nullptr, down = nullptr;  // host->device and device->host copies
    uint64_t budget = 0;
    uint64_t used = 0;
    std::vector<stntt::ServerWorker> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;  // arrival order
    std::unordered_map<uint64_t, stntt::ServerRequest> reqs;
    std::deque<stntt::ServerRound> rounds;
    std::map<size_t, std::vector<void *>> cache;  // size class -> free device buffers
    std::vector<hipEvent_t> free_events;
    std::vector<double> latencies;
    uint64_t next_id = 1;
    uint64_t completed = 0;
    uint64_t cached = 0;     // bytes of idle buffers in `cache` (held, charged to the budget)
    uint64_t peak_used = 0;  // peak of used + cached
    uint64_t rounds_launched = 0, batched_requests = 0, grouped_requests = 0;
    bool stop = false;
    std::thread thread;
    stntt::Clock::time_point t_first{}, t_last{};
    bool any = false;
};
