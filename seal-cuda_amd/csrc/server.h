// server.h -- continuous-batching request server over the batched operations.
//
// The reference's serving-shaped caller is app/matmul.cuh:98-121: one request at a
// time, every operation on the default stream.  Here a Server owns one device arena
// (allocated once), a pinned host staging arena (allocated once) and `lanes` HIP
// streams (default priority), each driven by a host thread.  Requests are submitted
// from any thread with host input and output buffers; a lane that is free takes the
// type of the oldest waiting request and admits as many waiting requests of that type
// as fit its batch limit and the device memory free at that moment, runs them as one
// batch on its stream (upload, batched operations, download) and completes them.
// Batches of different lanes overlap on the device; requests that arrive while others
// run are admitted as soon as a lane and memory free up.
#pragma once
#include "context.h"
#include "range_alloc.h"
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace shc {

// One kind of request the server runs in batches.
struct ServerProgram {
    enum Kind { MATVEC = 0, MULRELIN = 1 };
    Kind kind = MATVEC;
    const Context *ctx = nullptr;
    uint32_t L = 0;
    // MATVEC: diagonal BSGS product (baby steps b, giant steps g)
    uint32_t b = 1, g = 1;
    const uint64_t *plains = nullptr;            // [g][b][L][N], NTT form
    std::vector<const uint64_t *> baby_keys;     // b - 1 Galois keys (steps 1..b-1)
    std::vector<const uint64_t *> baby_hoisted;  // the same, prepared for hoisting at level L
    std::vector<const uint64_t *> giant_keys;    // g - 1 Galois keys (steps b, 2b, ...)
    std::vector<uint32_t> baby_elts, giant_elts;
    // MULRELIN: multiply two ciphertexts, relinearize, optionally rescale
    const uint64_t *relin_key = nullptr;
    bool rescale = false;

    size_t in_words() const;   // per request (host input)
    size_t out_words() const;  // per request (host output)
    // device bytes of a batch of m requests: the packed inputs, then the scratch the
    // batched operations use (results included)
    size_t footprint(uint32_t m) const;
};

struct RequestInfo {
    double submit_s = 0, admit_s = 0, done_s = 0;  // seconds on the server's clock
    uint32_t batch_size = 0, lane = 0;
    uint64_t batch_id = 0;
};

struct ServerStats {
    uint64_t submitted = 0, completed = 0, failed = 0, batches = 0;
    uint32_t max_batches_in_flight = 0;  // largest number of batches running at once
    size_t peak_device_bytes = 0;        // largest device arena use
};

class Server {
public:
    Server(int device, size_t device_bytes, size_t staging_bytes, uint32_t lanes, uint32_t max_batch);
    ~Server();
    uint32_t add_program(ServerProgram p);
    const ServerProgram &program(uint32_t type) const;
    // Queues a request.  An input or output buffer inside a block from host_alloc is
    // used in place (uploaded from / downloaded into directly; the input must stay
    // unchanged until the request is done); other buffers go through the staging arena
    // (the input copied before submit returns; blocks while staging is full).
    uint64_t submit(uint32_t type, const uint64_t *h_in, uint64_t *h_out);
    // Pinned host memory the server uploads from and downloads into without staging.
    void *host_alloc(size_t bytes);
    void host_free(void *p);
    // true when done (info filled), false on timeout; throws the request's error.
    bool wait(uint64_t id, double timeout_s, RequestInfo *info);
    double now() const;
    ServerStats stats() const;

private:
    struct Request {
        uint64_t id;
        uint32_t type;
        uint64_t *h_out;
        const char *src = nullptr;  // pinned input (a host_alloc block or staging)
        char *dst = nullptr;        // pinned output (a host_alloc block or staging)
        size_t staging_off = SIZE_MAX;  // staging block, if any
        RequestInfo info;
        bool done = false;
        std::string error;
    };
    bool pinned_range(const void *p, size_t bytes) const;  // inside one host_alloc block
    void lane_main(uint32_t lane);
    void run_batch(uint32_t lane, hipStream_t st, const ServerProgram &p, std::vector<Request *> &batch,
                   size_t dev_off);

    int device_;
    uint32_t max_batch_;
    char *d_arena_ = nullptr;
    char *h_staging_ = nullptr;
    RangeAllocator dev_alloc_, staging_alloc_;
    std::deque<ServerProgram> programs_;  // references stay valid as types are added
    std::vector<std::thread> threads_;
    std::vector<hipStream_t> streams_;
    mutable std::mutex mu_;
    std::condition_variable work_cv_, done_cv_, staging_cv_;
    std::deque<Request *> pending_;
    std::map<uint64_t, std::unique_ptr<Request>> requests_;
    std::map<const char *, size_t> host_blocks_;  // host_alloc blocks: address -> bytes
    uint64_t next_id_ = 1, next_batch_ = 1;
    uint32_t in_flight_ = 0;
    ServerStats stats_;
    bool stopping_ = false;
    std::chrono::steady_clock::time_point t0_;
};

}  // namespace shc
