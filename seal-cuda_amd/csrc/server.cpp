// server.cpp -- the continuous-batching request server (server.h).
#include "server.h"
#include "behz.h"
#include "elementwise.h"
#include "galois.h"
#include "ntt.h"
#include "rns_ops.h"
#include "rotations.h"
#include <algorithm>
#include <cstring>

namespace shc {

namespace {
constexpr size_t GRAIN = 256;
size_t round_up(size_t x) { return (x + GRAIN - 1) / GRAIN * GRAIN; }
}  // namespace

// ------------------------------------------------------------------ programs

namespace {

size_t ct_words(const ServerProgram &p) { return 2 * (size_t)p.L * p.ctx->n; }

size_t matvec_ws(const ServerProgram &p, uint32_t m) {
    return std::max(rotate_sum_workspace(*p.ctx, m, p.L), galois_keyswitch_workspace(*p.ctx, m, p.L));
}

struct MulRelinLayout {
    size_t prod, ks, mul, tmp, out, rs, total;
};

MulRelinLayout mulrelin_layout(const ServerProgram &p, uint32_t m) {
    const Context &c = *p.ctx;
    const size_t limb = c.n * 8, L = p.L;
    MulRelinLayout l{};
    size_t off = 0;
    auto put = [&](size_t bytes) { size_t o = off; off += round_up(bytes); return o; };
    l.prod = put(m * 3 * L * limb);
    l.ks = put(keyswitch_workspace(c, m, p.L));
    l.mul = put(c.scheme == SCHEME_BFV ? bfv_multiply_workspace(c, m, p.L) : 0);
    if (p.rescale) {
        l.tmp = put(m * 2 * L * limb);
        l.out = put(m * 2 * (L - 1) * limb);
        l.rs = put(rescale_workspace(c, m, 2, p.L));
    }
    l.total = off;
    return l;
}

}  // namespace

size_t ServerProgram::in_words() const { return kind == MULRELIN ? 2 * ct_words(*this) : ct_words(*this); }

size_t ServerProgram::out_words() const {
    return kind == MULRELIN && rescale ? 2 * (size_t)(L - 1) * ctx->n : ct_words(*this);
}

size_t ServerProgram::footprint(uint32_t m) const {
    const size_t in = round_up(m * in_words() * 8);
    if (kind == MULRELIN) return in + mulrelin_layout(*this, m).total;
    const size_t ct = m * ct_words(*this) * 8;
    return in + round_up(ct) + round_up(b * ct) + round_up(g * ct) + round_up(matvec_ws(*this, m));
}

// Runs one batch; returns the first result and sets *stride (words between results).
static const uint64_t *run_program(const ServerProgram &p, uint64_t *d_in, char *scratch, uint32_t m,
                                   hipStream_t st, size_t *stride) {
    const Context &c = *p.ctx;
    const size_t cw = ct_words(p), n = c.n;
    const uint32_t L = p.L;
    if (p.kind == ServerProgram::MULRELIN) {
        const MulRelinLayout l = mulrelin_layout(p, m);
        uint64_t *prod = reinterpret_cast<uint64_t *>(scratch + l.prod);
        const uint64_t *A = d_in, *B = d_in + (size_t)m * cw;
        if (c.scheme == SCHEME_BFV)
            bfv_multiply(c, A, B, prod, m, L, scratch + l.mul, bfv_multiply_workspace(c, m, L), st);
        else
            ckks_tensor(c, A, B, prod, m, L, st);
        const size_t poly = (size_t)L * n;
        switch_key(c, prod + 2 * poly, 3 * poly, prod, 3 * poly, p.relin_key, m, L, scratch + l.ks,
                   keyswitch_workspace(c, m, L), st);
        if (!p.rescale) {
            *stride = 3 * poly;
            return prod;
        }
        uint64_t *tmp = reinterpret_cast<uint64_t *>(scratch + l.tmp);
        uint64_t *out = reinterpret_cast<uint64_t *>(scratch + l.out);
        SHC_HIP(hipMemcpy2DAsync(tmp, 2 * poly * 8, prod, 3 * poly * 8, 2 * poly * 8, m, hipMemcpyDeviceToDevice, st));
        rescale_to_next(c, tmp, out, m, 2, L, scratch + l.rs, rescale_workspace(c, m, 2, L), st);
        *stride = 2 * (size_t)(L - 1) * n;
        return out;
    }
    // diagonal BSGS product: M v = sum_h rot_{hb}( sum_j diag'_{h,j} rot_j(v) )
    const size_t ct = (size_t)m * cw;
    char *q = scratch;
    uint64_t *out = reinterpret_cast<uint64_t *>(q);
    q += round_up(ct * 8);
    uint64_t *baby = reinterpret_cast<uint64_t *>(q);
    q += round_up(p.b * ct * 8);
    uint64_t *inner = reinterpret_cast<uint64_t *>(q);
    q += round_up(p.g * ct * 8);
    const size_t ws_bytes = matvec_ws(p, m);
    const bool bfv = c.scheme == SCHEME_BFV;
    SHC_HIP(hipMemcpyAsync(baby, d_in, ct * 8, hipMemcpyDeviceToDevice, st));
    if (p.b > 1)
        rotate_many(c, d_in, m, L, p.baby_elts.data(), p.baby_hoisted.data(), p.baby_keys.data(), p.b - 1, baby + ct,
                    q, ws_bytes, st);
    if (bfv) ntt_forward(baby, (size_t)p.b * m * 2 * L, L, c.primes(), c.logn, NTT_FULL, st);
    multiply_plain_accumulate_many(c, baby, p.plains, p.b, p.g, m * 2, L, inner, st);
    if (bfv) ntt_inverse(inner, (size_t)p.g * m * 2 * L, L, c.primes(), c.logn, NTT_FULL, st);
    for (uint32_t h = 1; h < p.g; h++)
        apply_galois_keyswitch(c, inner + h * ct, m, L, p.giant_elts[h - 1], p.giant_keys[h - 1], q, ws_bytes, st);
    if (p.g == 1) {
        SHC_HIP(hipMemcpyAsync(out, inner, ct * 8, hipMemcpyDeviceToDevice, st));
    } else {
        poly_add(c, inner, inner + ct, out, m * 2, L, st);
        for (uint32_t h = 2; h < p.g; h++) poly_add(c, out, inner + h * ct, out, m * 2, L, st);
    }
    *stride = cw;
    return out;
}

// ------------------------------------------------------------------ Server

Server::Server(int device, size_t device_bytes, size_t staging_bytes, uint32_t lanes, uint32_t max_batch)
    : device_(device), max_batch_(max_batch), t0_(std::chrono::steady_clock::now()) {
    if (lanes < 1 || lanes > 4) fail(ERR_INVALID_ARGUMENT, "lanes must be in [1, 4]");
    if (max_batch < 1) fail(ERR_INVALID_ARGUMENT, "max_batch must be at least 1");
    if (device_bytes < GRAIN || staging_bytes < GRAIN) fail(ERR_INVALID_ARGUMENT, "arena sizes too small");
    SHC_HIP(hipSetDevice(device));
    try {
        SHC_HIP(hipMalloc((void **)&d_arena_, device_bytes));
        SHC_HIP(hipHostMalloc((void **)&h_staging_, staging_bytes, hipHostMallocDefault));
        for (uint32_t i = 0; i < lanes; i++) {
            hipStream_t s;
            SHC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            streams_.push_back(s);
        }
    } catch (...) {
        for (auto s : streams_) (void)hipStreamDestroy(s);
        if (h_staging_) (void)hipHostFree(h_staging_);
        if (d_arena_) (void)hipFree(d_arena_);
        throw;
    }
    dev_alloc_.reset(device_bytes);
    staging_alloc_.reset(staging_bytes);
    for (uint32_t i = 0; i < lanes; i++) threads_.emplace_back(&Server::lane_main, this, i);
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> g(mu_);
        stopping_ = true;
    }
    work_cv_.notify_all();
    staging_cv_.notify_all();
    for (auto &t : threads_) t.join();
    for (auto s : streams_) (void)hipStreamDestroy(s);
    for (auto &b : host_blocks_) (void)hipHostFree(const_cast<char *>(b.first));
    (void)hipHostFree(h_staging_);
    (void)hipFree(d_arena_);
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

uint32_t Server::add_program(ServerProgram p) {
    if (!p.ctx) fail(ERR_INVALID_ARGUMENT, "program has no context");
    const Context &c = *p.ctx;
    if (p.L < 1 || p.L > c.key_count() - 1 || c.key_count() < 2)
        fail(ERR_INVALID_ARGUMENT, "limb_count out of range for key switching");
    if (c.scheme != SCHEME_BFV && c.scheme != SCHEME_CKKS) fail(ERR_INVALID_ARGUMENT, "unsupported scheme");
    if (p.kind == ServerProgram::MULRELIN) {
        if (!p.relin_key) fail(ERR_INVALID_ARGUMENT, "null relinearization key");
        if (p.rescale && p.L < 2) fail(ERR_INVALID_ARGUMENT, "cannot rescale a ciphertext of one limb");
    } else {
        if (p.b < 1 || p.g < 1 || !p.plains) fail(ERR_INVALID_ARGUMENT, "matvec needs b, g >= 1 and diagonals");
        if (p.baby_keys.size() != p.b - 1 || p.baby_hoisted.size() != p.b - 1 || p.giant_keys.size() != p.g - 1)
            fail(ERR_INVALID_ARGUMENT, "matvec needs b - 1 baby keys (and prepared forms) and g - 1 giant keys");
        for (auto k : p.baby_keys) if (!k) fail(ERR_INVALID_ARGUMENT, "null key pointer");
        for (auto k : p.baby_hoisted) if (!k) fail(ERR_INVALID_ARGUMENT, "null key pointer");
        for (auto k : p.giant_keys) if (!k) fail(ERR_INVALID_ARGUMENT, "null key pointer");
        p.baby_elts.clear();
        p.giant_elts.clear();
        for (uint32_t j = 1; j < p.b; j++) p.baby_elts.push_back(hm::galois_elt_from_step(c.n, (int)j));
        for (uint32_t h = 1; h < p.g; h++) p.giant_elts.push_back(hm::galois_elt_from_step(c.n, (int)(h * p.b)));
    }
    if (p.footprint(1) > dev_alloc_.capacity())
        fail(ERR_OUT_OF_RANGE, "one request of this type does not fit the server's device arena");
    if ((p.in_words() + p.out_words()) * 8 > staging_alloc_.capacity())
        fail(ERR_OUT_OF_RANGE, "one request of this type does not fit the server's staging arena");
    std::lock_guard<std::mutex> g(mu_);
    programs_.push_back(std::move(p));
    return (uint32_t)programs_.size() - 1;
}

const ServerProgram &Server::program(uint32_t type) const {
    std::lock_guard<std::mutex> g(mu_);
    if (type >= programs_.size()) fail(ERR_INVALID_ARGUMENT, "no such request type");
    return programs_[type];
}

void *Server::host_alloc(size_t bytes) {
    if (!bytes) fail(ERR_INVALID_ARGUMENT, "host_alloc of zero bytes");
    void *p = nullptr;
    SHC_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    std::lock_guard<std::mutex> g(mu_);
    host_blocks_[static_cast<const char *>(p)] = bytes;
    return p;
}

void Server::host_free(void *p) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = host_blocks_.find(static_cast<const char *>(p));
    if (it == host_blocks_.end()) fail(ERR_INVALID_ARGUMENT, "not a block from this server's host_alloc");
    SHC_HIP(hipHostFree(p));
    host_blocks_.erase(it);
}

bool Server::pinned_range(const void *p, size_t bytes) const {
    const char *c = static_cast<const char *>(p);
    auto it = host_blocks_.upper_bound(c);
    if (it == host_blocks_.begin()) return false;
    --it;
    return c >= it->first && c + bytes <= it->first + it->second;
}

uint64_t Server::submit(uint32_t type, const uint64_t *h_in, uint64_t *h_out) {
    if (!h_in || !h_out) fail(ERR_INVALID_ARGUMENT, "null request buffer");
    auto r = std::make_unique<Request>();
    r->type = type;
    r->h_out = h_out;
    std::unique_lock<std::mutex> lk(mu_);
    r->info.submit_s = now();
    if (type >= programs_.size()) fail(ERR_INVALID_ARGUMENT, "no such request type");
    const ServerProgram &p = programs_[type];
    const size_t in_b = p.in_words() * 8, out_b = p.out_words() * 8;
    const bool in_direct = pinned_range(h_in, in_b), out_direct = pinned_range(h_out, out_b);
    const size_t bytes = (in_direct ? 0 : in_b) + (out_direct ? 0 : out_b);
    if (bytes) {
        size_t off;
        while ((off = staging_alloc_.take(bytes)) == SIZE_MAX) {
            if (stopping_) fail(ERR_LOGIC, "server is stopping");
            staging_cv_.wait(lk);
        }
        r->staging_off = off;
    }
    char *stage = r->staging_off == SIZE_MAX ? nullptr : h_staging_ + r->staging_off;
    r->src = in_direct ? reinterpret_cast<const char *>(h_in) : stage;
    r->dst = out_direct ? reinterpret_cast<char *>(h_out) : stage + (in_direct ? 0 : in_b);
    if (!in_direct) {
        lk.unlock();
        std::memcpy(stage, h_in, in_b);
        lk.lock();
    }
    r->id = next_id_++;
    Request *raw = r.get();
    requests_[raw->id] = std::move(r);
    pending_.push_back(raw);
    stats_.submitted++;
    lk.unlock();
    work_cv_.notify_one();
    return raw->id;
}

bool Server::wait(uint64_t id, double timeout_s, RequestInfo *info) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = requests_.find(id);
    if (it == requests_.end()) fail(ERR_INVALID_ARGUMENT, "no such request (unknown or already waited for)");
    Request *r = it->second.get();
    auto ready = [&] { return r->done; };
    if (timeout_s < 0) {
        done_cv_.wait(lk, ready);
    } else if (!done_cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
        return false;
    }
    if (info) *info = r->info;
    const std::string err = r->error;
    requests_.erase(it);
    if (!err.empty()) fail(ERR_RUNTIME, "request failed: " + err);
    return true;
}

ServerStats Server::stats() const {
    std::lock_guard<std::mutex> g(mu_);
    return stats_;
}

void Server::lane_main(uint32_t lane) {
    (void)hipSetDevice(device_);
    const hipStream_t st = streams_[lane];
    std::unique_lock<std::mutex> lk(mu_);
    while (true) {
        // admission: the oldest waiting request fixes the type; take every waiting
        // request of that type (up to max_batch) whose batch still fits the device
        // memory free now; wait for memory when not even one fits
        std::vector<Request *> batch;
        size_t dev_off = SIZE_MAX;
        while (!stopping_) {
            if (!pending_.empty()) {
                const uint32_t type = pending_.front()->type;
                const ServerProgram &p = programs_[type];
                uint32_t waiting = 0;
                for (Request *r : pending_) waiting += r->type == type;
                uint32_t m = std::min(waiting, max_batch_);
                const size_t room = dev_alloc_.largest_free();
                while (m > 0 && p.footprint(m) > room) m--;
                if (m > 0) {
                    dev_off = dev_alloc_.take(p.footprint(m));
                    for (auto it = pending_.begin(); it != pending_.end() && batch.size() < m;) {
                        if ((*it)->type == type) {
                            batch.push_back(*it);
                            it = pending_.erase(it);
                        } else {
                            ++it;
                        }
                    }
                    break;
                }
            }
            work_cv_.wait(lk);
        }
        if (batch.empty()) return;  // stopping
        const uint32_t type = batch.front()->type;
        const ServerProgram &p = programs_[type];
        const uint64_t batch_id = next_batch_++;
        const double t_admit = now();
        for (Request *r : batch) {
            r->info.admit_s = t_admit;
            r->info.batch_size = (uint32_t)batch.size();
            r->info.lane = lane;
            r->info.batch_id = batch_id;
        }
        in_flight_++;
        stats_.batches++;
        stats_.max_batches_in_flight = std::max(stats_.max_batches_in_flight, in_flight_);
        stats_.peak_device_bytes = std::max(stats_.peak_device_bytes, dev_alloc_.in_use());
        // more requests of another type may still be admissible by another lane
        if (!pending_.empty()) work_cv_.notify_one();
        lk.unlock();
        std::string error;
        try {
            run_batch(lane, st, p, batch, dev_off);
        } catch (const std::exception &e) {
            error = e.what();
            (void)hipStreamSynchronize(st);
        }
        // results leave staging for the callers' buffers outside the lock
        if (error.empty())
            for (Request *r : batch)
                if (r->dst != reinterpret_cast<char *>(r->h_out)) std::memcpy(r->h_out, r->dst, p.out_words() * 8);
        lk.lock();
        const double t_done = now();
        dev_alloc_.give(dev_off);
        for (Request *r : batch) {
            if (r->staging_off != SIZE_MAX) staging_alloc_.give(r->staging_off);
            r->info.done_s = t_done;
            r->error = error;
            r->done = true;
            if (error.empty()) stats_.completed++;
            else stats_.failed++;
        }
        in_flight_--;
        done_cv_.notify_all();
        staging_cv_.notify_all();
        work_cv_.notify_all();  // memory freed: a waiting lane may admit now
    }
}

void Server::run_batch(uint32_t, hipStream_t st, const ServerProgram &p, std::vector<Request *> &batch,
                       size_t dev_off) {
    const uint32_t m = (uint32_t)batch.size();
    const size_t in_w = p.in_words(), out_w = p.out_words();
    uint64_t *d_in = reinterpret_cast<uint64_t *>(d_arena_ + dev_off);
    char *scratch = d_arena_ + dev_off + round_up(m * in_w * 8);
    // upload: a MULRELIN request is (a, b); the batch holds all a's, then all b's
    const uint32_t parts = p.kind == ServerProgram::MULRELIN ? 2 : 1;
    const size_t part_w = in_w / parts;
    for (uint32_t k = 0; k < m; k++) {
        const char *src = batch[k]->src;
        for (uint32_t part = 0; part < parts; part++)
            SHC_HIP(hipMemcpyAsync(d_in + (size_t)part * m * part_w + k * part_w, src + part * part_w * 8,
                                   part_w * 8, hipMemcpyHostToDevice, st));
    }
    size_t stride = 0;
    const uint64_t *res = run_program(p, d_in, scratch, m, st, &stride);
    for (uint32_t k = 0; k < m; k++)
        SHC_HIP(hipMemcpyAsync(batch[k]->dst, res + k * stride, out_w * 8, hipMemcpyDeviceToHost, st));
    SHC_HIP(hipStreamSynchronize(st));
}

}  // namespace shc
