// capi.cpp -- extern "C" entry points declared in include/sealhip.h.
#include "../../include/sealhip.h"
#include "autotune.h"
#include "context.h"
#include "ntt.h"
#include "ckks_encoder.h"
#include "crypto.h"
#include "elementwise.h"
#include "galois.h"
#include "rns_ops.h"
#include "rotations.h"
#include "scratch.h"
#include "server.h"
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

struct shc_context {
    std::unique_ptr<shc::Context> impl;
};

struct shc_server {
    std::unique_ptr<shc::Server> impl;
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F &&f) {
    try {
        f();
        g_last_error.clear();
        return SHC_OK;
    } catch (const shc::Error &e) {
        g_last_error = e.what();
        return e.kind == shc::ERR_HIP ? SHC_HIP_ERROR : (int)e.kind;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return SHC_RUNTIME_ERROR;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return SHC_RUNTIME_ERROR;
    }
}

const shc::Context &ctx_of(const shc_context *c) {
    if (!c || !c->impl) shc::fail(shc::ERR_INVALID_ARGUMENT, "context is null");
    return *c->impl;
}

void check_limbs(const shc::Context &c, uint32_t limbs) {
    if (limbs == 0 || limbs > c.key_count())
        shc::fail(shc::ERR_INVALID_ARGUMENT, "limb_count must be in [1, coeff_modulus_size]");
}
}  // namespace

extern "C" {

const char *shc_last_error(void) { return g_last_error.c_str(); }

int shc_init(int device) {
    return guarded([&] {
        int count = 0;
        SHC_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) shc::fail(shc::ERR_INVALID_ARGUMENT, "no such device");
        SHC_HIP(hipSetDevice(device));
        SHC_HIP(hipFree(nullptr));
        shc::noalloc_scratch_setup();  // this device's address of the fixed scratch array
    });
}

int shc_context_create(int scheme, uint32_t n, const uint64_t *moduli, uint32_t count, uint64_t plain,
                       shc_context **out) {
    return guarded([&] {
        if (!out || !moduli) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        if (n < 2 || (n & (n - 1))) shc::fail(shc::ERR_INVALID_ARGUMENT, "poly_modulus_degree must be a power of two");
        int logn = 0;
        while ((1u << logn) < n) logn++;
        auto impl = shc::make_context((shc::Scheme)scheme, logn, std::vector<uint64_t>(moduli, moduli + count), plain);
        *out = new shc_context{std::move(impl)};
    });
}

void shc_context_destroy(shc_context *ctx) { delete ctx; }

int shc_coeff_modulus_create(uint32_t n, const int *bits, uint32_t count, uint64_t plain, uint64_t *out) {
    return guarded([&] {
        auto v = shc::hm::coeff_modulus_create(n, std::vector<int>(bits, bits + count), plain);
        for (size_t i = 0; i < v.size(); i++) out[i] = v[i];
    });
}

int shc_ntt_forward(const shc_context *ctx, uint64_t *d, uint32_t polys, uint32_t limbs, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::ntt_forward(d, polys * limbs, limbs, c.primes(), c.logn, shc::NTT_FULL, (hipStream_t)stream);
    });
}

int shc_ntt_inverse(const shc_context *ctx, uint64_t *d, uint32_t polys, uint32_t limbs, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::ntt_inverse(d, polys * limbs, limbs, c.primes(), c.logn, shc::NTT_FULL, (hipStream_t)stream);
    });
}

namespace {
int ntt_multi(const shc_ntt_job *jobs, uint32_t n, void *stream, bool inverse) {
    return guarded([&] {
        if (n > shc::MAX_GROUPS) shc::fail(shc::ERR_INVALID_ARGUMENT, "at most 16 transform jobs per call");
        if (n && !jobs) shc::fail(shc::ERR_INVALID_ARGUMENT, "null job list");
        std::vector<shc::NttGroup> gs;
        std::vector<int> logns;
        for (uint32_t k = 0; k < n; k++) {
            const auto &c = ctx_of(jobs[k].ctx);
            const uint32_t limbs = jobs[k].limb_count, polys = jobs[k].poly_count;
            check_limbs(c, limbs);
            if (!polys) continue;
            if (!jobs[k].d_data) shc::fail(shc::ERR_INVALID_ARGUMENT, "null transform data");
            if (c.logn < 13) {  // one-tile transforms: a launch of their own
                if (inverse)
                    shc::ntt_inverse(jobs[k].d_data, polys * limbs, limbs, c.primes(), c.logn, shc::NTT_FULL,
                                     (hipStream_t)stream);
                else
                    shc::ntt_forward(jobs[k].d_data, polys * limbs, limbs, c.primes(), c.logn, shc::NTT_FULL,
                                     (hipStream_t)stream);
                continue;
            }
            const shc::LimbGrid lg{1, limbs, limbs, polys};
            gs.push_back(shc::NttGroup{jobs[k].d_data, nullptr, lg, lg, c.primes()});
            logns.push_back(c.logn);
        }
        if (inverse)
            shc::ntt_inverse_multi(gs.data(), logns.data(), (uint32_t)gs.size(), shc::NTT_FULL, (hipStream_t)stream);
        else
            shc::ntt_forward_multi(gs.data(), logns.data(), (uint32_t)gs.size(), shc::NTT_FULL, (hipStream_t)stream);
    });
}
}  // namespace

int shc_ntt_forward_multi(const shc_ntt_job *jobs, uint32_t n, void *stream) { return ntt_multi(jobs, n, stream, false); }
int shc_ntt_inverse_multi(const shc_ntt_job *jobs, uint32_t n, void *stream) { return ntt_multi(jobs, n, stream, true); }

int shc_ckks_multiply(const shc_context *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out, uint32_t count,
                      uint32_t limbs, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::ckks_tensor(c, a, b, out, count, limbs, (hipStream_t)stream);
    });
}

int shc_multiply_workspace_size(const shc_context *ctx, uint32_t count, uint32_t limbs, size_t *bytes) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        *bytes = c.scheme == shc::SCHEME_BFV ? shc::bfv_multiply_workspace(c, count, limbs) : 0;
    });
}

int shc_multiply(const shc_context *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out, uint32_t count,
                 uint32_t limbs, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        if (c.scheme == shc::SCHEME_BFV)
            shc::bfv_multiply(c, a, b, out, count, limbs, ws, ws_bytes, (hipStream_t)stream);
        else if (c.scheme == shc::SCHEME_CKKS)
            shc::ckks_tensor(c, a, b, out, count, limbs, (hipStream_t)stream);
        else
            shc::fail(shc::ERR_INVALID_ARGUMENT, "unsupported scheme");
    });
}

int shc_keyswitch_workspace_size(const shc_context *ctx, uint32_t count, uint32_t limbs, size_t *bytes) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        *bytes = shc::keyswitch_workspace(c, count, limbs);
    });
}

int shc_switch_key(const shc_context *ctx, const uint64_t *target, size_t target_stride, uint64_t *ct,
                   size_t ct_stride, const uint64_t *key, uint32_t count, uint32_t limbs, void *ws, size_t ws_bytes,
                   void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::switch_key(c, target, target_stride, ct, ct_stride, key, count, limbs, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_relinearize(const shc_context *ctx, uint64_t *ct, const uint64_t *key, uint32_t count, uint32_t limbs,
                    void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        const size_t poly = (size_t)limbs * c.n;
        shc::switch_key(c, ct + 2 * poly, 3 * poly, ct, 3 * poly, key, count, limbs, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_relinearize_noalloc(const shc_context *ctx, uint64_t *ct, const uint64_t *key, uint32_t count,
                            uint32_t limbs, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::relinearize_noalloc(c, ct, key, count, limbs, (hipStream_t)stream);
    });
}

int shc_autotune_relinearize(const shc_context *ctx, uint64_t *ct, const uint64_t *key, uint32_t count,
                             uint32_t limbs, void *ws, size_t ws_bytes, int reps, void *stream, shc_ks_shape *best,
                             float *best_ms, float *builtin_ms) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        const shc::TuneResult r = shc::autotune_relinearize(c, ct, key, count, limbs, ws, ws_bytes,
                                                            (hipStream_t)stream, reps);
        if (best) *best = shc_ks_shape{r.best.cols_targets, r.best.rows_group};
        if (best_ms) *best_ms = r.best_ms;
        if (builtin_ms) *builtin_ms = r.builtin_ms;
    });
}

int shc_keyswitch_shape_set(const shc_context *ctx, uint32_t limbs, uint32_t count, const shc_ks_shape *shape) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        if (!shape) {
            shc::keyswitch_shape_set(c, limbs, count, nullptr);
            return;
        }
        const shc::KsShape s{shape->cols_targets, shape->rows_group};
        shc::keyswitch_shape_set(c, limbs, count, &s);
    });
}

int shc_keyswitch_shape_get(const shc_context *ctx, uint32_t limbs, uint32_t count, shc_ks_shape *shape) {
    return guarded([&] {
        if (!shape) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        const shc::KsShape s = ctx_of(ctx).ks_shape(limbs, count);
        *shape = shc_ks_shape{s.cols_targets, s.rows_group};
    });
}

int shc_keyswitch_shapes_export(const shc_context *ctx, char *buf, size_t capacity, size_t *length) {
    return guarded([&] {
        const std::string t = shc::keyswitch_shapes_export(ctx_of(ctx));
        if (length) *length = t.size() + 1;
        if (buf && capacity) {
            if (capacity < t.size() + 1) shc::fail(shc::ERR_OUT_OF_RANGE, "buffer too small for the table");
            std::memcpy(buf, t.c_str(), t.size() + 1);
        }
    });
}

int shc_keyswitch_shapes_import(const shc_context *ctx, const char *text) {
    return guarded([&] {
        if (!text) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        shc::keyswitch_shapes_import(ctx_of(ctx), text);
    });
}

static void check_groups(const void *p, uint32_t n) {
    if (n && !p) shc::fail(shc::ERR_INVALID_ARGUMENT, "null group array");
}

int shc_ckks_multiply_groups(const shc_context *ctx, const shc_ct_group *a, const shc_ct_group *b,
                             const shc_ct_group *out, uint32_t ngroups, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_groups(a, ngroups);
        check_groups(b, ngroups);
        check_groups(out, ngroups);
        std::vector<shc::TensorGroup> g(ngroups);
        for (uint32_t i = 0; i < ngroups; i++) {
            if (a[i].count != b[i].count || a[i].count != out[i].count || a[i].limb_count != b[i].limb_count ||
                a[i].limb_count != out[i].limb_count)
                shc::fail(shc::ERR_INVALID_ARGUMENT, "groups of a, b and out differ in count or level");
            check_limbs(c, a[i].limb_count);
            g[i] = shc::TensorGroup{a[i].d_ct, b[i].d_ct, out[i].d_ct, a[i].count, a[i].limb_count};
        }
        shc::ckks_tensor_groups(c, g.data(), ngroups, (hipStream_t)stream);
    });
}

static std::vector<shc::RelinGroup> relin_groups(const shc::Context &c, const shc_ct_group *cts, uint32_t n) {
    check_groups(cts, n);
    std::vector<shc::RelinGroup> g(n);
    for (uint32_t i = 0; i < n; i++) {
        check_limbs(c, cts[i].limb_count);
        g[i] = shc::RelinGroup{cts[i].d_ct, cts[i].count, cts[i].limb_count};
    }
    return g;
}

int shc_relinearize_groups_workspace_size(const shc_context *ctx, const shc_ct_group *cts, uint32_t ngroups,
                                          size_t *bytes) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        const auto g = relin_groups(c, cts, ngroups);
        *bytes = shc::relinearize_groups_workspace(c, g.data(), ngroups);
    });
}

int shc_relinearize_groups(const shc_context *ctx, const shc_ct_group *cts, uint32_t ngroups, const uint64_t *key,
                           void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        const auto g = relin_groups(c, cts, ngroups);
        shc::relinearize_groups(c, g.data(), ngroups, key, ws, ws_bytes, (hipStream_t)stream);
    });
}

static std::vector<shc::RescaleGroup> rescale_groups_of(const shc::Context &c, const shc_ct_group *in,
                                                        const shc_ct_group *out, uint32_t n) {
    check_groups(in, n);
    std::vector<shc::RescaleGroup> g(n);
    for (uint32_t i = 0; i < n; i++) {
        check_limbs(c, in[i].limb_count);
        if (out && (out[i].count != in[i].count || out[i].limb_count + 1 != in[i].limb_count))
            shc::fail(shc::ERR_INVALID_ARGUMENT, "an output group must hold the same count one level down");
        g[i] = shc::RescaleGroup{in[i].d_ct, out ? out[i].d_ct : nullptr, in[i].count, in[i].limb_count};
    }
    return g;
}

int shc_rescale_groups_workspace_size(const shc_context *ctx, const shc_ct_group *in, uint32_t ngroups,
                                      size_t *bytes) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        const auto g = rescale_groups_of(c, in, nullptr, ngroups);
        *bytes = shc::rescale_groups_workspace(c, g.data(), ngroups);
    });
}

int shc_rescale_to_next_groups(const shc_context *ctx, const shc_ct_group *in, uint32_t in_size,
                               const shc_ct_group *out, uint32_t ngroups, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_groups(out, ngroups);
        const auto g = rescale_groups_of(c, in, out, ngroups);
        shc::rescale_groups(c, g.data(), ngroups, in_size, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_galois_elt_from_step(uint32_t n, int step, uint32_t *elt) {
    return guarded([&] {
        if (n < 2 || (n & (n - 1))) shc::fail(shc::ERR_INVALID_ARGUMENT, "poly_modulus_degree must be a power of two");
        *elt = shc::hm::galois_elt_from_step(n, step);
    });
}

int shc_apply_galois(const shc_context *ctx, const uint64_t *in, uint64_t *out, uint32_t polys, uint32_t limbs,
                     uint32_t elt, int ntt_form, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        const size_t poly = (size_t)limbs * c.n;
        shc::apply_galois(c, in, poly, out, poly, polys, limbs, elt, ntt_form != 0, (hipStream_t)stream);
    });
}

int shc_galois_workspace_size(const shc_context *ctx, uint32_t count, uint32_t limbs, size_t *bytes) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        *bytes = shc::galois_keyswitch_workspace(c, count, limbs);
    });
}

int shc_apply_galois_keyswitch(const shc_context *ctx, uint64_t *ct, const uint64_t *key, uint32_t elt, uint32_t count,
                               uint32_t limbs, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::apply_galois_keyswitch(c, ct, count, limbs, elt, key, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_context_set_sampler_key(shc_context *ctx, const uint8_t *key) {
    return guarded([&] {
        if (!ctx || !ctx->impl) shc::fail(shc::ERR_INVALID_ARGUMENT, "context is null");
        shc::Context &c = *ctx->impl;
        if (!key) {
            c.sampler_keyed = false;
            std::fill(c.sampler_key, c.sampler_key + 8, 0u);
            return;
        }
        for (int i = 0; i < 8; i++)
            c.sampler_key[i] = (uint32_t)key[4 * i] | ((uint32_t)key[4 * i + 1] << 8) |
                               ((uint32_t)key[4 * i + 2] << 16) | ((uint32_t)key[4 * i + 3] << 24);
        c.sampler_keyed = true;
    });
}

int shc_sample_poly(const shc_context *ctx, int kind, uint64_t seed, uint32_t limbs, uint64_t *out, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        if (kind < 0 || kind > 2) shc::fail(shc::ERR_INVALID_ARGUMENT, "unknown sample kind");
        shc::sample_poly(c, (shc::SampleKind)kind, seed, c.primes(), limbs, out, (hipStream_t)stream);
    });
}

int shc_keygen_secret(const shc_context *ctx, uint64_t seed, uint64_t *sk, void *stream) {
    return guarded([&] { shc::keygen_secret(ctx_of(ctx), seed, sk, (hipStream_t)stream); });
}

int shc_keygen_public(const shc_context *ctx, const uint64_t *sk, uint64_t seed, uint64_t *pk, void *stream) {
    return guarded([&] { shc::keygen_public(ctx_of(ctx), sk, seed, pk, (hipStream_t)stream); });
}

int shc_keygen_workspace_size(const shc_context *ctx, size_t *bytes) {
    return guarded([&] { *bytes = shc::keygen_workspace(ctx_of(ctx)); });
}

int shc_keygen_relin(const shc_context *ctx, const uint64_t *sk, uint64_t seed, uint64_t *key, void *ws,
                     size_t ws_bytes, void *stream) {
    return guarded([&] { shc::keygen_relin(ctx_of(ctx), sk, seed, key, ws, ws_bytes, (hipStream_t)stream); });
}

int shc_keygen_galois(const shc_context *ctx, const uint64_t *sk, uint32_t elt, uint64_t seed, uint64_t *key, void *ws,
                      size_t ws_bytes, void *stream) {
    return guarded([&] { shc::keygen_galois(ctx_of(ctx), sk, elt, seed, key, ws, ws_bytes, (hipStream_t)stream); });
}

int shc_encrypt_workspace_size(const shc_context *ctx, uint32_t limbs, size_t *bytes) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        *bytes = shc::encrypt_workspace(c, limbs);
    });
}

int shc_encrypt_zero_asymmetric(const shc_context *ctx, const uint64_t *pk, const uint64_t *u, const uint64_t *e0,
                                const uint64_t *e1, uint32_t limbs, uint64_t *out, void *ws, size_t ws_bytes,
                                void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::encrypt_zero_asymmetric(c, pk, u, e0, e1, limbs, out, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_encrypt(const shc_context *ctx, const uint64_t *pk, const uint64_t *plain, uint32_t limbs, uint64_t seed,
                uint64_t *out, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::encrypt(c, pk, plain, limbs, seed, out, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_encrypt_many_workspace_size(const shc_context *ctx, uint32_t limbs, uint32_t count, size_t *bytes) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        *bytes = shc::encrypt_many_workspace(c, limbs, count);
    });
}

int shc_encrypt_many(const shc_context *ctx, const uint64_t *pk, const uint64_t *plain, uint32_t count, uint32_t limbs,
                     uint64_t seed, uint64_t *out, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::encrypt_many(c, pk, plain, count, limbs, seed, out, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_encrypt_zero_symmetric(const shc_context *ctx, const uint64_t *sk, uint32_t limbs, uint64_t seed,
                               uint64_t *out, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::encrypt_zero_symmetric(c, sk, c.primes(), limbs, c.scheme == shc::SCHEME_CKKS, seed, out,
                                    (hipStream_t)stream);
    });
}

int shc_decrypt_workspace_size(const shc_context *ctx, uint32_t limbs, size_t *bytes) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        *bytes = shc::decrypt_workspace(c, limbs);
    });
}

int shc_decrypt(const shc_context *ctx, const uint64_t *ct, uint32_t size, const uint64_t *sk, uint32_t limbs,
                uint64_t *out, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::decrypt(c, ct, size, sk, limbs, out, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_decrypt_many_workspace_size(const shc_context *ctx, uint32_t limbs, uint32_t count, size_t *bytes) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        *bytes = shc::decrypt_many_workspace(c, limbs, count);
    });
}

int shc_decrypt_many(const shc_context *ctx, const uint64_t *ct, uint32_t size, uint32_t count, const uint64_t *sk,
                     uint32_t limbs, uint64_t *out, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::decrypt_many(c, ct, size, count, sk, limbs, out, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_batch_encode(const shc_context *ctx, const uint64_t *values, uint32_t count, uint64_t *plain, void *stream) {
    return guarded([&] { shc::batch_encode(ctx_of(ctx), values, count, plain, (hipStream_t)stream); });
}

int shc_batch_decode(const shc_context *ctx, const uint64_t *plain, uint64_t *values, uint64_t *scratch,
                     void *stream) {
    return guarded([&] { shc::batch_decode(ctx_of(ctx), plain, values, scratch, (hipStream_t)stream); });
}

int shc_batch_encode_many(const shc_context *ctx, const uint64_t *values, uint32_t count, size_t values_stride,
                          uint32_t plaintexts, uint64_t *plain, void *stream) {
    return guarded([&] {
        shc::batch_encode_many(ctx_of(ctx), values, count, values_stride, plaintexts, plain, (hipStream_t)stream);
    });
}

int shc_batch_decode_many(const shc_context *ctx, const uint64_t *plain, uint32_t plaintexts, uint64_t *values,
                          uint64_t *scratch, void *stream) {
    return guarded([&] {
        shc::batch_decode_many(ctx_of(ctx), plain, plaintexts, values, scratch, (hipStream_t)stream);
    });
}

int shc_ckks_encode_workspace_size(const shc_context *ctx, size_t *bytes) {
    return guarded([&] { *bytes = shc::ckks_encode_workspace(ctx_of(ctx)); });
}

int shc_ckks_encode(const shc_context *ctx, const double *values, uint32_t count, double scale, uint32_t limbs,
                    uint64_t *plain, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::ckks_encode(c, values, count, scale, limbs, plain, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_ckks_encode_deferred(const shc_context *ctx, const double *values, uint32_t count, double scale,
                             uint32_t limbs, uint64_t *plain, int32_t *d_status, void *ws, size_t ws_bytes,
                             void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::ckks_encode_deferred(c, values, count, scale, limbs, plain, d_status, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_ckks_encode_many_workspace_size(const shc_context *ctx, uint32_t plaintexts, size_t *bytes) {
    return guarded([&] { *bytes = shc::ckks_encode_workspace_many(ctx_of(ctx), plaintexts); });
}

int shc_ckks_encode_many(const shc_context *ctx, const double *values, uint32_t count, size_t values_stride,
                         uint32_t plaintexts, double scale, uint32_t limbs, uint64_t *plain, int32_t *d_status,
                         void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::ckks_encode_many(c, values, count, values_stride, plaintexts, scale, limbs, plain, d_status, ws, ws_bytes,
                              (hipStream_t)stream);
    });
}

int shc_ckks_decode_many_workspace_size(const shc_context *ctx, uint32_t limbs, uint32_t plaintexts, size_t *bytes) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        *bytes = shc::ckks_decode_workspace_many(c, limbs, plaintexts);
    });
}

int shc_ckks_decode_many(const shc_context *ctx, const uint64_t *plain, uint32_t limbs, uint32_t plaintexts,
                         double scale, double *values, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::ckks_decode_many(c, plain, limbs, plaintexts, scale, values, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_ckks_decode_workspace_size(const shc_context *ctx, uint32_t limbs, size_t *bytes) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        *bytes = shc::ckks_decode_workspace(c, limbs);
    });
}

int shc_ckks_decode(const shc_context *ctx, const uint64_t *plain, uint32_t limbs, double scale, double *values,
                    void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::ckks_decode(c, plain, limbs, scale, values, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_ckks_encode_polynomial(const shc_context *ctx, const double *values, uint32_t count, double scale,
                               uint32_t limbs, uint64_t *plain, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::ckks_encode_polynomial(c, values, count, scale, limbs, plain, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_ckks_decode_polynomial(const shc_context *ctx, const uint64_t *plain, uint32_t limbs, double scale,
                               double *values, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::ckks_decode_polynomial(c, plain, limbs, scale, values, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_rescale_workspace_size(const shc_context *ctx, uint32_t count, uint32_t polys, uint32_t limbs, size_t *bytes) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        *bytes = shc::rescale_workspace(c, count, polys, limbs);
    });
}

int shc_rescale_to_next(const shc_context *ctx, const uint64_t *in, uint64_t *out, uint32_t count, uint32_t polys,
                        uint32_t limbs, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        check_limbs(c, limbs);
        shc::rescale_to_next(c, in, out, count, polys, limbs, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_add(const shc_context *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out, uint32_t polys,
            uint32_t limbs, void *stream) {
    return guarded([&] { shc::poly_add(ctx_of(ctx), a, b, out, polys, limbs, (hipStream_t)stream); });
}

int shc_sub(const shc_context *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out, uint32_t polys,
            uint32_t limbs, void *stream) {
    return guarded([&] { shc::poly_sub(ctx_of(ctx), a, b, out, polys, limbs, (hipStream_t)stream); });
}

int shc_negate(const shc_context *ctx, const uint64_t *a, uint64_t *out, uint32_t polys, uint32_t limbs,
               void *stream) {
    return guarded([&] { shc::poly_negate(ctx_of(ctx), a, out, polys, limbs, (hipStream_t)stream); });
}

int shc_multiply_plain_ntt(const shc_context *ctx, const uint64_t *a, const uint64_t *plain, uint64_t *out,
                           uint32_t polys, uint32_t limbs, void *stream) {
    return guarded([&] { shc::multiply_plain_ntt(ctx_of(ctx), a, plain, out, polys, limbs, (hipStream_t)stream); });
}

int shc_add_plain(const shc_context *ctx, const uint64_t *ct, const uint64_t *plain, uint64_t *out, uint32_t count,
                  uint32_t size, uint32_t limbs, int subtract, void *stream) {
    return guarded([&] {
        shc::add_plain_poly0(ctx_of(ctx), ct, plain, out, count, size, limbs, subtract != 0, (hipStream_t)stream);
    });
}

int shc_bfv_add_plain(const shc_context *ctx, const uint64_t *ct, const uint64_t *plain, uint64_t *out,
                      uint32_t count, uint32_t size, uint32_t limbs, int subtract, void *stream) {
    return guarded([&] {
        shc::bfv_add_plain(ctx_of(ctx), ct, plain, out, count, size, limbs, subtract != 0, (hipStream_t)stream);
    });
}

int shc_bfv_multiply_plain_workspace_size(const shc_context *ctx, uint32_t limbs, size_t *bytes) {
    return guarded([&] {
        if (!bytes) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        *bytes = shc::bfv_multiply_plain_workspace(ctx_of(ctx), limbs);
    });
}

int shc_bfv_multiply_plain(const shc_context *ctx, const uint64_t *a, const uint64_t *plain, uint64_t *out,
                           uint32_t polys, uint32_t limbs, void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        shc::bfv_multiply_plain(ctx_of(ctx), a, plain, out, polys, limbs, ws, ws_bytes, (hipStream_t)stream);
    });
}

int shc_bfv_plain_to_ntt(const shc_context *ctx, const uint64_t *plain, uint32_t limbs, uint64_t *out, void *stream) {
    return guarded([&] { shc::bfv_plain_to_ntt(ctx_of(ctx), plain, limbs, out, (hipStream_t)stream); });
}

int shc_hoisted_key_size(const shc_context *ctx, uint32_t limbs, size_t *bytes) {
    return guarded([&] {
        if (!bytes) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        *bytes = shc::hoisted_key_words(ctx_of(ctx), limbs) * 8;
    });
}

int shc_hoisted_key_workspace_size(const shc_context *ctx, uint32_t limbs, size_t *bytes) {
    return guarded([&] {
        if (!bytes) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        *bytes = shc::hoisted_key_workspace(ctx_of(ctx), limbs);
    });
}

int shc_hoisted_key_prepare(const shc_context *ctx, const uint64_t *key, uint32_t elt, uint32_t limbs, uint64_t *out,
                            void *ws, size_t ws_bytes, void *stream) {
    return guarded(
        [&] { shc::hoisted_key_prepare(ctx_of(ctx), key, elt, limbs, out, ws, ws_bytes, (hipStream_t)stream); });
}

int shc_rotate_sum_workspace_size(const shc_context *ctx, uint32_t count, uint32_t limbs, size_t *bytes) {
    return guarded([&] {
        if (!bytes) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        *bytes = shc::rotate_sum_workspace(ctx_of(ctx), count, limbs);
    });
}

int shc_rotate_sum(const shc_context *ctx, const uint64_t *in, uint32_t count, uint32_t limbs, const uint32_t *elts,
                   const uint64_t *const *hoisted, const uint64_t *const *keys, uint32_t rotations, uint64_t *out,
                   void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        shc::rotate_sum(ctx_of(ctx), in, count, limbs, elts, hoisted, keys, rotations, out, ws, ws_bytes,
                        (hipStream_t)stream);
    });
}

int shc_rotate_many(const shc_context *ctx, const uint64_t *in, uint32_t count, uint32_t limbs, const uint32_t *elts,
                    const uint64_t *const *hoisted, const uint64_t *const *keys, uint32_t rotations, uint64_t *out,
                    void *ws, size_t ws_bytes, void *stream) {
    return guarded([&] {
        shc::rotate_many(ctx_of(ctx), in, count, limbs, elts, hoisted, keys, rotations, out, ws, ws_bytes,
                         (hipStream_t)stream);
    });
}

int shc_multiply_plain_accumulate(const shc_context *ctx, const uint64_t *cts, const uint64_t *plains, uint32_t terms,
                                  uint32_t polys, uint32_t limbs, uint64_t *out, void *stream) {
    return guarded([&] {
        shc::multiply_plain_accumulate(ctx_of(ctx), cts, plains, terms, polys, limbs, out, (hipStream_t)stream);
    });
}

int shc_multiply_plain_accumulate_many(const shc_context *ctx, const uint64_t *cts, const uint64_t *plains,
                                       uint32_t terms, uint32_t outputs, uint32_t polys, uint32_t limbs, uint64_t *out,
                                       void *stream) {
    return guarded([&] {
        shc::multiply_plain_accumulate_many(ctx_of(ctx), cts, plains, terms, outputs, polys, limbs, out,
                                            (hipStream_t)stream);
    });
}

int shc_probe_keyswitch(shc_context *ctx, void *begin_event, void *end_event) {
    return guarded([&] {
        const auto &c = ctx_of(ctx);
        if (!begin_event || !end_event) shc::fail(shc::ERR_INVALID_ARGUMENT, "null event");
        c.arm_probe(shc::KernelProbe{(hipEvent_t)begin_event, (hipEvent_t)end_event});
    });
}

int shc_server_create(int device, size_t device_bytes, size_t staging_bytes, uint32_t lanes, uint32_t max_batch,
                      shc_server **out) {
    return guarded([&] {
        if (!out) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        auto impl = std::make_unique<shc::Server>(device, device_bytes, staging_bytes, lanes, max_batch);
        *out = new shc_server{std::move(impl)};
    });
}

void shc_server_destroy(shc_server *server) { delete server; }

static shc::Server &server_of(const shc_server *s) {
    if (!s || !s->impl) shc::fail(shc::ERR_INVALID_ARGUMENT, "server is null");
    return *s->impl;
}

int shc_server_add_matvec(shc_server *server, const shc_context *ctx, uint32_t limbs, uint32_t baby, uint32_t giant,
                          const uint64_t *d_diagonals, const uint64_t *const *d_baby_keys,
                          const uint64_t *const *d_baby_hoisted, const uint64_t *const *d_giant_keys,
                          uint32_t *type) {
    return guarded([&] {
        if (!type) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        if (baby < 1 || giant < 1) shc::fail(shc::ERR_INVALID_ARGUMENT, "baby and giant steps must be >= 1");
        if ((baby > 1 && (!d_baby_keys || !d_baby_hoisted)) || (giant > 1 && !d_giant_keys))
            shc::fail(shc::ERR_INVALID_ARGUMENT, "null key array");
        shc::ServerProgram p;
        p.kind = shc::ServerProgram::MATVEC;
        p.ctx = &ctx_of(ctx);
        p.L = limbs;
        p.b = baby;
        p.g = giant;
        p.plains = d_diagonals;
        if (baby > 1) {
            p.baby_keys.assign(d_baby_keys, d_baby_keys + (baby - 1));
            p.baby_hoisted.assign(d_baby_hoisted, d_baby_hoisted + (baby - 1));
        }
        if (giant > 1) p.giant_keys.assign(d_giant_keys, d_giant_keys + (giant - 1));
        *type = server_of(server).add_program(std::move(p));
    });
}

int shc_server_add_mulrelin(shc_server *server, const shc_context *ctx, uint32_t limbs, const uint64_t *d_relin_key,
                            int rescale, uint32_t *type) {
    return guarded([&] {
        if (!type) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        shc::ServerProgram p;
        p.kind = shc::ServerProgram::MULRELIN;
        p.ctx = &ctx_of(ctx);
        p.L = limbs;
        p.relin_key = d_relin_key;
        p.rescale = rescale != 0;
        *type = server_of(server).add_program(std::move(p));
    });
}

int shc_server_request_size(const shc_server *server, uint32_t type, size_t *in_words, size_t *out_words,
                            size_t *device_bytes) {
    return guarded([&] {
        const shc::ServerProgram &p = server_of(server).program(type);
        if (in_words) *in_words = p.in_words();
        if (out_words) *out_words = p.out_words();
        if (device_bytes) *device_bytes = p.footprint(1);
    });
}

int shc_server_submit(shc_server *server, uint32_t type, const uint64_t *h_in, uint64_t *h_out, uint64_t *id) {
    return guarded([&] {
        if (!id) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        *id = server_of(server).submit(type, h_in, h_out);
    });
}

int shc_server_wait(shc_server *server, uint64_t id, double timeout_s, shc_request_info *info) {
    bool done = false;
    const int rc = guarded([&] {
        shc::RequestInfo ri;
        done = server_of(server).wait(id, timeout_s, &ri);
        if (done && info) {
            info->submit_s = ri.submit_s;
            info->admit_s = ri.admit_s;
            info->done_s = ri.done_s;
            info->batch_size = ri.batch_size;
            info->lane = ri.lane;
            info->batch_id = ri.batch_id;
        }
    });
    return rc == SHC_OK && !done ? SHC_TIMEOUT : rc;
}

int shc_server_host_alloc(shc_server *server, size_t bytes, void **out) {
    return guarded([&] {
        if (!out) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        *out = server_of(server).host_alloc(bytes);
    });
}

int shc_server_host_free(shc_server *server, void *ptr) {
    return guarded([&] { server_of(server).host_free(ptr); });
}

double shc_server_clock(const shc_server *server) { return server && server->impl ? server->impl->now() : 0.0; }

int shc_server_get_stats(const shc_server *server, shc_server_stats *out) {
    return guarded([&] {
        if (!out) shc::fail(shc::ERR_INVALID_ARGUMENT, "null argument");
        const shc::ServerStats s = server_of(server).stats();
        out->submitted = s.submitted;
        out->completed = s.completed;
        out->failed = s.failed;
        out->batches = s.batches;
        out->max_batches_in_flight = s.max_batches_in_flight;
        out->peak_device_bytes = s.peak_device_bytes;
    });
}

}  // extern "C"
