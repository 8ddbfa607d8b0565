// ckks_encoder.h -- CKKS encode / decode on the GPU (fp64).
#pragma once
#include "context.h"

namespace shc {

// Per-level constants of the decode CRT composition.
struct CkksLevelDev {
    uint32_t L;
    uint64_t Q_lo, Q_hi;             // prod(q) mod 2^128
    const uint64_t *qhat_inv;        // [L][2] Shoup pairs: (Q/q_j)^-1 mod q_j
    const uint64_t *qhat128;         // [L][2] (Q/q_j) mod 2^128 (lo, hi)
    const double *inv_q;             // [L]    1 / q_j
    int total_bits;                  // significant bits of prod(q)
};

struct CkksTables {
    DevBuf roots;                    // double2 [N] root_powers, then [N] inv_root_powers
    std::vector<CkksLevelDev> levels;  // index L
    std::vector<DevBuf> bufs;
};

void build_ckks_tables(Context &c);

size_t ckks_encode_workspace(const Context &c);
// values: `count` (re, im) double pairs (device), count <= N/2 -> plain [L][N] NTT form.
void ckks_encode(const Context &c, const double *values, uint32_t count, double scale, uint32_t L, uint64_t *plain,
                 void *ws, size_t ws_bytes, hipStream_t st);
// The same, with the range check written to *d_status on the device (0, or -1 with a
// zero plaintext) instead of raised: nothing waits for the stream.
void ckks_encode_deferred(const Context &c, const double *values, uint32_t count, double scale, uint32_t L,
                          uint64_t *plain, int *d_status, void *ws, size_t ws_bytes, hipStream_t st);
// Batches: `batch` plaintexts in each launch (the DWT in two LDS-tiled passes).  Encode: values of plaintext p
// at values + p * vstride (count pairs each), plain [batch][L][N]; d_status (optional, one
// int per plaintext) takes the range checks on the device.  Decode: plain [batch][L][N]
// -> values [batch][N/2] pairs.
size_t ckks_encode_workspace_many(const Context &c, uint32_t batch);
void ckks_encode_many(const Context &c, const double *values, uint32_t count, size_t vstride, uint32_t batch,
                      double scale, uint32_t L, uint64_t *plain, int *d_status, void *ws, size_t ws_bytes,
                      hipStream_t st);
size_t ckks_decode_workspace_many(const Context &c, uint32_t L, uint32_t batch);
void ckks_decode_many(const Context &c, const uint64_t *plain, uint32_t L, uint32_t batch, double scale,
                      double *values, void *ws, size_t ws_bytes, hipStream_t st);
size_t ckks_decode_workspace(const Context &c, uint32_t L);
// plain [L][N] NTT form -> N/2 (re, im) double pairs.
void ckks_decode(const Context &c, const uint64_t *plain, uint32_t L, double scale, double *values, void *ws,
                 size_t ws_bytes, hipStream_t st);
// CKKSEncoderCuda::encodePolynomial / decodePolynomial (src/ckks_cuda.cu:455-572,
// 983-1053): count <= N real coefficients times scale, rounded half away from zero, in
// NTT form (workspace: ckks_encode_workspace); decode gives the N real coefficients over
// scale (workspace: ckks_decode_workspace).
void ckks_encode_polynomial(const Context &c, const double *values, uint32_t count, double scale, uint32_t L,
                            uint64_t *plain, void *ws, size_t ws_bytes, hipStream_t st);
void ckks_decode_polynomial(const Context &c, const uint64_t *plain, uint32_t L, double scale, double *values,
                            void *ws, size_t ws_bytes, hipStream_t st);

}  // namespace shc
