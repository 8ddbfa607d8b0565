// ckks_encoder.hip -- CKKS canonical-embedding encode / decode in fp64 on gfx950.
//
// Same algorithm as CKKSEncoder (src/ckks.cpp:98-277 encode, :474-580 decode; GPU
// twin src/ckks_cuda.cu:314-453, 900-971): conjugate-symmetric scatter through the
// slot index map, inverse DWT over the complex 2N-th roots with scale/N folded into
// the last layer, rounding, reduction into each limb, NTT; decode runs the inverse.
// Decode composes the CRT value without multi-precision loops: the wrap count
// k = round(sum_j y_j / q_j) comes from a double sum (exact while |x| << Q) and the
// centred integer x = sum_j y_j Qhat_j - k Q is formed exactly modulo 2^128.
#include "ckks_encoder.h"
#include "modarith.h"
#include "ntt.h"
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace shc {
namespace {

constexpr int TPB = 256;

inline uint32_t blocks_for(size_t n) {
    const size_t b = (n + TPB - 1) / TPB;
    return (uint32_t)(b < (size_t(1) << 20) ? b : (size_t(1) << 20));
}

// slot -> coefficient position (CKKSEncoder matrix_reps_index_map_, src/ckks.cpp:44-56)
__device__ __forceinline__ uint32_t slot_pos(uint32_t i, int logn) {
    const uint64_t n = 1ull << logn, slots = n >> 1, mask = 2 * n - 1;
    const uint32_t e = i < slots ? i : i - (uint32_t)slots;
    uint64_t pos = 1, b = 3;
    for (uint32_t k = e; k; k >>= 1) {
        if (k & 1) pos = (pos * b) & mask;
        b = (b * b) & mask;
    }
    const uint64_t idx = i < slots ? (pos - 1) >> 1 : (2 * n - pos - 1) >> 1;
    return logn ? __brev((uint32_t)idx) >> (32 - logn) : 0;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Every kernel here takes plaintext blockIdx.y of a batch (its data at a fixed stride).
__global__ void scatter_kernel(const double *__restrict__ values, uint32_t count, size_t vstride,
                               double2 *__restrict__ a, int logn) {
    const uint32_t slots = 1u << (logn - 1);
    values += blockIdx.y * vstride;
    a += (size_t)blockIdx.y << logn;
    for (uint32_t i = blockIdx.x * TPB + threadIdx.x; i < slots; i += gridDim.x * TPB) {
        double2 v = make_double2(0.0, 0.0);
        if (i < count) v = make_double2(values[2 * i], values[2 * i + 1]);
        a[slot_pos(i, logn)] = v;
        a[slot_pos(i + slots, logn)] = make_double2(v.x, -v.y);
    }
}

// ---- DWT (DWTHandler::transformFromRev / transformToRev, src/utils/dwthandler.h) ----
// Two passes over HBM instead of one per layer.  A block holds
// FFT_TILE values of one plaintext (blockIdx.y) in LDS and runs every layer whose
// butterflies stay inside its lines: the row pass takes 2^kr contiguous values
// (kr = min(logn, 12)), the layers with gap G < 2^kr; the column pass takes W = FFT_TILE
// >> kc neighbouring columns of 2^kc values at stride 2^kr (kc = logn - kr), the layers
// with G >= 2^kr.  Each butterfly is the reference's (Gentleman-Sande for the inverse,
// scale folded into the last layer; Cooley-Tukey forward) with the same twiddle, so the
// words are those of the layer-by-layer transform.
constexpr int FFT_TILE = 4096;

// inverse (Gentleman-Sande) layer of gap G reads irp[N - N/G + 1 + group]; forward
// (Cooley-Tukey) layer of gap G reads rp[N/(2G) + group]; group = position / (2G).
template <bool INV>
__device__ __forceinline__ void dwt_butterfly(double2 *t, uint32_t x, uint32_t y, double2 w, bool last,
                                              double scalar) {
    const double2 u = t[x];
    if (INV) {
        const double2 v = t[y];
        double2 s = make_double2(u.x + v.x, u.y + v.y), d = make_double2(u.x - v.x, u.y - v.y);
        if (last) {
            w = make_double2(w.x * scalar, w.y * scalar);
            s = make_double2(s.x * scalar, s.y * scalar);
        }
        t[x] = s;
        t[y] = cmul(d, w);
    } else {
        const double2 v = cmul(t[y], w);
        t[x] = make_double2(u.x + v.x, u.y + v.y);
        t[y] = make_double2(u.x - v.x, u.y - v.y);
    }
}

template <bool INV>
__global__ __launch_bounds__(TPB) void fft_rows_kernel(double2 *__restrict__ a, const double2 *__restrict__ r,
                                                      int logn, int kr, double scalar) {
    __shared__ double2 t[FFT_TILE];
    const uint32_t n = 1u << logn, len = 1u << kr, base = blockIdx.x << kr;
    a += ((size_t)blockIdx.y << logn) + base;
    for (uint32_t k = threadIdx.x; k < len; k += TPB) t[k] = a[k];
    __syncthreads();
    for (uint32_t step = 0; step < (uint32_t)kr; step++) {
        const uint32_t G = INV ? 1u << step : len >> (step + 1);
        const uint32_t ridx = INV ? n - n / G + 1 : n / (2 * G);
        const bool last = INV && 2 * G == n;
        for (uint32_t b = threadIdx.x; b < len / 2; b += TPB) {
            const uint32_t il = b / G, j = b % G, x = 2 * G * il + j;
            dwt_butterfly<INV>(t, x, x + G, r[ridx + (base + x) / (2 * G)], last, scalar);
        }
        __syncthreads();
    }
    for (uint32_t k = threadIdx.x; k < len; k += TPB) a[k] = t[k];
}

template <bool INV>
__global__ __launch_bounds__(TPB) void fft_cols_kernel(double2 *__restrict__ a, const double2 *__restrict__ r,
                                                      int logn, int kr, double scalar) {
    __shared__ double2 t[FFT_TILE];
    const int kc = logn - kr;
    const uint32_t n = 1u << logn, rows = 1u << kc, W = FFT_TILE >> kc, c0 = blockIdx.x * W;
    a += ((size_t)blockIdx.y << logn) + c0;
    for (uint32_t e = threadIdx.x; e < FFT_TILE; e += TPB) {
        const uint32_t k = e / W, c = e % W;
        t[e] = a[((size_t)k << kr) + c];
    }
    __syncthreads();
    for (uint32_t step = 0; step < (uint32_t)kc; step++) {
        const uint32_t g = INV ? 1u << step : rows >> (step + 1), G = g << kr;
        const uint32_t ridx = INV ? n - n / G + 1 : n / (2 * G);
        const bool last = INV && 2 * G == n;
        for (uint32_t b = threadIdx.x; b < FFT_TILE / 2; b += TPB) {
            const uint32_t c = b % W, kb = b / W, il = kb / g, k = 2 * g * il + kb % g;
            dwt_butterfly<INV>(t, k * W + c, (k + g) * W + c, r[ridx + il], last, scalar);
        }
        __syncthreads();
    }
    for (uint32_t e = threadIdx.x; e < FFT_TILE; e += TPB) {
        const uint32_t k = e / W, c = e % W;
        a[((size_t)k << kr) + c] = t[e];
    }
}

// The whole DWT of `batch` plaintexts of N = 2^logn values at a (inverse: scalar folded
// into the last layer).
template <bool INV>
void launch_dwt(double2 *a, const double2 *r, int logn, uint32_t batch, double scalar, hipStream_t st) {
    const int kr = logn < 12 ? logn : 12, kc = logn - kr;
    const dim3 rows_grid(1u << kc, batch), cols_grid(kc ? (1u << kr) / (FFT_TILE >> kc) : 0, batch);
    if (INV) {
        fft_rows_kernel<true><<<rows_grid, TPB, 0, st>>>(a, r, logn, kr, scalar);
        if (kc) fft_cols_kernel<true><<<cols_grid, TPB, 0, st>>>(a, r, logn, kr, scalar);
    } else {
        if (kc) fft_cols_kernel<false><<<cols_grid, TPB, 0, st>>>(a, r, logn, kr, scalar);
        fft_rows_kernel<false><<<rows_grid, TPB, 0, st>>>(a, r, logn, kr, scalar);
    }
}

// round the real parts (half away from zero, as std::round in the reference's
// src/ckks.cpp:176 and src/ckks_cuda.cu:220), reduce into each limb; record max
// |coefficient| (positive doubles order like their bit patterns, so an integer
// atomicMax works)
// atomicMax of the active lanes' values: with the whole wave active, its maximum first
// (a butterfly of lane shuffles) and one atomic from lane 0.  A non-negative double's
// bit pattern orders like the double.
__device__ __forceinline__ void wave_max_atomic(unsigned long long *dst, unsigned long long v) {
    const uint64_t active = __ballot(1);
    if (active == ~0ull) {
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const unsigned long long o = __shfl_xor(v, m);
            v = o > v ? o : v;
        }
        if (__lane_id() == 0) atomicMax(dst, v);
    } else {
        atomicMax(dst, v);  // partial wave (the grid's tail): one atomic per lane
    }
}

__global__ void round_reduce_kernel(const double2 *__restrict__ a, uint64_t *__restrict__ out,
                                    const DevPrime *__restrict__ primes, uint32_t L, int logn,
                                    unsigned long long *maxbits) {
    const size_t n = (size_t)1 << logn;
    const double two64 = 18446744073709551616.0;
    a += blockIdx.y * n;
    out += blockIdx.y * L * n;
    maxbits += blockIdx.y;
    for (size_t c = (size_t)blockIdx.x * TPB + threadIdx.x; c < n; c += (size_t)gridDim.x * TPB) {
        double d = round(a[c].x);
        const bool neg = signbit(d);
        d = fabs(d);
        wave_max_atomic(maxbits, (unsigned long long)__double_as_longlong(d));
        if (!(d < two64 * two64)) d = 0;  // too large, inf or NaN: rejected on the host after the range check
        // d = hi 2^64 + lo exactly: scaling by a power of two is exact, and d mod 2^64 (the
        // low bits of a 53-bit mantissa) is representable, so the subtraction is exact
        const double hid = floor(d * (1.0 / two64));
        const uint64_t hi = (uint64_t)hid;
        const uint64_t lo = (uint64_t)(d - hid * two64);
        for (uint32_t l = 0; l < L; l++) {
            const DevPrime &P = primes[l];
            const uint64_t r = barrett128(u128{lo, hi}, P.q, P.ratio_lo, P.ratio_hi);
            out[l * n + c] = neg ? negmod(r, P.q) : r;
        }
    }
}

// Polynomial encoding: coefficient i is values[i] * scale (zero past count); the
// maximum is taken before rounding, as the reference's (src/ckks_cuda.cu:490-522).
__global__ void poly_scale_kernel(const double *__restrict__ values, uint32_t count, double scale,
                                  double2 *__restrict__ a, int logn, unsigned long long *maxbits) {
    const size_t n = (size_t)1 << logn;
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) {
        const double v = i < count ? values[i] * scale : 0.0;
        a[i] = make_double2(v, 0.0);
        wave_max_atomic(maxbits, (unsigned long long)__double_as_longlong(fabs(v)));
    }
}

// Polynomial decoding: the real coefficients of a composed plaintext.
__global__ void real_parts_kernel(const double2 *__restrict__ a, double *__restrict__ values, int logn) {
    const size_t n = (size_t)1 << logn;
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) values[i] = a[i].x;
}

__global__ void clear_words_kernel(unsigned long long *w, uint32_t words) {
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) w[i] = 0;
}

// The encoder's range check on the device: status = 0, or SHC_INVALID_ARGUMENT (-1) when
// the largest |coefficient| is not finite or needs total_bits or more bits, or more than
// 128 (CKKSEncoder::encodeInternal's "encoded values are too large", src/ckks.cpp:167-171).
// (one thread per plaintext: plaintext p's maximum and status word)
__global__ void encode_check_kernel(const unsigned long long *maxbits, int total_bits, int *status, uint32_t count) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    maxbits += p;
    status += p;
    const double maxc = __longlong_as_double((long long)*maxbits);
    int st = 0;
    if (!isfinite(maxc)) {
        st = -1;
    } else {
        const int bits = (int)ceil(log2(fmax(maxc, 1.0))) + 1;
        if (bits >= total_bits || bits > 128) st = -1;
    }
    *status = st;
}

// A plaintext the check rejected is left all zero.
__global__ void zero_if_rejected_kernel(uint64_t *__restrict__ plain, const int *__restrict__ status, size_t words) {
    plain += blockIdx.y * words;
    if (status[blockIdx.y] == 0) return;
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < words; i += (size_t)gridDim.x * TPB) plain[i] = 0;
}

__global__ void compose_kernel(const uint64_t *__restrict__ c, double2 *__restrict__ a, CkksLevelDev T,
                               const DevPrime *__restrict__ primes, int logn, double inv_scale) {
    const size_t n = (size_t)1 << logn;
    const double two64 = 18446744073709551616.0;
    c += blockIdx.y * T.L * n;
    a += blockIdx.y * n;
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) {
        double f = 0.0;
        uint64_t lo = 0, hi = 0;
        for (uint32_t j = 0; j < T.L; j++) {
            const uint64_t y = shoup(c[j * n + i], T.qhat_inv[2 * j], T.qhat_inv[2 * j + 1], primes[j].q);
            f += (double)y * T.inv_q[j];
            // (lo, hi) += y * qhat_j mod 2^128
            const u128 p = mul64x64(y, T.qhat128[2 * j]);
            const uint64_t nlo = lo + p.lo;
            hi += p.hi + y * T.qhat128[2 * j + 1] + (nlo < lo ? 1ull : 0ull);
            lo = nlo;
        }
        const uint64_t k = (uint64_t)llrint(f);
        // x = acc - k Q (mod 2^128), read as a signed 128-bit integer
        const u128 kq = mul64x64(k, T.Q_lo);
        const uint64_t kq_hi = kq.hi + k * T.Q_hi;
        const uint64_t xlo = lo - kq.lo;
        const uint64_t xhi = hi - kq_hi - (lo < kq.lo ? 1ull : 0ull);
        double v;
        if (xhi >> 63) {
            const uint64_t nlo = ~xlo + 1, nhi = ~xhi + (nlo == 0 ? 1ull : 0ull);
            v = -((double)nhi * two64 + (double)nlo);
        } else {
            v = (double)xhi * two64 + (double)xlo;
        }
        a[i] = make_double2(v * inv_scale, 0.0);
    }
}

__global__ void gather_kernel(const double2 *__restrict__ a, double *__restrict__ values, int logn) {
    const uint32_t slots = 1u << (logn - 1);
    a += (size_t)blockIdx.y << logn;
    values += (size_t)blockIdx.y << logn;
    for (uint32_t i = blockIdx.x * TPB + threadIdx.x; i < slots; i += gridDim.x * TPB) {
        const double2 v = a[slot_pos(i, logn)];
        values[2 * i] = v.x;
        values[2 * i + 1] = v.y;
    }
}

// ComplexRoots::getRoot (src/utils/croots.cpp:31-60)
std::complex<double> root_of(size_t m, size_t index) {
    index &= m - 1;
    if (index <= m / 8) return std::polar(1.0, 2 * M_PI * (double)index / (double)m);
    if (index <= m / 4) {
        auto r = root_of(m, m / 4 - index);
        return {r.imag(), r.real()};
    }
    if (index <= m / 2) return -std::conj(root_of(m, m / 2 - index));
    if (index <= 3 * m / 4) return -root_of(m, index - m / 2);
    return std::conj(root_of(m, m - index));
}

}  // namespace

void build_ckks_tables(Context &c) {
    auto t = std::make_shared<CkksTables>();
    const size_t n = c.n, m = 2 * n;
    const int logn = c.logn;
    std::vector<double> rt(4 * n, 0.0);  // [rp re, im]... then [irp]
    auto put = [&](size_t base, size_t i, std::complex<double> v) {
        rt[2 * (base + i)] = v.real();
        rt[2 * (base + i) + 1] = v.imag();
    };
    put(0, 0, 1.0);
    put(n, 0, 1.0);
    if (m >= 8) {
        for (size_t i = 1; i < n; i++) {
            put(0, i, root_of(m, hm::reverse_bits((uint32_t)i, logn)));
            put(n, i, std::conj(root_of(m, hm::reverse_bits((uint32_t)(i - 1), logn) + 1)));
        }
    } else if (m == 4) {
        put(0, 1, {0, 1});
        put(n, 1, {0, -1});
    }
    t->roots = DevBuf(rt.size() * 8);
    SHC_HIP(hipMemcpy(t->roots.ptr, rt.data(), rt.size() * 8, hipMemcpyHostToDevice));
    const size_t K = c.moduli.size();
    t->levels.resize(K + 1);
    for (size_t L = 1; L <= K; L++) {
        std::vector<uint64_t> q(c.moduli.begin(), c.moduli.begin() + L);
        std::vector<uint64_t> words;
        std::vector<double> invq;
        const hm::Big Q = hm::big_product(q);
        for (size_t j = 0; j < L; j++) {
            const hm::Big qh = hm::big_product_except(q, j);
            const uint64_t inv = hm::invmod(hm::big_mod_word(qh, q[j]), q[j]);
            words.push_back(inv);
            words.push_back(hm::shoup(inv, q[j]));
        }
        for (size_t j = 0; j < L; j++) {
            const hm::Big qh = hm::big_product_except(q, j);
            words.push_back(qh.size() > 0 ? qh[0] : 0);
            words.push_back(qh.size() > 1 ? qh[1] : 0);
            invq.push_back(1.0 / (double)q[j]);
        }
        DevBuf b(words.size() * 8 + invq.size() * 8);
        SHC_HIP(hipMemcpy(b.ptr, words.data(), words.size() * 8, hipMemcpyHostToDevice));
        SHC_HIP(hipMemcpy(static_cast<char *>(b.ptr) + words.size() * 8, invq.data(), invq.size() * 8,
                          hipMemcpyHostToDevice));
        CkksLevelDev &d = t->levels[L];
        d.L = (uint32_t)L;
        d.Q_lo = Q.size() > 0 ? Q[0] : 0;
        d.Q_hi = Q.size() > 1 ? Q[1] : 0;
        d.qhat_inv = b.as<const uint64_t>();
        d.qhat128 = b.as<const uint64_t>() + 2 * L;
        d.inv_q = reinterpret_cast<const double *>(static_cast<char *>(b.ptr) + words.size() * 8);
        d.total_bits = hm::big_bit_count(Q);
        t->bufs.push_back(std::move(b));
    }
    c.ckks = t;
}

size_t ckks_encode_workspace(const Context &c) { return ckks_encode_workspace_many(c, 1); }

namespace {

// Reads the per-plaintext maxima back and raises the reference's range errors
// (CKKSEncoder::encodeInternal, src/ckks.cpp:167-171).
void host_range_check(const unsigned long long *maxbits, uint32_t batch, const CkksLevelDev &T, hipStream_t st) {
    std::vector<unsigned long long> mb(batch);
    SHC_HIP(hipMemcpyAsync(mb.data(), maxbits, 8 * (size_t)batch, hipMemcpyDeviceToHost, st));
    SHC_HIP(hipStreamSynchronize(st));
    for (uint32_t p = 0; p < batch; p++) {
        double maxc;
        std::memcpy(&maxc, &mb[p], 8);
        // an infinite or NaN input makes the maximum non-finite (NaN orders above +inf as
        // an unsigned bit pattern), which no coefficient modulus can hold
        if (!std::isfinite(maxc)) fail(ERR_INVALID_ARGUMENT, "encoded values are not finite");
        const int bits = (int)std::ceil(std::log2(std::max(maxc, 1.0))) + 1;
        if (bits >= T.total_bits || bits > 128) fail(ERR_INVALID_ARGUMENT, "encoded values are too large");
    }
}

// Encode `batch` plaintexts up to the forward NTT: values of plaintext p at
// values + p * vstride, plaintext p at plain + p * L * N.  d_status: the range checks go
// to d_status[p] on the device (a rejected plaintext is zero) and nothing waits for the
// stream; otherwise they are read back and raised like the reference's.
void encode_body(const Context &c, const double *values, uint32_t count, size_t vstride, uint32_t batch, double scale,
                 uint32_t L, uint64_t *plain, void *ws, size_t ws_bytes, int *d_status, hipStream_t st) {
    if (c.scheme != SCHEME_CKKS || !c.ckks) fail(ERR_INVALID_ARGUMENT, "unsupported scheme");
    if (c.logn < 1) fail(ERR_INVALID_ARGUMENT, "poly_modulus_degree too small");
    if (count > c.n / 2) fail(ERR_INVALID_ARGUMENT, "values_size is too large");
    if (batch == 0) return;
    if (batch > 65535) fail(ERR_INVALID_ARGUMENT, "at most 65535 plaintexts per call");
    if (!ws || ws_bytes < ckks_encode_workspace_many(c, batch))
        fail(ERR_INVALID_ARGUMENT, "workspace too small for encoding");
    const CkksLevelDev &T = c.ckks->levels[L];
    // CKKSEncoder::encodeInternal scale check (src/ckks.cpp:120-123)
    if (!(scale > 0) || (int)std::log2(scale) + 1 >= T.total_bits) fail(ERR_INVALID_ARGUMENT, "scale out of bounds");
    const size_t n = c.n;
    double2 *a = static_cast<double2 *>(ws);
    auto *maxbits = reinterpret_cast<unsigned long long *>(static_cast<char *>(ws) + batch * n * 16);
    const double2 *irp = c.ckks->roots.as<const double2>() + n;
    scatter_kernel<<<dim3(blocks_for(n / 2), batch), TPB, 0, st>>>(values, count, vstride, a, c.logn);
    launch_dwt<true>(a, irp, c.logn, batch, scale / (double)n, st);
    clear_words_kernel<<<1, TPB, 0, st>>>(maxbits, batch);
    round_reduce_kernel<<<dim3(blocks_for(n), batch), TPB, 0, st>>>(a, plain, c.primes(), L, c.logn, maxbits);
    launch_check("ckks encode");
    if (d_status) {
        encode_check_kernel<<<(batch + TPB - 1) / TPB, TPB, 0, st>>>(maxbits, T.total_bits, d_status, batch);
        zero_if_rejected_kernel<<<dim3(blocks_for((size_t)L * n), batch), TPB, 0, st>>>(plain, d_status,
                                                                                         (size_t)L * n);
        launch_check("ckks encode check");
        return;
    }
    host_range_check(maxbits, batch, T, st);
}

}  // namespace

size_t ckks_encode_workspace_many(const Context &c, uint32_t batch) {
    return (size_t)batch * c.n * 16 + ((8 * (size_t)batch + 255) & ~size_t(255));
}

void ckks_encode(const Context &c, const double *values, uint32_t count, double scale, uint32_t L, uint64_t *plain,
                 void *ws, size_t ws_bytes, hipStream_t st) {
    encode_body(c, values, count, 0, 1, scale, L, plain, ws, ws_bytes, nullptr, st);
    ntt_forward(plain, L, L, c.primes(), c.logn, NTT_FULL, st);
}

void ckks_encode_deferred(const Context &c, const double *values, uint32_t count, double scale, uint32_t L,
                          uint64_t *plain, int *d_status, void *ws, size_t ws_bytes, hipStream_t st) {
    if (!d_status) fail(ERR_INVALID_ARGUMENT, "null status word");
    encode_body(c, values, count, 0, 1, scale, L, plain, ws, ws_bytes, d_status, st);
    ntt_forward(plain, L, L, c.primes(), c.logn, NTT_FULL, st);
}

void ckks_encode_many(const Context &c, const double *values, uint32_t count, size_t vstride, uint32_t batch,
                      double scale, uint32_t L, uint64_t *plain, int *d_status, void *ws, size_t ws_bytes,
                      hipStream_t st) {
    encode_body(c, values, count, vstride, batch, scale, L, plain, ws, ws_bytes, d_status, st);
    if (batch) ntt_forward(plain, batch * L, L, c.primes(), c.logn, NTT_FULL, st);
}

size_t ckks_decode_workspace(const Context &c, uint32_t L) { return ckks_decode_workspace_many(c, L, 1); }

size_t ckks_decode_workspace_many(const Context &c, uint32_t L, uint32_t batch) {
    return (((size_t)batch * L * c.n * 8 + 255) & ~size_t(255)) + (size_t)batch * c.n * 16;
}

void ckks_decode_many(const Context &c, const uint64_t *plain, uint32_t L, uint32_t batch, double scale,
                      double *values, void *ws, size_t ws_bytes, hipStream_t st) {
    if (c.scheme != SCHEME_CKKS || !c.ckks) fail(ERR_INVALID_ARGUMENT, "unsupported scheme");
    if (batch == 0) return;
    if (batch > 65535) fail(ERR_INVALID_ARGUMENT, "at most 65535 plaintexts per call");
    if (!ws || ws_bytes < ckks_decode_workspace_many(c, L, batch))
        fail(ERR_INVALID_ARGUMENT, "workspace too small for decoding");
    const CkksLevelDev &T = c.ckks->levels[L];
    if (!(scale > 0) || (int)std::log2(scale) >= T.total_bits) fail(ERR_INVALID_ARGUMENT, "scale out of bounds");
    const size_t n = c.n;
    uint64_t *cp = static_cast<uint64_t *>(ws);
    double2 *a = reinterpret_cast<double2 *>(static_cast<char *>(ws) + (((size_t)batch * L * n * 8 + 255) & ~size_t(255)));
    const LimbGrid lg{1, L, L, batch};  // the plaintexts' limbs, transformed out of place
    ntt_inverse_from(plain, lg, cp, lg, c.primes(), c.logn, NTT_FULL, st);
    compose_kernel<<<dim3(blocks_for(n), batch), TPB, 0, st>>>(cp, a, T, c.primes(), c.logn, 1.0 / scale);
    launch_dwt<false>(a, c.ckks->roots.as<const double2>(), c.logn, batch, 0.0, st);
    gather_kernel<<<dim3(blocks_for(n / 2), batch), TPB, 0, st>>>(a, values, c.logn);
    launch_check("ckks decode");
}

void ckks_decode(const Context &c, const uint64_t *plain, uint32_t L, double scale, double *values, void *ws,
                 size_t ws_bytes, hipStream_t st) {
    ckks_decode_many(c, plain, L, 1, scale, values, ws, ws_bytes, st);
}

void ckks_encode_polynomial(const Context &c, const double *values, uint32_t count, double scale, uint32_t L,
                            uint64_t *plain, void *ws, size_t ws_bytes, hipStream_t st) {
    if (c.scheme != SCHEME_CKKS || !c.ckks) fail(ERR_INVALID_ARGUMENT, "unsupported scheme");
    if (count > c.n) fail(ERR_INVALID_ARGUMENT, "values_size is too large");
    if (count && !values) fail(ERR_INVALID_ARGUMENT, "values cannot be null");
    if (!ws || ws_bytes < ckks_encode_workspace_many(c, 1)) fail(ERR_INVALID_ARGUMENT, "workspace too small for encoding");
    const CkksLevelDev &T = c.ckks->levels[L];
    const size_t n = c.n;
    double2 *a = static_cast<double2 *>(ws);
    auto *maxbits = reinterpret_cast<unsigned long long *>(static_cast<char *>(ws) + n * 16);
    clear_words_kernel<<<1, TPB, 0, st>>>(maxbits, 1);
    poly_scale_kernel<<<blocks_for(n), TPB, 0, st>>>(values, count, scale, a, c.logn, maxbits);
    round_reduce_kernel<<<dim3(blocks_for(n), 1), TPB, 0, st>>>(a, plain, c.primes(), L, c.logn, maxbits);
    launch_check("ckks polynomial encode");
    host_range_check(maxbits, 1, T, st);
    ntt_forward(plain, L, L, c.primes(), c.logn, NTT_FULL, st);
}

void ckks_decode_polynomial(const Context &c, const uint64_t *plain, uint32_t L, double scale, double *values,
                            void *ws, size_t ws_bytes, hipStream_t st) {
    if (c.scheme != SCHEME_CKKS || !c.ckks) fail(ERR_INVALID_ARGUMENT, "unsupported scheme");
    if (!ws || ws_bytes < ckks_decode_workspace_many(c, L, 1)) fail(ERR_INVALID_ARGUMENT, "workspace too small for decoding");
    const CkksLevelDev &T = c.ckks->levels[L];
    if (!(scale > 0) || (int)std::log2(scale) >= T.total_bits) fail(ERR_INVALID_ARGUMENT, "scale out of bounds");
    const size_t n = c.n;
    uint64_t *cp = static_cast<uint64_t *>(ws);
    double2 *a = reinterpret_cast<double2 *>(static_cast<char *>(ws) + (((size_t)L * n * 8 + 255) & ~size_t(255)));
    SHC_HIP(hipMemcpyAsync(cp, plain, (size_t)L * n * 8, hipMemcpyDeviceToDevice, st));
    ntt_inverse(cp, L, L, c.primes(), c.logn, NTT_FULL, st);
    compose_kernel<<<dim3(blocks_for(n), 1), TPB, 0, st>>>(cp, a, T, c.primes(), c.logn, 1.0 / scale);
    real_parts_kernel<<<blocks_for(n), TPB, 0, st>>>(a, values, c.logn);
    launch_check("ckks polynomial decode");
}

}  // namespace shc
