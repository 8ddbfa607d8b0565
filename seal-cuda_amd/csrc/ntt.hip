// ntt.hip -- batched negacyclic NTT / inverse NTT for gfx950.
//
// Same transform as the reference's nttNegacyclicHarvey / inverseNttNegacyclicHarvey
// (src/utils/ntt.cpp:158-213, src/utils/dwthandler.h:88-372; GPU launchers
// src/kernelutils.cu:330-476), scheduled for CDNA4 instead of one launch per layer:
//
//   * A length-2^logn limb is viewed as a 2^K1 x 2^K2 matrix.  The first K1 layers
//     act along columns, the last K2 layers inside contiguous rows, so a forward
//     transform is two kernels (columns pass, rows pass) and each reads and writes
//     the limb once.  Up to 2^12 points the whole limb is one tile (one kernel).
//   * A tile lives in LDS; each thread pulls 2^E values into VGPRs and runs E
//     butterfly layers there before writing them back, so one LDS round trip
//     covers E layers.
//   * Harvey lazy butterflies with Shoup twiddles keep values in [0,4q) (forward)
//     or [0,2q) (inverse); the full reduction and the N^-1 scaling are folded into
//     the store of the last pass.
//
// Twiddle indexing: forward layer l (m = 2^l groups) uses root_powers[m + group];
// inverse layer l uses inv_root_powers[N - 2m + 1 + group], group = x >> (logn - l).
#include "ntt.h"
#include "modarith.h"

namespace shc {
namespace {

__host__ __device__ constexpr int lds_pad(int a) { return a + (a >> 4); }

template <bool INV>
__device__ __forceinline__ void butterfly(uint64_t &x, uint64_t &y, ulonglong2 w, uint64_t q, uint64_t two_q) {
    if (!INV) {
        uint64_t u = x >= two_q ? x - two_q : x;
        uint64_t v = shoup_lazy(y, w.x, w.y, q);
        x = u + v;
        y = u + two_q - v;
    } else {
        uint64_t u = x, v = y;
        uint64_t s = u + v;
        x = s >= two_q ? s - two_q : s;
        y = shoup_lazy(u + two_q - v, w.x, w.y, q);
    }
}

// Position inside a 2^K-point sub-transform of register j of thread t, for the phase
// that runs local layers [S0, S0+EP): the EP butterfly bits sit at bit positions
// [K-S0-EP, K-S0) of the index and come from the low EP bits of j.
template <int K, int E, int S0, int EP>
__device__ __forceinline__ int phase_index(int t, int j) {
    constexpr int LOWB = K - S0 - EP;
    const int jl = j & ((1 << EP) - 1);
    const int jh = j >> EP;
    const int other = (t << (E - EP)) | jh;
    return ((other >> LOWB) << (K - S0)) | (jl << LOWB) | (other & ((1 << LOWB) - 1));
}

// One phase on the thread's registers.  g = index of this sub-transform among the
// 2^L0 sub-transforms the full transform splits into at global layer L0.
template <int LOGN, int K, int E, int S0, int EP, bool INV>
__device__ __forceinline__ void phase_compute(uint64_t (&r)[1 << E], int t, int g, int L0,
                                              const ulonglong2 *__restrict__ tw, uint64_t q, uint64_t two_q) {
#pragma unroll
    for (int ss = 0; ss < EP; ss++) {
        const int s = INV ? (S0 + EP - 1 - ss) : (S0 + ss);
        const int b = S0 + EP - 1 - s;
#pragma unroll
        for (int j = 0; j < (1 << E); j++) {
            if (j & (1 << b)) continue;
            const int idx = phase_index<K, E, S0, EP>(t, j);
            const int grp = (g << s) + (idx >> (K - s));
            const int l = L0 + s;
            const ulonglong2 w = INV ? tw[(1 << LOGN) - (2 << l) + 1 + grp] : tw[(1 << l) + grp];
            butterfly<INV>(r[j], r[j | (1 << b)], w, q, two_q);
        }
    }
}

// Runs all phases of a sub-transform whose values sit in LDS; `addr(idx)` maps an
// index of this thread's sub-transform to its LDS slot.
template <int LOGN, int K, int E, bool INV, int P = 0, class Addr>
__device__ __forceinline__ void run_phases(uint64_t *tile, const Addr &addr, int t, int g, int L0,
                                           const ulonglong2 *__restrict__ tw, uint64_t q, uint64_t two_q,
                                           bool active) {
    constexpr int NP = (K + E - 1) / E;
    if constexpr (P < NP) {
        // forward runs phases from layer 0 upwards, inverse from the top layer down
        constexpr int PI = INV ? (NP - 1 - P) : P;
        constexpr int S0 = PI * E;
        constexpr int EP = (S0 + E <= K) ? E : (K - S0);
        uint64_t r[1 << E];
        if (active) {
#pragma unroll
            for (int j = 0; j < (1 << E); j++) r[j] = tile[addr(phase_index<K, E, S0, EP>(t, j))];
            phase_compute<LOGN, K, E, S0, EP, INV>(r, t, g, L0, tw, q, two_q);
#pragma unroll
            for (int j = 0; j < (1 << E); j++) tile[addr(phase_index<K, E, S0, EP>(t, j))] = r[j];
        }
        __syncthreads();
        run_phases<LOGN, K, E, INV, P + 1>(tile, addr, t, g, L0, tw, q, two_q, active);
    }
}

// Final touch of a value leaving the last pass.
template <bool INV, bool FULL>
__device__ __forceinline__ uint64_t finish(uint64_t x, const DevPrime &P) {
    if (!INV) {
        if (FULL) {
            x = x >= P.two_q ? x - P.two_q : x;
            x = x >= P.q ? x - P.q : x;
        }
    } else {
        x = shoup_lazy(x, P.inv_n, P.inv_n_shoup, P.q);
        if (FULL) x = x >= P.q ? x - P.q : x;
    }
    return x;
}

// Rows pass: SUBS contiguous rows of 2^K values per block; the row is the last K
// layers of the limb (or the whole limb when K == LOGN).
template <int LOGN, int K, int E, int SUBS, bool INV, bool LAST, bool FULL>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_rows_kernel(uint64_t *__restrict__ data,
                                                                  const DevPrime *__restrict__ primes,
                                                                  uint32_t limbs_per_poly, uint32_t total_rows) {
    constexpr int T = 1 << (K - E);
    constexpr int SZ = 1 << K;
    constexpr int RPL = LOGN - K;  // log2(rows per limb)
    __shared__ uint64_t tile[lds_pad(SUBS * SZ)];
    const int tid = threadIdx.x;
    const int sub = tid / T, t = tid % T;
    const uint32_t row0 = blockIdx.x * SUBS;
    const int nvalid = (int)min<uint32_t>(SUBS, total_rows - row0) * SZ;
    uint64_t *base = data + (size_t)row0 * SZ;
    for (int e = tid; e < nvalid; e += SUBS * T) tile[lds_pad(e)] = base[e];
    __syncthreads();

    const uint32_t row = row0 + sub;
    const bool active = row < total_rows;
    const uint32_t limb = active ? (row >> RPL) : 0;
    const DevPrime &P = primes[limb % limbs_per_poly];
    const int g = (int)(row & ((1u << RPL) - 1));
    const ulonglong2 *tw = INV ? P.inv : P.fwd;
    const uint64_t q = P.q, two_q = P.two_q;
    auto addr = [sub](int idx) { return lds_pad(sub * SZ + idx); };
    run_phases<LOGN, K, E, INV>(tile, addr, t, g, RPL, tw, q, two_q, active);

    for (int e = tid; e < nvalid; e += SUBS * T) {
        uint64_t x = tile[lds_pad(e)];
        if (LAST) {
            const uint32_t lb = (row0 + (uint32_t)(e / SZ)) >> RPL;
            x = finish<INV, FULL>(x, primes[lb % limbs_per_poly]);
        }
        base[e] = x;
    }
}

// Columns pass: SUBS adjacent columns of one limb per block; a column is the first
// K layers of the limb (stride 2^(LOGN-K) between its values).
template <int LOGN, int K, int E, int SUBS, bool INV, bool LAST, bool FULL>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_cols_kernel(uint64_t *__restrict__ data,
                                                                  const DevPrime *__restrict__ primes,
                                                                  uint32_t limbs_per_poly) {
    constexpr int SZ = 1 << K;
    constexpr int C = 1 << (LOGN - K);
    constexpr int TILES = C / SUBS;
    __shared__ uint64_t tile[lds_pad(SUBS * SZ)];
    const int tid = threadIdx.x;
    const int col = tid % SUBS, t = tid / SUBS;
    const uint32_t limb = blockIdx.x / TILES;
    const int c0 = (int)(blockIdx.x % TILES) * SUBS;
    uint64_t *base = data + ((size_t)limb << LOGN) + c0;
    const DevPrime &P = primes[limb % limbs_per_poly];
    for (int e = tid; e < SUBS * SZ; e += SUBS << (K - E)) tile[lds_pad(e)] = base[(e / SUBS) * C + (e % SUBS)];
    __syncthreads();
    const ulonglong2 *tw = INV ? P.inv : P.fwd;
    auto addr = [col](int idx) { return lds_pad(idx * SUBS + col); };
    run_phases<LOGN, K, E, INV>(tile, addr, t, 0, 0, tw, P.q, P.two_q, true);
    for (int e = tid; e < SUBS * SZ; e += SUBS << (K - E)) {
        uint64_t x = tile[lds_pad(e)];
        if (LAST) x = finish<INV, FULL>(x, P);
        base[(e / SUBS) * C + (e % SUBS)] = x;
    }
}

// ModUp columns pass: limb m = (ct, j, i) of the output T[ct][j][i] is digit j of
// ciphertext ct, read from d[ct][j], reduced modulo target prime i on load, then run
// through the first K layers (the key switch's gather + forward NTT pass 1 in one).
template <int LOGN, int K, int E, int SUBS>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_cols_modup_kernel(const uint64_t *__restrict__ d,
                                                                        size_t d_ct_stride,
                                                                        uint64_t *__restrict__ T,
                                                                        const DevPrime *__restrict__ targets,
                                                                        uint32_t L) {
    constexpr int SZ = 1 << K;
    constexpr int C = 1 << (LOGN - K);
    constexpr int TILES = C / SUBS;
    __shared__ uint64_t tile[lds_pad(SUBS * SZ)];
    const int tid = threadIdx.x;
    const int col = tid % SUBS, t = tid / SUBS;
    const uint32_t limb = blockIdx.x / TILES;
    const int c0 = (int)(blockIdx.x % TILES) * SUBS;
    const uint32_t i = limb % (L + 1);
    const uint32_t r = limb / (L + 1);
    const uint32_t j = r % L;
    const size_t ct = r / L;
    const uint64_t *in = d + ct * d_ct_stride + ((size_t)j << LOGN) + c0;
    uint64_t *out = T + ((size_t)limb << LOGN) + c0;
    const DevPrime &P = targets[i];
    for (int e = tid; e < SUBS * SZ; e += SUBS << (K - E))
        tile[lds_pad(e)] = barrett64(in[(e / SUBS) * C + (e % SUBS)], P.q, P.ratio_hi);
    __syncthreads();
    auto addr = [col](int idx) { return lds_pad(idx * SUBS + col); };
    run_phases<LOGN, K, E, false>(tile, addr, t, 0, 0, P.fwd, P.q, P.two_q, true);
    for (int e = tid; e < SUBS * SZ; e += SUBS << (K - E)) out[(e / SUBS) * C + (e % SUBS)] = tile[lds_pad(e)];
}

template <int LOGN, bool INV, bool FULL>
void launch_ntt(uint64_t *data, uint32_t count, uint32_t lpp, const DevPrime *primes, hipStream_t st) {
    if constexpr (LOGN <= 12) {
        constexpr int K = LOGN;
        constexpr int E = K < 4 ? K : 4;
        constexpr int T = 1 << (K - E);
        constexpr int SUBS = (T >= 256) ? 1 : (256 / T);
        const uint32_t blocks = (count + SUBS - 1) / SUBS;
        ntt_rows_kernel<LOGN, K, E, SUBS, INV, true, FULL><<<blocks, SUBS * T, 0, st>>>(data, primes, lpp, count);
    } else {
        constexpr int K1 = LOGN / 2;
        constexpr int K2 = LOGN - K1;
        constexpr int E = 4;
        constexpr int CS = 16;                               // columns per block
        constexpr int RS = (4096 >> K2) < 1 ? 1 : (4096 >> K2); // rows per block
        const uint32_t col_blocks = count * ((1u << K2) / CS);
        const uint32_t rows = count << K1;
        const uint32_t row_blocks = (rows + RS - 1) / RS;
        if (!INV) {
            ntt_cols_kernel<LOGN, K1, E, CS, false, false, FULL>
                <<<col_blocks, CS << (K1 - E), 0, st>>>(data, primes, lpp);
            ntt_rows_kernel<LOGN, K2, E, RS, false, true, FULL>
                <<<row_blocks, RS << (K2 - E), 0, st>>>(data, primes, lpp, rows);
        } else {
            ntt_rows_kernel<LOGN, K2, E, RS, true, false, FULL>
                <<<row_blocks, RS << (K2 - E), 0, st>>>(data, primes, lpp, rows);
            ntt_cols_kernel<LOGN, K1, E, CS, true, true, FULL>
                <<<col_blocks, CS << (K1 - E), 0, st>>>(data, primes, lpp);
        }
    }
}

template <bool INV, bool FULL>
void dispatch_chunk(uint64_t *data, uint32_t count, uint32_t lpp, const DevPrime *primes, int logn,
                    hipStream_t st) {
    switch (logn) {
#define SHC_NTT_CASE(L) \
    case L: launch_ntt<L, INV, FULL>(data, count, lpp, primes, st); break;
        SHC_NTT_CASE(1) SHC_NTT_CASE(2) SHC_NTT_CASE(3) SHC_NTT_CASE(4) SHC_NTT_CASE(5) SHC_NTT_CASE(6)
        SHC_NTT_CASE(7) SHC_NTT_CASE(8) SHC_NTT_CASE(9) SHC_NTT_CASE(10) SHC_NTT_CASE(11) SHC_NTT_CASE(12)
        SHC_NTT_CASE(13) SHC_NTT_CASE(14) SHC_NTT_CASE(15) SHC_NTT_CASE(16) SHC_NTT_CASE(17)
#undef SHC_NTT_CASE
        default: fail(ERR_INVALID_ARGUMENT, "poly_modulus_degree must be 2^1 .. 2^17");
    }
    launch_check("ntt launch");
}

// Splits very large batches so one launch stays well below 2^31 work-items (every
// pass runs N/16 work-items per limb); chunks hold whole polynomials so that limb k
// keeps prime k % lpp.
template <bool INV, bool FULL>
void dispatch(uint64_t *data, uint32_t count, uint32_t lpp, const DevPrime *primes, int logn, hipStream_t st) {
    uint32_t max_limbs = (uint32_t)((uint64_t(1) << 34) >> logn);
    max_limbs = max_limbs < lpp ? lpp : max_limbs - max_limbs % lpp;
    for (uint32_t done = 0; done < count;) {
        const uint32_t c = count - done < max_limbs ? count - done : max_limbs;
        dispatch_chunk<INV, FULL>(data + ((size_t)done << logn), c, lpp, primes, logn, st);
        done += c;
    }
}

template <int LOGN>
void launch_modup(const uint64_t *d, size_t d_stride, uint64_t *T, uint32_t cts, uint32_t L, const DevPrime *targets,
                  hipStream_t st) {
    constexpr int K1 = LOGN / 2;
    constexpr int K2 = LOGN - K1;
    constexpr int E = 4;
    constexpr int CS = 16;
    constexpr int RS = 4096 >> K2;
    const uint32_t limbs = cts * L * (L + 1);
    ntt_cols_modup_kernel<LOGN, K1, E, CS>
        <<<limbs * ((1u << K2) / CS), CS << (K1 - E), 0, st>>>(d, d_stride, T, targets, L);
    const uint32_t rows = limbs << K1;
    ntt_rows_kernel<LOGN, K2, E, RS, false, true, true>
        <<<(rows + RS - 1) / RS, RS << (K2 - E), 0, st>>>(T, targets, L + 1, rows);
}

}  // namespace

bool ntt_forward_modup(const uint64_t *d, size_t d_ct_stride, uint64_t *T, uint32_t count, uint32_t L,
                       const DevPrime *targets, int logn, hipStream_t st) {
    if (logn < 13 || logn > 17 || count == 0) return false;
    const uint64_t per_ct = (uint64_t)L * (L + 1);
    uint32_t max_cts = (uint32_t)(((uint64_t(1) << 34) >> logn) / per_ct);
    if (max_cts == 0) max_cts = 1;
    for (uint32_t done = 0; done < count;) {
        const uint32_t c = count - done < max_cts ? count - done : max_cts;
        const uint64_t *dd = d + (size_t)done * d_ct_stride;
        uint64_t *TT = T + ((size_t)done * per_ct << logn);
        switch (logn) {
            case 13: launch_modup<13>(dd, d_ct_stride, TT, c, L, targets, st); break;
            case 14: launch_modup<14>(dd, d_ct_stride, TT, c, L, targets, st); break;
            case 15: launch_modup<15>(dd, d_ct_stride, TT, c, L, targets, st); break;
            case 16: launch_modup<16>(dd, d_ct_stride, TT, c, L, targets, st); break;
            case 17: launch_modup<17>(dd, d_ct_stride, TT, c, L, targets, st); break;
        }
        launch_check("ntt modup launch");
        done += c;
    }
    return true;
}

void ntt_forward(uint64_t *data, uint32_t count, uint32_t lpp, const DevPrime *primes, int logn, NttOut out,
                 hipStream_t st) {
    if (count == 0) return;
    if (out == NTT_FULL) dispatch<false, true>(data, count, lpp, primes, logn, st);
    else dispatch<false, false>(data, count, lpp, primes, logn, st);
}

void ntt_inverse(uint64_t *data, uint32_t count, uint32_t lpp, const DevPrime *primes, int logn, NttOut out,
                 hipStream_t st) {
    if (count == 0) return;
    if (out == NTT_FULL) dispatch<true, true>(data, count, lpp, primes, logn, st);
    else dispatch<true, false>(data, count, lpp, primes, logn, st);
}

}  // namespace shc
