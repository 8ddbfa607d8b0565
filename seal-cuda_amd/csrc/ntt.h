// ntt.h -- host launchers for the batched negacyclic NTT / inverse NTT.
#pragma once
#include "common.h"

namespace shc {

// Output range of a transform.
enum NttOut { NTT_FULL = 0,   // fully reduced, [0, q)
              NTT_LAZY = 1 }; // forward: [0, 4q); inverse: [0, 2q)

// Limb layout of a batch: limb (i, j) is limb number i * A + j * B of the buffer and is
// reduced by primes[i], for i < I and j < J.  The [poly][limb][coeff] layout with m
// limbs per polynomial is {A = 1, B = m, I = m, J = polys}; a target-major key-switch
// buffer [prime][ct][digit] is {A = cts * digits, B = 1, I = primes, J = cts * digits}.
struct LimbGrid {
    uint32_t A, B, I, J;
};

// In-place transform of `count` limbs stored back to back, each of length 2^logn.
// Limb k (0 <= k < count) is reduced modulo primes[k % limbs_per_poly], i.e. the data is
// a [poly][limb][coeff] array with `limbs_per_poly` limbs per polynomial (count must be
// a multiple of limbs_per_poly).
//   forward: natural order in (values < 8q), bit-reversed order out
//   inverse: bit-reversed order in (values < 4q), natural order out, times N^-1
// Every prime must be below 2^61.
void ntt_forward(uint64_t *data, uint32_t count, uint32_t limbs_per_poly, const DevPrime *primes, int logn,
                 NttOut out, hipStream_t stream);
void ntt_inverse(uint64_t *data, uint32_t count, uint32_t limbs_per_poly, const DevPrime *primes, int logn,
                 NttOut out, hipStream_t stream);
// Forward / inverse transform of the I x J limbs of a LimbGrid.
void ntt_forward_grid(uint64_t *data, LimbGrid grid, const DevPrime *primes, int logn, NttOut out,
                      hipStream_t stream);
void ntt_inverse_grid(uint64_t *data, LimbGrid grid, const DevPrime *primes, int logn, NttOut out,
                      hipStream_t stream);
// Inverse transform of src's limbs (grid src_grid) into data (grid grid, same I x J);
// src is left as it was.
void ntt_inverse_from(const uint64_t *src, LimbGrid src_grid, uint64_t *data, LimbGrid grid, const DevPrime *primes,
                      int logn, NttOut out, hipStream_t stream);

// Key-switching ModUp fused with the forward NTT: T[i][ct][j] = NTT_i(d[ct][j] mod t_i)
// for digits j < L (primes digits[j]) and targets i <= L (targets[i]), d[ct] at
// d + ct * d_ct_stride with L limbs.  Target primes must be below 2^60; output in
// [0, 4 t_i) when reduce_4q, else [0, 8 t_i).  skip_diag: limbs T[i][ct][i] are not
// produced and targets i < L hold L - 1 limbs per ciphertext, T[i][ct][j - (j > i)]
// at limb i * count * L + ct * (L - 1) + ...; the special target keeps all L.  Returns
// false (and does nothing) for N < 2^13, where the caller gathers first and runs the
// mapped transform.
// With a probe, its events are recorded right before and after the rows-pass launch.
// `shape` overrides the launch shape (KsShape; rows_group applies to the fused form).
bool ntt_forward_modup(const uint64_t *d, size_t d_ct_stride, uint64_t *T, uint32_t count, uint32_t L,
                       const DevPrime *targets, const DevPrime *digits, int logn, bool skip_diag, bool reduce_4q,
                       hipStream_t stream, const KernelProbe *probe = nullptr, KsShape shape = KsShape{});

// The same ModUp with its rows pass fused into the products by the key (N = 2^15, 2^16;
// returns false and does nothing otherwise): the columns pass fills T as above (diag !=
// nullptr: skip_diag), then P[ct][k][i] = sum_j NTT_i(d_j mod t_i) key[j][k][ki] for
// i < L and S[ct][k] for the special target (ki = key_count - 1), the own-prime terms
// read from diag + ct * diag_stride + i * N.  key: [L.. ][2][key_count][N] rows of the
// key-switching key.  With a probe, its events bracket the fused launch.
// With fold.ct (CKKS: the NTT-form ciphertext the switch adds into, [count] at fold.ct_stride
// words, polynomials k = 0, 1 of L limbs each), P[ct][k][i] also takes ct_k[i] * (p mod q_i)
// (fold.dc: the special prime's divide-and-round constants), so that the ModDown writes
// (P - u) p^-1 without reading the ciphertext (ModDownArgs::folded).
// c1_zero: the added-into ciphertext is (ct_0, 0) -- a Galois switch, whose sigma(c0)
// stays where it was written and whose c1 is never materialised as zeros.
struct KsFold {
    const uint64_t *ct;
    size_t ct_stride;
    const DivRoundConst *dc;
    bool c1_zero = false;
};
bool ntt_modup_keyprod(const uint64_t *d, size_t d_ct_stride, uint64_t *T, const uint64_t *key, uint64_t *P,
                       uint64_t *S, const uint64_t *diag, size_t diag_stride, uint32_t count, uint32_t L,
                       uint32_t key_count, const DevPrime *targets, const DevPrime *digits, int logn,
                       hipStream_t stream, const KernelProbe *probe = nullptr, KsShape shape = KsShape{},
                       KsFold fold = KsFold{});

// ModDown by the special prime p fused into the last transforms of a key switch, for
// `count` ciphertexts at L limbs (N >= 2^13; returns false and does nothing otherwise):
//   u_i = ([s + floor(p/2)]_p mod q_i) - (floor(p/2) mod q_i),  ct_k[i] += (P_k[i] - u_i) p^-1
// (folded: P already holds p ct_k[i], and ct_k[i] = (P_k[i] - u_i) p^-1 is written unread)
// with s = S[ct][k] (special limb, coefficient form) and P = [ct][k][L][N].
//   CKKS: buf = U scratch [count][2][L][N]; P in NTT form (u is transformed on the way).
//   BFV:  buf = P itself (coefficient form after the in-place inverse transform).
// p_lpp != 0 (CKKS only): P holds p_lpp limbs per polynomial, polynomial (ct, k) at
// ct * p_ct_stride + k * p_lpp * N words -- a ciphertext with a limb above the L kept
// ones, divided and rounded by its last prime (a rescale; folded).
struct ModDownArgs {
    const uint64_t *S;
    const uint64_t *P;
    uint64_t *ct;
    size_t ct_stride;
    const DivRoundConst *dc;
    uint64_t p;
    uint32_t L;
    bool folded;
    size_t p_ct_stride = 0;
    uint32_t p_lpp = 0;
};
bool ntt_moddown(bool ckks, uint64_t *buf, const ModDownArgs &args, uint32_t count, const DevPrime *primes, int logn,
                 hipStream_t stream);
// CKKS form over `polys` polynomials (limb (i, j), j = 2 ct + k, may be odd): the
// rescale of polys contiguous polynomials is one call with ct_stride = 2 L N and
// p_ct_stride = 2 p_lpp N.
bool ntt_moddown_polys(uint64_t *U, const ModDownArgs &args, uint32_t polys, const DevPrime *primes, int logn,
                       hipStream_t stream);

// CKKS public-key encryption of m ciphertexts fused into the forward transform of its
// error term (N >= 2^13; false and nothing done otherwise).  Polynomial j = 2g + k of
// ciphertext g: the columns pass computes x = e_k - r_k (mod q_i) on load -- r_k the
// divide-and-round correction of the dropped prime p from Pl[g][k] + e_k (mod p), when
// drop -- and the rows pass stores out[g][k][i] = (pk_k[i] U[g][i] + NTT_i(x)) (times
// p^-1 when drop) (+ plain[g][i] for k = 0).  Es: [m][2][N] signed errors; Pl: [m][2][N]
// coefficient form; U: [m][L1][N] NTT(u); pk: [2][K][N]; out: [m][2][L][N].
struct EncArgs {
    const int64_t *Es;
    const uint64_t *Pl;
    const uint64_t *pk;
    const uint64_t *U;
    const uint64_t *plain;
    uint64_t *out;
    const DivRoundConst *dc;
    uint64_t p;
    uint32_t K, L1, L;
    bool drop;
};
bool ntt_encrypt_ckks(const EncArgs &args, uint32_t m, const DevPrime *primes, int logn, hipStream_t stream);

// ---- grouped launches (ciphertexts at several levels, one launch per stage) ----------
// At most MAX_GROUPS groups per call; every group is one level's batch with its own
// buffers.  A grouped call gives each group the words of the one-level call.
constexpr uint32_t MAX_GROUPS = 16;

// A transform group: limbs of data (grid lg, primes[i] for limb (i, j)); src != nullptr:
// the first pass reads the limbs from src (grid slg, same shape) instead, so the
// transform lands in data and src is left as it was.
struct NttGroup {
    uint64_t *data;
    const uint64_t *src;
    LimbGrid lg, slg;
    const DevPrime *primes;
};
void ntt_forward_groups(const NttGroup *groups, uint32_t n, int logn, NttOut out, hipStream_t stream);
void ntt_inverse_groups(const NttGroup *groups, uint32_t n, int logn, NttOut out, hipStream_t stream);
// The same over groups of several ring degrees, group g at N = 2^logns[g] (13 .. 17): one
// launch per pass for all of them.
void ntt_forward_multi(const NttGroup *groups, const int *logns, uint32_t n, NttOut out, hipStream_t stream);
void ntt_inverse_multi(const NttGroup *groups, const int *logns, uint32_t n, NttOut out, hipStream_t stream);

// A key-switch group for ntt_modup_keyprod_groups (the arguments of ntt_modup_keyprod).
struct KsGroup {
    const uint64_t *d;
    size_t d_ct_stride;
    uint64_t *T, *P, *S;
    const uint64_t *diag;
    size_t diag_stride;
    const DevPrime *targets;
    uint32_t count, L;
    KsFold fold;
};
// The fused ModUp + key products of every group (N = 2^15, 2^16; false otherwise): one
// columns-pass launch and one rows-pass launch per ciphertext grouping (4, 2, 1).
bool ntt_modup_keyprod_groups(const KsGroup *groups, uint32_t n, const uint64_t *key, uint32_t key_count,
                              const DevPrime *digits, int logn, hipStream_t stream);

// ModDown groups (CKKS, N >= 2^13; false otherwise): buf = U scratch of the group.
struct ModDownGroup {
    uint64_t *U;
    ModDownArgs a;
    uint32_t count;
};
bool ntt_moddown_groups(const ModDownGroup *groups, uint32_t n, const DevPrime *primes, int logn, hipStream_t stream);

}  // namespace shc
