"""ctypes access to the two CPU checkers: the oracle restatement (oracle/liboracle.so)
and, where it was built in this container, the reference CPU path itself
(oracle/_ref/libtroyref.so, built by oracle/ref.mk from the reference sources).

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libtroyref.so")

u64p = ctypes.POINTER(ctypes.c_uint64)
_oracle = None
_ref = None


def ptr(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(u64p)


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run __graft_entry__.build())")
        lib = ctypes.CDLL(ORACLE_SO)
        lib.or_mulmod.restype = ctypes.c_uint64
        lib.or_mulmod.argtypes = [ctypes.c_uint64] * 3
        lib.or_powmod.restype = ctypes.c_uint64
        lib.or_powmod.argtypes = [ctypes.c_uint64] * 3
        lib.or_invmod.restype = ctypes.c_uint64
        lib.or_invmod.argtypes = [ctypes.c_uint64] * 2
        lib.or_shoup.restype = ctypes.c_uint64
        lib.or_shoup.argtypes = [ctypes.c_uint64] * 2
        lib.or_is_prime.argtypes = [ctypes.c_uint64]
        lib.or_get_primes.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t, u64p]
        lib.or_coeff_modulus_create.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_int),
                                                ctypes.c_size_t, ctypes.c_uint64, u64p]
        lib.or_minimal_primitive_root.argtypes = [ctypes.c_uint64, ctypes.c_uint64, u64p]
        lib.or_ntt_batch.argtypes = [ctypes.c_int, ctypes.c_uint64, u64p, ctypes.c_size_t, ctypes.c_int]
        lib.or_ntt_tables_export.argtypes = [ctypes.c_int, ctypes.c_uint64] + [u64p] * 5
        lib.or_galois_elt_from_step.restype = ctypes.c_uint32
        lib.or_galois_elt_from_step.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.or_apply_galois.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, u64p, u64p]
        lib.or_apply_galois_ntt.argtypes = [ctypes.c_int, ctypes.c_uint32, u64p, u64p]
        sz = ctypes.c_size_t
        lib.or_ckks_multiply.argtypes = [ctypes.c_int, u64p, sz, u64p, sz, u64p, sz, u64p]
        lib.or_switch_key.argtypes = [ctypes.c_int, ctypes.c_int, u64p, sz, sz, u64p, u64p, u64p]
        lib.or_relinearize.argtypes = [ctypes.c_int, ctypes.c_int, u64p, sz, sz, u64p, u64p]
        lib.or_rescale_to_next.argtypes = [ctypes.c_int, ctypes.c_int, u64p, sz, u64p, sz, u64p]
        lib.or_apply_galois_keyswitch.argtypes = [ctypes.c_int, ctypes.c_int, u64p, sz, sz, u64p, ctypes.c_uint32,
                                                  u64p]
        lib.or_encrypt_zero_asymmetric.argtypes = [ctypes.c_int, ctypes.c_int, u64p, sz, u64p, sz, u64p, u64p,
                                                   u64p, u64p]
        lib.or_encrypt_zero_symmetric.argtypes = [ctypes.c_int, ctypes.c_int, u64p, sz, u64p, sz, u64p, u64p, u64p]
        lib.or_generate_kswitch_key.argtypes = [ctypes.c_int, u64p, sz, u64p, u64p, u64p, u64p, u64p]
        lib.or_decrypt_phase.argtypes = [ctypes.c_int, ctypes.c_int, u64p, sz, u64p, sz, u64p, sz, u64p]
        lib.or_bfv_decrypt_scale_round.argtypes = [ctypes.c_int, u64p, sz, ctypes.c_uint64, u64p, u64p]
        lib.or_bfv_add_plain_scaled.argtypes = [ctypes.c_int, u64p, sz, ctypes.c_uint64, u64p, sz, u64p]
        lib.or_bfv_sub_plain_scaled.argtypes = [ctypes.c_int, u64p, sz, ctypes.c_uint64, u64p, sz, u64p]
        lib.or_multiply_plain_ntt.argtypes = [ctypes.c_int, u64p, sz, u64p, sz, u64p, u64p]
        lib.or_bfv_multiply_plain.argtypes = [ctypes.c_int, u64p, sz, ctypes.c_uint64, u64p, sz, u64p, u64p]
        lib.or_batch_index_map.argtypes = [ctypes.c_int, u64p]
        lib.or_batch_encode.argtypes = [ctypes.c_int, ctypes.c_uint64, u64p, sz, u64p]
        lib.or_batch_decode.argtypes = [ctypes.c_int, ctypes.c_uint64, u64p, u64p]
        dpp = ctypes.POINTER(ctypes.c_double)
        lib.or_ckks_encode.argtypes = [ctypes.c_int, u64p, sz, dpp, sz, ctypes.c_double, u64p]
        lib.or_ckks_decode.argtypes = [ctypes.c_int, u64p, sz, u64p, ctypes.c_double, dpp]
        lib.or_behz_bases.argtypes = [ctypes.c_int, u64p, sz, ctypes.c_uint64, u64p, u64p, u64p,
                                      ctypes.POINTER(sz)]
        lib.or_bfv_multiply.argtypes = [ctypes.c_int, u64p, sz, ctypes.c_uint64, u64p, u64p, u64p]
        u32p = ctypes.POINTER(ctypes.c_uint32)
        lib.or_chacha20_block.argtypes = [u32p, ctypes.c_uint64, ctypes.c_uint64, u32p]
        lib.or_chacha20_block.restype = None
        lib.or_sample_poly.argtypes = [ctypes.c_int, u32p, ctypes.c_uint64, ctypes.c_int, u64p, sz, u64p]
        _oracle = lib
    return _oracle


def ref():
    """The reference CPU library, or None where it was not built (a machine without the reference sources)."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        lib = ctypes.CDLL(REF_SO)
        lib.ref_coeff_modulus_create.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.c_uint64, u64p]
        lib.ref_get_primes.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64, u64p]
        lib.ref_ntt_tables.argtypes = [ctypes.c_int, ctypes.c_uint64] + [u64p] * 5
        lib.ref_ntt.argtypes = [ctypes.c_int, ctypes.c_uint64, u64p, ctypes.c_uint64, ctypes.c_int]
        lib.ref_galois_elt_from_step.restype = ctypes.c_int64
        lib.ref_galois_elt_from_step.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.ref_apply_galois.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, u64p, u64p,
                                         ctypes.c_int]
        u64 = ctypes.c_uint64
        lib.ref_session_create.restype = ctypes.c_void_p
        lib.ref_session_create.argtypes = [ctypes.c_int, u64, u64p, u64, u64]
        lib.ref_session_destroy.argtypes = [ctypes.c_void_p]
        lib.ref_multiply.argtypes = [ctypes.c_void_p, u64, u64p, u64p, u64p]
        lib.ref_relinearize.argtypes = [ctypes.c_void_p, u64, u64p, u64p, u64p]
        lib.ref_rescale.argtypes = [ctypes.c_void_p, u64, u64, u64p, u64p]
        lib.ref_apply_galois_keyswitch.argtypes = [ctypes.c_void_p, u64, u64p, ctypes.c_uint32, u64p, u64p]
        dp = ctypes.POINTER(ctypes.c_double)
        lib.ref_decrypt.argtypes = [ctypes.c_void_p, u64, u64p, u64p, u64p]
        lib.ref_encrypt.argtypes = [ctypes.c_void_p, u64, u64p, u64p, u64p]
        lib.ref_keygen_relin.argtypes = [ctypes.c_void_p]
        lib.ref_batch_encode.argtypes = [ctypes.c_void_p, u64p, u64, u64p]
        lib.ref_batch_decode.argtypes = [ctypes.c_void_p, u64p, u64p]
        lib.ref_ckks_encode.argtypes = [ctypes.c_void_p, u64, dp, u64, ctypes.c_double, u64p]
        lib.ref_ckks_decode.argtypes = [ctypes.c_void_p, u64, u64p, ctypes.c_double, dp]
        lib.ref_ct_binary.argtypes = [ctypes.c_void_p, u64, u64, ctypes.c_int, u64p, u64p, u64p]
        lib.ref_plain_op.argtypes = [ctypes.c_void_p, u64, u64, ctypes.c_int, u64p, u64p, u64p]
        lib.ref_parms_id.argtypes = [ctypes.c_void_p, u64, u64p]
        lib.ref_plaintext_hex.argtypes = [ctypes.c_char_p, u64p, ctypes.POINTER(u64), ctypes.c_char_p, u64]
        _ref = lib
    return _ref


# ------------------------------------------------------------------ oracle helpers

def coeff_modulus_create(n, bits, plain_modulus=0):
    out = np.zeros(len(bits), dtype=np.uint64)
    b = (ctypes.c_int * len(bits))(*bits)
    rc = oracle().or_coeff_modulus_create(n, b, len(bits), plain_modulus, ptr(out))
    if rc:
        raise ValueError("coeff_modulus_create failed")
    return [int(x) for x in out]


def get_primes(factor, bit_size, count):
    out = np.zeros(count, dtype=np.uint64)
    if oracle().or_get_primes(factor, bit_size, count, ptr(out)):
        raise ValueError("not enough primes")
    return [int(x) for x in out]


def ntt(data, q, logn, inverse=False):
    """Fully reduced negacyclic NTT (or inverse) of every length-2^logn row of `data`."""
    a = np.ascontiguousarray(data, dtype=np.uint64).copy()
    n = 1 << logn
    assert a.size % n == 0
    if oracle().or_ntt_batch(logn, q, ptr(a), a.size // n, 1 if inverse else 0):
        raise ValueError("bad modulus for NTT")
    return a


def ntt_rns(data, moduli, logn, inverse=False):
    """NTT of a [..., limb, N] array where limb i is modulo moduli[i]."""
    a = np.ascontiguousarray(data, dtype=np.uint64).copy()
    L = len(moduli)
    n = 1 << logn
    v = a.reshape(-1, L, n)
    for i, q in enumerate(moduli):
        v[:, i, :] = ntt(v[:, i, :], q, logn, inverse)
    return a


def _arr(moduli):
    return np.ascontiguousarray(np.array(moduli, dtype=np.uint64))


def ckks_multiply(a, b, moduli, logn):
    """a, b: [2][L][N] -> [3][L][N]"""
    L = len(moduli)
    out = np.zeros((3, L, 1 << logn), dtype=np.uint64)
    m = _arr(moduli)
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    assert oracle().or_ckks_multiply(logn, ptr(m), L, ptr(a), 2, ptr(b), 2, ptr(out)) == 0
    return out


def bfv_multiply(a, b, moduli, t, logn):
    """a, b: [2][L][N] coefficient form -> [3][L][N] (BEHZ)"""
    L = len(moduli)
    out = np.zeros((3, L, 1 << logn), dtype=np.uint64)
    m = _arr(moduli)
    assert oracle().or_bfv_multiply(logn, ptr(m), L, t, ptr(np.ascontiguousarray(a, dtype=np.uint64)),
                                    ptr(np.ascontiguousarray(b, dtype=np.uint64)), ptr(out)) == 0
    return out


def behz_bases(moduli, t, logn):
    m = _arr(moduli)
    msk, gamma = np.zeros(1, dtype=np.uint64), np.zeros(1, dtype=np.uint64)
    B = np.zeros(70, dtype=np.uint64)
    bsz = ctypes.c_size_t()
    assert oracle().or_behz_bases(logn, ptr(m), len(moduli), t, ptr(msk), ptr(gamma), ptr(B),
                                  ctypes.byref(bsz)) == 0
    return int(msk[0]), int(gamma[0]), [int(x) for x in B[:bsz.value]]


def relinearize(scheme, ct3, key, key_moduli, L, logn):
    """ct3: [3][L][N] (CKKS NTT form / BFV coefficient form); key [K-1][2][K][N] -> [2][L][N]"""
    c = np.ascontiguousarray(ct3, dtype=np.uint64).copy()
    m = _arr(key_moduli)
    k = np.ascontiguousarray(key, dtype=np.uint64)
    assert oracle().or_relinearize(scheme, logn, ptr(m), len(key_moduli), L, ptr(c), ptr(k)) == 0
    return c[:2].copy()


def apply_galois_keyswitch(scheme, ct, elt, key, key_moduli, L, logn):
    """ct: [2][L][N] -> sigma_elt then key switch back (applyGaloisInplace)"""
    c = np.ascontiguousarray(ct, dtype=np.uint64).copy()
    m = _arr(key_moduli)
    assert oracle().or_apply_galois_keyswitch(scheme, logn, ptr(m), len(key_moduli), L, ptr(c), elt,
                                              ptr(np.ascontiguousarray(key, dtype=np.uint64))) == 0
    return c


def encrypt_zero_asymmetric(scheme, pk, u, e0, e1, moduli, logn):
    """pk: [2][K][N] NTT form; u, e0, e1: [L][N] coefficient form -> [2][L][N]"""
    L = len(moduli)
    out = np.zeros((2, L, 1 << logn), dtype=np.uint64)
    m = _arr(moduli)
    c = lambda x: ptr(np.ascontiguousarray(x, dtype=np.uint64))
    assert oracle().or_encrypt_zero_asymmetric(scheme, logn, ptr(m), L, c(pk), pk.shape[1], c(u), c(e0), c(e1),
                                               ptr(out)) == 0
    return out


def encrypt_zero_symmetric(scheme, sk, a, e, moduli, logn):
    L = len(moduli)
    out = np.zeros((2, L, 1 << logn), dtype=np.uint64)
    m = _arr(moduli)
    c = lambda x: ptr(np.ascontiguousarray(x, dtype=np.uint64))
    assert oracle().or_encrypt_zero_symmetric(scheme, logn, ptr(m), L, c(sk), sk.shape[0], c(a), c(e),
                                              ptr(out)) == 0
    return out


def generate_kswitch_key(sk, new_key, a, e, moduli, logn):
    K = len(moduli)
    out = np.zeros((K - 1, 2, K, 1 << logn), dtype=np.uint64)
    m = _arr(moduli)
    c = lambda x: ptr(np.ascontiguousarray(x, dtype=np.uint64))
    assert oracle().or_generate_kswitch_key(logn, ptr(m), K, c(sk), c(new_key), c(a), c(e), ptr(out)) == 0
    return out


def decrypt_phase(scheme, ct, sk, moduli, logn):
    size, L, n = ct.shape
    out = np.zeros((L, n), dtype=np.uint64)
    m = _arr(moduli[:L])
    c = lambda x: ptr(np.ascontiguousarray(x, dtype=np.uint64))
    assert oracle().or_decrypt_phase(scheme, logn, ptr(m), L, c(ct), size, c(sk), sk.shape[0], ptr(out)) == 0
    return out


def bfv_decrypt(ct, sk, moduli, t, logn):
    L = ct.shape[1]
    phase = decrypt_phase(1, ct, sk, moduli, logn)
    out = np.zeros(1 << logn, dtype=np.uint64)
    assert oracle().or_bfv_decrypt_scale_round(logn, ptr(_arr(moduli[:L])), L, t, ptr(phase), ptr(out)) == 0
    return out


def bfv_add_plain_scaled(c0, plain, moduli, t, logn):
    c = np.ascontiguousarray(c0, dtype=np.uint64).copy()
    L = c.shape[0]
    p = np.ascontiguousarray(plain, dtype=np.uint64)
    assert oracle().or_bfv_add_plain_scaled(logn, ptr(_arr(moduli[:L])), L, t, ptr(p), len(p), ptr(c)) == 0
    return c


def bfv_sub_plain_scaled(c0, plain, moduli, t, logn):
    c = np.ascontiguousarray(c0, dtype=np.uint64).copy()
    L = c.shape[0]
    p = np.ascontiguousarray(plain, dtype=np.uint64)
    assert oracle().or_bfv_sub_plain_scaled(logn, ptr(_arr(moduli[:L])), L, t, ptr(p), len(p), ptr(c)) == 0
    return c


def multiply_plain_ntt(a, plain, moduli, logn):
    """a: [polys][L][N] NTT form, plain [L][N] NTT form"""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    polys, L, n = a.shape
    out = np.zeros_like(a)
    assert oracle().or_multiply_plain_ntt(logn, ptr(_arr(moduli[:L])), L, ptr(a), polys,
                                          ptr(np.ascontiguousarray(plain, dtype=np.uint64)), ptr(out)) == 0
    return out


def bfv_multiply_plain(a, plain, moduli, t, logn):
    """a: [polys][L][N] coefficient form, plain [N] mod t"""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    polys, L, n = a.shape
    out = np.zeros_like(a)
    assert oracle().or_bfv_multiply_plain(logn, ptr(_arr(moduli[:L])), L, t, ptr(a), polys,
                                          ptr(np.ascontiguousarray(plain, dtype=np.uint64)), ptr(out)) == 0
    return out


def mod_add(a, b, moduli):
    """[..., L, N] + [..., L, N] mod moduli[l] (closed form)"""
    q = np.array(moduli[:a.shape[-2]], dtype=np.uint64)[:, None]
    s = a.astype(np.uint64) + b.astype(np.uint64)
    return np.where(s >= q, s - q, s)


def mod_sub(a, b, moduli):
    q = np.array(moduli[:a.shape[-2]], dtype=np.uint64)[:, None]
    a = a.astype(np.uint64)
    b = b.astype(np.uint64)
    return np.where(a >= b, a - b, a + q - b)


def mod_neg(a, moduli):
    q = np.array(moduli[:a.shape[-2]], dtype=np.uint64)[:, None]
    a = a.astype(np.uint64)
    return np.where(a == 0, a, q - a)


def batch_encode(values, t, logn):
    v = np.ascontiguousarray(values, dtype=np.uint64)
    out = np.zeros(1 << logn, dtype=np.uint64)
    assert oracle().or_batch_encode(logn, t, ptr(v), len(v), ptr(out)) == 0
    return out


def batch_decode(plain, t, logn):
    out = np.zeros(1 << logn, dtype=np.uint64)
    assert oracle().or_batch_decode(logn, t, ptr(np.ascontiguousarray(plain, dtype=np.uint64)), ptr(out)) == 0
    return out


def ckks_encode(values, moduli, scale, logn):
    """complex values (<= N/2) -> NTT-form plaintext [L][N]"""
    v = np.ascontiguousarray(np.stack([np.real(values), np.imag(values)], axis=1).reshape(-1), dtype=np.float64)
    out = np.zeros((len(moduli), 1 << logn), dtype=np.uint64)
    rc = oracle().or_ckks_encode(logn, ptr(_arr(moduli)), len(moduli), dptr(v), len(values), scale, ptr(out))
    assert rc == 0, rc
    return out


def ckks_decode(plain, moduli, scale, logn):
    out = np.zeros(1 << logn, dtype=np.float64)
    assert oracle().or_ckks_decode(logn, ptr(_arr(moduli)), len(moduli), ptr(np.ascontiguousarray(plain)), scale,
                                   dptr(out)) == 0
    return out[0::2] + 1j * out[1::2]


def rns_of_small(values, moduli):
    """[N] signed integers -> [L][N] residues (one integer reduced into every limb)."""
    v = np.asarray(values, dtype=object)
    return np.array([[int(x) % q for x in v] for q in moduli], dtype=np.uint64)


def chacha20_block(key, counter, nonce):
    """ChaCha20 block (16 words) for key = 32 bytes, 64-bit counter and nonce."""
    k = np.frombuffer(bytes(key), dtype="<u4").astype(np.uint32).copy()
    out = np.zeros(16, dtype=np.uint32)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    oracle().or_chacha20_block(k.ctypes.data_as(u32p), counter, nonce, out.ctypes.data_as(u32p))
    return out


def sample_poly(kind, key, seed, logn, moduli):
    """The library's sampler stream: key = 32 bytes (ChaCha20) or None (splitmix64)."""
    L = len(moduli)
    out = np.zeros((L, 1 << logn), dtype=np.uint64)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    k = None if key is None else np.frombuffer(bytes(key), dtype="<u4").astype(np.uint32).copy()
    kp = None if k is None else k.ctypes.data_as(u32p)
    assert oracle().or_sample_poly(kind, kp, seed, logn, ptr(_arr(moduli)), L, ptr(out)) == 0
    return out


def galois_elt(logn, step):
    return int(oracle().or_galois_elt_from_step(logn, step))


def rescale(scheme, ct, moduli, logn):
    """ct: [polys][L][N] with moduli[:L] -> [polys][L-1][N]"""
    ct = np.ascontiguousarray(ct, dtype=np.uint64)
    polys, L, n = ct.shape
    out = np.zeros((polys, L - 1, n), dtype=np.uint64)
    m = _arr(moduli[:L])
    assert oracle().or_rescale_to_next(scheme, logn, ptr(m), L, ptr(ct), polys, ptr(out)) == 0
    return out


class RefSession:
    """The reference's own Evaluator (oracle/_ref) on injected words."""

    def __init__(self, scheme, n, moduli, plain=0):
        m = _arr(moduli)
        self.h = ref().ref_session_create(scheme, n, ptr(m), len(moduli), plain)
        assert self.h, "reference rejected the parameters"
        self.n, self.K, self.scheme = n, len(moduli), scheme

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_session_destroy(self.h)

    def multiply(self, a, b, L):
        out = np.zeros((3, L, self.n), dtype=np.uint64)
        assert ref().ref_multiply(self.h, L, ptr(np.ascontiguousarray(a)), ptr(np.ascontiguousarray(b)), ptr(out)) == 0
        return out

    def relinearize(self, ct3, key, L):
        out = np.zeros((2, L, self.n), dtype=np.uint64)
        assert ref().ref_relinearize(self.h, L, ptr(np.ascontiguousarray(ct3)), ptr(np.ascontiguousarray(key)),
                                     ptr(out)) == 0
        return out

    def apply_galois(self, ct, elt, key, L):
        out = np.zeros((2, L, self.n), dtype=np.uint64)
        assert ref().ref_apply_galois_keyswitch(self.h, L, ptr(np.ascontiguousarray(ct)), elt,
                                                ptr(np.ascontiguousarray(key)), ptr(out)) == 0
        return out

    def decrypt(self, ct, sk, L):
        shape = (L, self.n) if self.scheme == 2 else (self.n,)
        out = np.zeros(shape, dtype=np.uint64)
        assert ref().ref_decrypt(self.h, L, ptr(np.ascontiguousarray(ct)), ptr(np.ascontiguousarray(sk)),
                                 ptr(out)) == 0
        return out

    def encrypt(self, pk, plain, L):
        """The reference's public-key encryption (its own randomness: for timing)."""
        out = np.zeros((2, L, self.n), dtype=np.uint64)
        assert ref().ref_encrypt(self.h, L, ptr(np.ascontiguousarray(pk)), ptr(np.ascontiguousarray(plain)),
                                 ptr(out)) == 0
        return out

    def keygen_relin(self):
        """A fresh KeyGenerator's relinearization key (its own randomness: for timing)."""
        assert ref().ref_keygen_relin(self.h) == 0

    def batch_encode(self, values):
        v = np.ascontiguousarray(values, dtype=np.uint64)
        out = np.zeros(self.n, dtype=np.uint64)
        assert ref().ref_batch_encode(self.h, ptr(v), len(v), ptr(out)) == 0
        return out

    def batch_decode(self, plain):
        out = np.zeros(self.n, dtype=np.uint64)
        assert ref().ref_batch_decode(self.h, ptr(np.ascontiguousarray(plain, dtype=np.uint64)), ptr(out)) == 0
        return out

    def ckks_encode(self, values, L, scale):
        v = np.ascontiguousarray(np.stack([np.real(values), np.imag(values)], axis=1).reshape(-1), dtype=np.float64)
        out = np.zeros((L, self.n), dtype=np.uint64)
        assert ref().ref_ckks_encode(self.h, L, dptr(v), len(values), scale, ptr(out)) == 0
        return out

    def ckks_decode(self, plain, L, scale):
        out = np.zeros(self.n, dtype=np.float64)
        assert ref().ref_ckks_decode(self.h, L, ptr(np.ascontiguousarray(plain, dtype=np.uint64)), scale,
                                     dptr(out)) == 0
        return out[0::2] + 1j * out[1::2]

    def ct_binary(self, op, a, b, L):
        """op 0 add, 1 sub, 2 negate (b ignored) on ciphertexts [size][L][N]"""
        a = np.ascontiguousarray(a, dtype=np.uint64)
        out = np.zeros_like(a)
        bb = a if b is None else np.ascontiguousarray(b, dtype=np.uint64)
        assert ref().ref_ct_binary(self.h, L, a.shape[0], op, ptr(a), ptr(bb), ptr(out)) == 0
        return out

    def plain_op(self, op, ct, plain, L):
        """op 0 addPlain, 1 subPlain, 2 multiplyPlain on ct [size][L][N]"""
        ct = np.ascontiguousarray(ct, dtype=np.uint64)
        out = np.zeros_like(ct)
        assert ref().ref_plain_op(self.h, L, ct.shape[0], op, ptr(ct), ptr(np.ascontiguousarray(plain, dtype=np.uint64)),
                                  ptr(out)) == 0
        return out

    def parms_id(self, L):
        out = np.zeros(4, dtype=np.uint64)
        assert ref().ref_parms_id(self.h, L, ptr(out)) == 0
        return [int(x) for x in out]

    def rescale(self, ct, L):
        size = ct.shape[0]
        out = np.zeros((size, L - 1, self.n), dtype=np.uint64)
        assert ref().ref_rescale(self.h, L, size, ptr(np.ascontiguousarray(ct)), ptr(out)) == 0
        return out


def parallel_map(fn, items, workers=None):
    """[fn(x) for x in items] on a thread pool.  The oracle keeps no mutable global
    state and ctypes releases the GIL during a call, so oracle calls on large shapes
    run side by side; at most 16 threads (the CPU share of one GPU's host)."""
    from concurrent.futures import ThreadPoolExecutor
    items = list(items)
    workers = workers or max(1, min(16, os.cpu_count() or 1, len(items)))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        return list(pool.map(fn, items))


def random_key(rng, moduli, logn):
    """Uniform key-switching key words [K-1][2][K][N] at key level (injected keys)."""
    K = len(moduli)
    return uniform_residues(rng, (K - 1, 2, K, 1 << logn), moduli)


# ------------------------------------------------------------------ seeded inputs

def splitmix64(seed):
    """splitmix64 stream; the same generator the bench uses for synthetic inputs."""
    state = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def uniform_residues(rng, shape, moduli):
    """Array of shape (..., len(moduli), N) with limb i uniform in [0, moduli[i])."""
    a = np.empty(shape, dtype=np.uint64)
    L = len(moduli)
    assert shape[-2] == L
    v = a.reshape(-1, L, shape[-1])
    for i, q in enumerate(moduli):
        v[:, i, :] = rng.integers(0, q, size=v[:, i, :].shape, dtype=np.uint64)
    return a
