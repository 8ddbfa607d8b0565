"""ctypes binding of the C ABI in include/sealhip.h (libsealhip.so, built in-tree).

Device buffers are passed as raw device pointers; callers typically hold them in
torch tensors (torch is only the allocator and stream provider here).  Every call
raises SealHipError with the library's message when the status is not SHC_OK, and
loading fails loudly when the HIP library has not been built.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsealhip.so")

SCHEME_BFV, SCHEME_CKKS, SCHEME_BGV = 1, 2, 3

_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p
_lib = None


class SealHipError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"sealhip status {status}: {msg}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH}; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.shc_last_error.restype = ctypes.c_char_p
        L.shc_init.argtypes = [ctypes.c_int]
        L.shc_context_create.argtypes = [ctypes.c_int, ctypes.c_uint32, _u64p, ctypes.c_uint32, ctypes.c_uint64,
                                         ctypes.POINTER(_vp)]
        L.shc_context_destroy.argtypes = [_vp]
        L.shc_context_destroy.restype = None
        L.shc_coeff_modulus_create.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_int), ctypes.c_uint32,
                                               ctypes.c_uint64, _u64p]
        for name in ("shc_ntt_forward", "shc_ntt_inverse"):
            getattr(L, name).argtypes = [_vp, _vp, ctypes.c_uint32, ctypes.c_uint32, _vp]
        for name in ("shc_ntt_forward_multi", "shc_ntt_inverse_multi"):
            getattr(L, name).argtypes = [ctypes.POINTER(NttJob), ctypes.c_uint32, _vp]
        u32, sz = ctypes.c_uint32, ctypes.c_size_t
        L.shc_ckks_multiply.argtypes = [_vp, _vp, _vp, _vp, u32, u32, _vp]
        L.shc_multiply_workspace_size.argtypes = [_vp, u32, u32, ctypes.POINTER(sz)]
        L.shc_multiply.argtypes = [_vp, _vp, _vp, _vp, u32, u32, _vp, sz, _vp]
        L.shc_keyswitch_workspace_size.argtypes = [_vp, u32, u32, ctypes.POINTER(sz)]
        L.shc_switch_key.argtypes = [_vp, _vp, sz, _vp, sz, _vp, u32, u32, _vp, sz, _vp]
        L.shc_relinearize.argtypes = [_vp, _vp, _vp, u32, u32, _vp, sz, _vp]
        L.shc_relinearize_noalloc.argtypes = [_vp, _vp, _vp, u32, u32, _vp]
        L.shc_galois_elt_from_step.argtypes = [u32, ctypes.c_int, ctypes.POINTER(u32)]
        L.shc_apply_galois.argtypes = [_vp, _vp, _vp, u32, u32, u32, ctypes.c_int, _vp]
        L.shc_galois_workspace_size.argtypes = [_vp, u32, u32, ctypes.POINTER(sz)]
        L.shc_apply_galois_keyswitch.argtypes = [_vp, _vp, _vp, u32, u32, u32, _vp, sz, _vp]
        u64 = ctypes.c_uint64
        L.shc_sample_poly.argtypes = [_vp, ctypes.c_int, u64, u32, _vp, _vp]
        L.shc_context_set_sampler_key.argtypes = [_vp, ctypes.c_char_p]
        L.shc_keygen_secret.argtypes = [_vp, u64, _vp, _vp]
        L.shc_keygen_public.argtypes = [_vp, _vp, u64, _vp, _vp]
        L.shc_keygen_workspace_size.argtypes = [_vp, ctypes.POINTER(sz)]
        L.shc_keygen_relin.argtypes = [_vp, _vp, u64, _vp, _vp, sz, _vp]
        L.shc_keygen_galois.argtypes = [_vp, _vp, u32, u64, _vp, _vp, sz, _vp]
        L.shc_encrypt_workspace_size.argtypes = [_vp, u32, ctypes.POINTER(sz)]
        L.shc_encrypt_zero_asymmetric.argtypes = [_vp, _vp, _vp, _vp, _vp, u32, _vp, _vp, sz, _vp]
        L.shc_encrypt.argtypes = [_vp, _vp, _vp, u32, u64, _vp, _vp, sz, _vp]
        L.shc_encrypt_zero_symmetric.argtypes = [_vp, _vp, u32, u64, _vp, _vp]
        L.shc_encrypt_many_workspace_size.argtypes = [_vp, u32, u32, ctypes.POINTER(sz)]
        L.shc_encrypt_many.argtypes = [_vp, _vp, _vp, u32, u32, u64, _vp, _vp, sz, _vp]
        L.shc_decrypt_many_workspace_size.argtypes = [_vp, u32, u32, ctypes.POINTER(sz)]
        L.shc_decrypt_many.argtypes = [_vp, _vp, u32, u32, _vp, u32, _vp, _vp, sz, _vp]
        L.shc_decrypt_workspace_size.argtypes = [_vp, u32, ctypes.POINTER(sz)]
        L.shc_decrypt.argtypes = [_vp, _vp, u32, _vp, u32, _vp, _vp, sz, _vp]
        L.shc_batch_encode.argtypes = [_vp, _vp, u32, _vp, _vp]
        L.shc_batch_decode.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.shc_batch_encode_many.argtypes = [_vp, _vp, u32, sz, u32, _vp, _vp]
        L.shc_batch_decode_many.argtypes = [_vp, _vp, u32, _vp, _vp, _vp]
        L.shc_ckks_encode_workspace_size.argtypes = [_vp, ctypes.POINTER(sz)]
        L.shc_ckks_encode.argtypes = [_vp, _vp, u32, ctypes.c_double, u32, _vp, _vp, sz, _vp]
        L.shc_ckks_encode_many_workspace_size.argtypes = [_vp, u32, ctypes.POINTER(sz)]
        L.shc_ckks_encode_many.argtypes = [_vp, _vp, u32, sz, u32, ctypes.c_double, u32, _vp, _vp, _vp, sz, _vp]
        L.shc_ckks_decode_many_workspace_size.argtypes = [_vp, u32, u32, ctypes.POINTER(sz)]
        L.shc_ckks_decode_many.argtypes = [_vp, _vp, u32, u32, ctypes.c_double, _vp, _vp, sz, _vp]
        L.shc_ckks_encode_deferred.argtypes = [_vp, _vp, u32, ctypes.c_double, u32, _vp, _vp, _vp, sz, _vp]
        L.shc_ckks_decode_workspace_size.argtypes = [_vp, u32, ctypes.POINTER(sz)]
        L.shc_ckks_decode.argtypes = [_vp, _vp, u32, ctypes.c_double, _vp, _vp, sz, _vp]
        L.shc_ckks_encode_polynomial.argtypes = [_vp, _vp, u32, ctypes.c_double, u32, _vp, _vp, sz, _vp]
        L.shc_ckks_decode_polynomial.argtypes = [_vp, _vp, u32, ctypes.c_double, _vp, _vp, sz, _vp]
        L.shc_rescale_workspace_size.argtypes = [_vp, u32, u32, u32, ctypes.POINTER(sz)]
        L.shc_rescale_to_next.argtypes = [_vp, _vp, _vp, u32, u32, u32, _vp, sz, _vp]
        L.shc_probe_keyswitch.argtypes = [_vp, _vp, _vp]
        for name in ("shc_add", "shc_sub"):
            getattr(L, name).argtypes = [_vp, _vp, _vp, _vp, u32, u32, _vp]
        L.shc_negate.argtypes = [_vp, _vp, _vp, u32, u32, _vp]
        L.shc_multiply_plain_ntt.argtypes = [_vp, _vp, _vp, _vp, u32, u32, _vp]
        L.shc_add_plain.argtypes = [_vp, _vp, _vp, _vp, u32, u32, u32, ctypes.c_int, _vp]
        L.shc_bfv_add_plain.argtypes = [_vp, _vp, _vp, _vp, u32, u32, u32, ctypes.c_int, _vp]
        L.shc_bfv_multiply_plain_workspace_size.argtypes = [_vp, u32, ctypes.POINTER(sz)]
        L.shc_bfv_multiply_plain.argtypes = [_vp, _vp, _vp, _vp, u32, u32, _vp, sz, _vp]
        L.shc_bfv_plain_to_ntt.argtypes = [_vp, _vp, u32, _vp, _vp]
        L.shc_hoisted_key_size.argtypes = [_vp, u32, ctypes.POINTER(sz)]
        L.shc_hoisted_key_workspace_size.argtypes = [_vp, u32, ctypes.POINTER(sz)]
        L.shc_hoisted_key_prepare.argtypes = [_vp, _vp, u32, u32, _vp, _vp, sz, _vp]
        L.shc_rotate_sum_workspace_size.argtypes = [_vp, u32, u32, ctypes.POINTER(sz)]
        L.shc_rotate_sum.argtypes = [_vp, _vp, u32, u32, ctypes.POINTER(u32), ctypes.POINTER(_vp),
                                     ctypes.POINTER(_vp), u32, _vp, _vp, sz, _vp]
        L.shc_rotate_many.argtypes = L.shc_rotate_sum.argtypes
        L.shc_multiply_plain_accumulate.argtypes = [_vp, _vp, _vp, u32, u32, u32, _vp, _vp]
        L.shc_multiply_plain_accumulate_many.argtypes = [_vp, _vp, _vp, u32, u32, u32, u32, _vp, _vp]
        L.shc_autotune_relinearize.argtypes = [_vp, _vp, _vp, u32, u32, _vp, sz, ctypes.c_int, _vp,
                                               ctypes.POINTER(KsShape), ctypes.POINTER(ctypes.c_float),
                                               ctypes.POINTER(ctypes.c_float)]
        L.shc_keyswitch_shape_set.argtypes = [_vp, u32, u32, ctypes.POINTER(KsShape)]
        L.shc_keyswitch_shape_get.argtypes = [_vp, u32, u32, ctypes.POINTER(KsShape)]
        L.shc_keyswitch_shapes_export.argtypes = [_vp, ctypes.c_char_p, sz, ctypes.POINTER(sz)]
        L.shc_keyswitch_shapes_import.argtypes = [_vp, ctypes.c_char_p]
        gp = ctypes.POINTER(CtGroup)
        L.shc_ckks_multiply_groups.argtypes = [_vp, gp, gp, gp, u32, _vp]
        L.shc_relinearize_groups_workspace_size.argtypes = [_vp, gp, u32, ctypes.POINTER(sz)]
        L.shc_relinearize_groups.argtypes = [_vp, gp, u32, _vp, _vp, sz, _vp]
        L.shc_rescale_groups_workspace_size.argtypes = [_vp, gp, u32, ctypes.POINTER(sz)]
        L.shc_rescale_to_next_groups.argtypes = [_vp, gp, u32, gp, u32, _vp, sz, _vp]
        L.shc_server_create.argtypes = [ctypes.c_int, sz, sz, u32, u32, ctypes.POINTER(_vp)]
        L.shc_server_destroy.argtypes = [_vp]
        L.shc_server_destroy.restype = None
        L.shc_server_add_matvec.argtypes = [_vp, _vp, u32, u32, u32, _vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                            ctypes.POINTER(_vp), ctypes.POINTER(u32)]
        L.shc_server_add_mulrelin.argtypes = [_vp, _vp, u32, _vp, ctypes.c_int, ctypes.POINTER(u32)]
        L.shc_server_request_size.argtypes = [_vp, u32, ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(sz)]
        L.shc_server_submit.argtypes = [_vp, u32, _vp, _vp, ctypes.POINTER(u64)]
        L.shc_server_wait.argtypes = [_vp, u64, ctypes.c_double, ctypes.POINTER(RequestInfo)]
        L.shc_server_host_alloc.argtypes = [_vp, sz, ctypes.POINTER(_vp)]
        L.shc_server_host_free.argtypes = [_vp, _vp]
        L.shc_server_clock.argtypes = [_vp]
        L.shc_server_clock.restype = ctypes.c_double
        L.shc_server_get_stats.argtypes = [_vp, ctypes.POINTER(ServerStats)]
        _lib = L
    return _lib


class CtGroup(ctypes.Structure):
    _fields_ = [("d_ct", ctypes.c_void_p), ("count", ctypes.c_uint32), ("limb_count", ctypes.c_uint32)]


def groups(items):
    """A ctypes array of CtGroup from (device pointer, count, limb_count) triples."""
    arr = (CtGroup * max(len(items), 1))()
    for k, (p, c, l) in enumerate(items):
        arr[k] = CtGroup(p, c, l)
    return arr


class NttJob(ctypes.Structure):
    """shc_ntt_job: poly_count polynomials of limb_count limbs at d_data, in ctx's ring."""
    _fields_ = [("ctx", _vp), ("d_data", _vp), ("poly_count", ctypes.c_uint32), ("limb_count", ctypes.c_uint32)]


def _ntt_multi(fn, jobs, stream):
    arr = (NttJob * max(1, len(jobs)))(*[NttJob(c.handle, d, p, l) for c, d, p, l in jobs])
    check(fn(arr, len(jobs), stream))


def ntt_forward_multi(jobs, stream=None):
    """Forward transforms of (Context, d_ptr, polys, limbs) jobs of mixed ring degrees, one
    launch per pass for those at N >= 2^13."""
    _ntt_multi(lib().shc_ntt_forward_multi, jobs, stream)


def ntt_inverse_multi(jobs, stream=None):
    _ntt_multi(lib().shc_ntt_inverse_multi, jobs, stream)


class KsShape(ctypes.Structure):
    _fields_ = [("cols_targets", ctypes.c_uint32), ("rows_group", ctypes.c_uint32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("batch_size", ctypes.c_uint32), ("lane", ctypes.c_uint32), ("batch_id", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("max_batches_in_flight", ctypes.c_uint32),
                ("peak_device_bytes", ctypes.c_size_t)]


SHC_TIMEOUT = 1


def check(status):
    if status != 0:
        raise SealHipError(status, lib().shc_last_error().decode())


class HipEvent:
    """A timing hipEvent_t from the HIP runtime the library is linked against (for
    shc_probe_keyswitch, which records on the library's launch stream)."""
    _rt = None

    def __init__(self):
        if HipEvent._rt is None:
            lib()
            rt = ctypes.CDLL("libamdhip64.so")
            rt.hipEventCreate.argtypes = [ctypes.POINTER(_vp)]
            rt.hipEventDestroy.argtypes = [_vp]
            rt.hipEventSynchronize.argtypes = [_vp]
            rt.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), _vp, _vp]
            HipEvent._rt = rt
        h = _vp()
        if HipEvent._rt.hipEventCreate(ctypes.byref(h)) != 0:
            raise SealHipError(-5, "hipEventCreate failed")
        self.handle = h.value

    def elapsed_ms(self, end):
        """Milliseconds from this event to `end` (waits for `end`)."""
        if HipEvent._rt.hipEventSynchronize(end.handle) != 0:
            raise SealHipError(-5, "hipEventSynchronize failed")
        ms = ctypes.c_float()
        if HipEvent._rt.hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle) != 0:
            raise SealHipError(-5, "hipEventElapsedTime failed")
        return ms.value

    def __del__(self):
        if getattr(self, "handle", None) and HipEvent._rt is not None:
            HipEvent._rt.hipEventDestroy(self.handle)


def init(device=0):
    check(lib().shc_init(device))


def coeff_modulus_create(n, bits, plain_modulus=0):
    out = (ctypes.c_uint64 * len(bits))()
    check(lib().shc_coeff_modulus_create(n, (ctypes.c_int * len(bits))(*bits), len(bits), plain_modulus, out))
    return list(out)


class Context:
    """Device-resident tables for one parameter set (SEALContextCuda analogue)."""

    def __init__(self, scheme, n, coeff_modulus, plain_modulus=0):
        self.scheme, self.n = scheme, n
        self.moduli = [int(q) for q in coeff_modulus]
        self.plain_modulus = plain_modulus
        arr = (ctypes.c_uint64 * len(self.moduli))(*self.moduli)
        h = _vp()
        check(lib().shc_context_create(scheme, n, arr, len(self.moduli), plain_modulus, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().shc_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ntt_forward(self, d_ptr, polys, limbs, stream=None):
        check(lib().shc_ntt_forward(self.handle, d_ptr, polys, limbs, stream))

    def ntt_inverse(self, d_ptr, polys, limbs, stream=None):
        check(lib().shc_ntt_inverse(self.handle, d_ptr, polys, limbs, stream))

    def ckks_multiply(self, d_a, d_b, d_out, count, limbs, stream=None):
        check(lib().shc_ckks_multiply(self.handle, d_a, d_b, d_out, count, limbs, stream))

    def multiply_workspace_size(self, count, limbs):
        b = ctypes.c_size_t()
        check(lib().shc_multiply_workspace_size(self.handle, count, limbs, ctypes.byref(b)))
        return b.value

    def multiply(self, d_a, d_b, d_out, count, limbs, d_ws, ws_bytes, stream=None):
        check(lib().shc_multiply(self.handle, d_a, d_b, d_out, count, limbs, d_ws, ws_bytes, stream))

    def keyswitch_workspace_size(self, count, limbs):
        b = ctypes.c_size_t()
        check(lib().shc_keyswitch_workspace_size(self.handle, count, limbs, ctypes.byref(b)))
        return b.value

    def switch_key(self, d_target, target_stride, d_ct, ct_stride, d_key, count, limbs, d_ws, ws_bytes, stream=None):
        check(lib().shc_switch_key(self.handle, d_target, target_stride, d_ct, ct_stride, d_key, count, limbs, d_ws,
                                   ws_bytes, stream))

    def relinearize(self, d_ct, d_key, count, limbs, d_ws, ws_bytes, stream=None):
        check(lib().shc_relinearize(self.handle, d_ct, d_key, count, limbs, d_ws, ws_bytes, stream))

    def relinearize_noalloc(self, d_ct, d_key, count, limbs, stream=None):
        """shc_relinearize with the library's fixed device scratch (no workspace)."""
        check(lib().shc_relinearize_noalloc(self.handle, d_ct, d_key, count, limbs, stream))

    def galois_elt_from_step(self, step):
        e = ctypes.c_uint32()
        check(lib().shc_galois_elt_from_step(self.n, step, ctypes.byref(e)))
        return e.value

    def apply_galois(self, d_in, d_out, polys, limbs, elt, ntt_form, stream=None):
        check(lib().shc_apply_galois(self.handle, d_in, d_out, polys, limbs, elt, 1 if ntt_form else 0, stream))

    def galois_workspace_size(self, count, limbs):
        b = ctypes.c_size_t()
        check(lib().shc_galois_workspace_size(self.handle, count, limbs, ctypes.byref(b)))
        return b.value

    def apply_galois_keyswitch(self, d_ct, d_key, elt, count, limbs, d_ws, ws_bytes, stream=None):
        check(lib().shc_apply_galois_keyswitch(self.handle, d_ct, d_key, elt, count, limbs, d_ws, ws_bytes, stream))

    # ---- keys, encryption, decryption, batch encoding
    def _size(self, fn, *args):
        b = ctypes.c_size_t()
        check(fn(self.handle, *args, ctypes.byref(b)))
        return b.value

    def set_sampler_key(self, key):
        """32-byte ChaCha20 key for every sampler stream of this context (None: splitmix64)."""
        if key is not None and len(key) != 32:
            raise ValueError("sampler key must be 32 bytes")
        check(lib().shc_context_set_sampler_key(self.handle, None if key is None else bytes(key)))

    def sample_poly(self, kind, seed, limbs, d_out, stream=None):
        check(lib().shc_sample_poly(self.handle, kind, seed, limbs, d_out, stream))

    def keygen_secret(self, seed, d_sk, stream=None):
        check(lib().shc_keygen_secret(self.handle, seed, d_sk, stream))

    def keygen_public(self, d_sk, seed, d_pk, stream=None):
        check(lib().shc_keygen_public(self.handle, d_sk, seed, d_pk, stream))

    def keygen_workspace_size(self):
        return self._size(lib().shc_keygen_workspace_size)

    def keygen_relin(self, d_sk, seed, d_key, d_ws, ws_bytes, stream=None):
        check(lib().shc_keygen_relin(self.handle, d_sk, seed, d_key, d_ws, ws_bytes, stream))

    def keygen_galois(self, d_sk, elt, seed, d_key, d_ws, ws_bytes, stream=None):
        check(lib().shc_keygen_galois(self.handle, d_sk, elt, seed, d_key, d_ws, ws_bytes, stream))

    def encrypt_workspace_size(self, limbs):
        return self._size(lib().shc_encrypt_workspace_size, limbs)

    def encrypt_zero_asymmetric(self, d_pk, d_u, d_e0, d_e1, limbs, d_out, d_ws, ws_bytes, stream=None):
        check(lib().shc_encrypt_zero_asymmetric(self.handle, d_pk, d_u, d_e0, d_e1, limbs, d_out, d_ws, ws_bytes,
                                                stream))

    def encrypt(self, d_pk, d_plain, limbs, seed, d_out, d_ws, ws_bytes, stream=None):
        check(lib().shc_encrypt(self.handle, d_pk, d_plain, limbs, seed, d_out, d_ws, ws_bytes, stream))

    def encrypt_zero_symmetric(self, d_sk, limbs, seed, d_out, stream=None):
        check(lib().shc_encrypt_zero_symmetric(self.handle, d_sk, limbs, seed, d_out, stream))

    def encrypt_many_workspace_size(self, limbs, count):
        return self._size(lib().shc_encrypt_many_workspace_size, limbs, count)

    def encrypt_many(self, d_pk, d_plain, count, limbs, seed, d_out, d_ws, ws_bytes, stream=None):
        """Ciphertext g = encrypt(seed + g) of plaintext g (d_plain None/0: zeros)."""
        check(lib().shc_encrypt_many(self.handle, d_pk, d_plain, count, limbs, seed, d_out, d_ws, ws_bytes, stream))

    def decrypt_many_workspace_size(self, limbs, count):
        return self._size(lib().shc_decrypt_many_workspace_size, limbs, count)

    def decrypt_many(self, d_ct, size, count, d_sk, limbs, d_out, d_ws, ws_bytes, stream=None):
        check(lib().shc_decrypt_many(self.handle, d_ct, size, count, d_sk, limbs, d_out, d_ws, ws_bytes, stream))

    def decrypt_workspace_size(self, limbs):
        return self._size(lib().shc_decrypt_workspace_size, limbs)

    def decrypt(self, d_ct, size, d_sk, limbs, d_out, d_ws, ws_bytes, stream=None):
        check(lib().shc_decrypt(self.handle, d_ct, size, d_sk, limbs, d_out, d_ws, ws_bytes, stream))

    def batch_encode(self, d_values, count, d_plain, stream=None):
        check(lib().shc_batch_encode(self.handle, d_values, count, d_plain, stream))

    def batch_decode(self, d_plain, d_values, d_scratch, stream=None):
        check(lib().shc_batch_decode(self.handle, d_plain, d_values, d_scratch, stream))

    def batch_encode_many(self, d_values, count, values_stride, plaintexts, d_plain, stream=None):
        check(lib().shc_batch_encode_many(self.handle, d_values, count, values_stride, plaintexts, d_plain, stream))

    def batch_decode_many(self, d_plain, plaintexts, d_values, d_scratch, stream=None):
        check(lib().shc_batch_decode_many(self.handle, d_plain, plaintexts, d_values, d_scratch, stream))

    def ckks_encode_workspace_size(self):
        return self._size(lib().shc_ckks_encode_workspace_size)

    def ckks_encode(self, d_values, count, scale, limbs, d_plain, d_ws, ws_bytes, stream=None):
        check(lib().shc_ckks_encode(self.handle, d_values, count, scale, limbs, d_plain, d_ws, ws_bytes, stream))

    def ckks_encode_deferred(self, d_values, count, scale, limbs, d_plain, d_status, d_ws, ws_bytes, stream=None):
        """Encode with the range check written to the int32 device word d_status."""
        check(lib().shc_ckks_encode_deferred(self.handle, d_values, count, scale, limbs, d_plain, d_status, d_ws,
                                             ws_bytes, stream))

    def ckks_encode_many_workspace_size(self, plaintexts):
        return self._size(lib().shc_ckks_encode_many_workspace_size, plaintexts)

    def ckks_encode_many(self, d_values, count, values_stride, plaintexts, scale, limbs, d_plain, d_status, d_ws,
                         ws_bytes, stream=None):
        check(lib().shc_ckks_encode_many(self.handle, d_values, count, values_stride, plaintexts, scale, limbs,
                                         d_plain, d_status, d_ws, ws_bytes, stream))

    def ckks_decode_many_workspace_size(self, limbs, plaintexts):
        return self._size(lib().shc_ckks_decode_many_workspace_size, limbs, plaintexts)

    def ckks_decode_many(self, d_plain, limbs, plaintexts, scale, d_values, d_ws, ws_bytes, stream=None):
        check(lib().shc_ckks_decode_many(self.handle, d_plain, limbs, plaintexts, scale, d_values, d_ws, ws_bytes,
                                         stream))

    def ckks_decode_workspace_size(self, limbs):
        return self._size(lib().shc_ckks_decode_workspace_size, limbs)

    def ckks_decode(self, d_plain, limbs, scale, d_values, d_ws, ws_bytes, stream=None):
        check(lib().shc_ckks_decode(self.handle, d_plain, limbs, scale, d_values, d_ws, ws_bytes, stream))

    def ckks_encode_polynomial(self, d_values, count, scale, limbs, d_plain, d_ws, ws_bytes, stream=None):
        """count <= N real coefficients times scale (no FFT); workspace as ckks_encode."""
        check(lib().shc_ckks_encode_polynomial(self.handle, d_values, count, scale, limbs, d_plain, d_ws, ws_bytes,
                                               stream))

    def ckks_decode_polynomial(self, d_plain, limbs, scale, d_values, d_ws, ws_bytes, stream=None):
        """The N real coefficients over scale; workspace as ckks_decode."""
        check(lib().shc_ckks_decode_polynomial(self.handle, d_plain, limbs, scale, d_values, d_ws, ws_bytes, stream))

    def rescale_workspace_size(self, count, polys, limbs):
        b = ctypes.c_size_t()
        check(lib().shc_rescale_workspace_size(self.handle, count, polys, limbs, ctypes.byref(b)))
        return b.value

    def rescale_to_next(self, d_in, d_out, count, polys, limbs, d_ws, ws_bytes, stream=None):
        check(lib().shc_rescale_to_next(self.handle, d_in, d_out, count, polys, limbs, d_ws, ws_bytes, stream))

    def add(self, a, b, out, polys, limbs, stream=None):
        check(lib().shc_add(self.handle, a, b, out, polys, limbs, stream))

    def sub(self, a, b, out, polys, limbs, stream=None):
        check(lib().shc_sub(self.handle, a, b, out, polys, limbs, stream))

    def negate(self, a, out, polys, limbs, stream=None):
        check(lib().shc_negate(self.handle, a, out, polys, limbs, stream))

    def multiply_plain_ntt(self, a, plain, out, polys, limbs, stream=None):
        check(lib().shc_multiply_plain_ntt(self.handle, a, plain, out, polys, limbs, stream))

    def add_plain(self, ct, plain, out, count, size, limbs, subtract=False, stream=None):
        check(lib().shc_add_plain(self.handle, ct, plain, out, count, size, limbs, int(subtract), stream))

    def bfv_add_plain(self, ct, plain, out, count, size, limbs, subtract=False, stream=None):
        check(lib().shc_bfv_add_plain(self.handle, ct, plain, out, count, size, limbs, int(subtract), stream))

    def bfv_multiply_plain_workspace_size(self, limbs):
        return self._size(lib().shc_bfv_multiply_plain_workspace_size, limbs)

    def bfv_multiply_plain(self, a, plain, out, polys, limbs, d_ws, ws_bytes, stream=None):
        check(lib().shc_bfv_multiply_plain(self.handle, a, plain, out, polys, limbs, d_ws, ws_bytes, stream))

    def bfv_plain_to_ntt(self, plain, limbs, out, stream=None):
        check(lib().shc_bfv_plain_to_ntt(self.handle, plain, limbs, out, stream))

    def hoisted_key_size(self, limbs):
        return self._size(lib().shc_hoisted_key_size, limbs)

    def hoisted_key_workspace_size(self, limbs):
        return self._size(lib().shc_hoisted_key_workspace_size, limbs)

    def hoisted_key_prepare(self, d_key, elt, limbs, d_out, d_ws, ws_bytes, stream=None):
        check(lib().shc_hoisted_key_prepare(self.handle, d_key, elt, limbs, d_out, d_ws, ws_bytes, stream))

    def rotate_sum_workspace_size(self, count, limbs):
        return self._size(lib().shc_rotate_sum_workspace_size, count, limbs)

    def rotate_sum(self, d_in, count, limbs, elts, d_hoisted, d_keys, d_out, d_ws, ws_bytes, stream=None):
        r = len(elts)
        e = (ctypes.c_uint32 * r)(*elts)
        h = (_vp * r)(*d_hoisted)
        k = (_vp * r)(*d_keys)
        check(lib().shc_rotate_sum(self.handle, d_in, count, limbs, e, h, k, r, d_out, d_ws, ws_bytes, stream))

    def rotate_many(self, d_in, count, limbs, elts, d_hoisted, d_keys, d_out, d_ws, ws_bytes, stream=None):
        r = len(elts)
        e = (ctypes.c_uint32 * r)(*elts)
        h = (_vp * r)(*d_hoisted)
        k = (_vp * r)(*d_keys)
        check(lib().shc_rotate_many(self.handle, d_in, count, limbs, e, h, k, r, d_out, d_ws, ws_bytes, stream))

    def multiply_plain_accumulate(self, d_cts, d_plains, terms, polys, limbs, d_out, stream=None):
        check(lib().shc_multiply_plain_accumulate(self.handle, d_cts, d_plains, terms, polys, limbs, d_out, stream))

    def multiply_plain_accumulate_many(self, d_cts, d_plains, terms, outputs, polys, limbs, d_out, stream=None):
        check(lib().shc_multiply_plain_accumulate_many(self.handle, d_cts, d_plains, terms, outputs, polys, limbs,
                                                       d_out, stream))

    def autotune_relinearize(self, d_ct, d_key, count, limbs, d_ws, ws_bytes, reps=2, stream=None):
        """Times relinearizations of d_ct (scratch) with every launch shape and keeps the
        fastest for (limbs, count); returns {"best": (cols_targets, rows_group),
        "best_ms": ..., "builtin_ms": ...}."""
        best, bms, dms = KsShape(), ctypes.c_float(), ctypes.c_float()
        check(lib().shc_autotune_relinearize(self.handle, d_ct, d_key, count, limbs, d_ws, ws_bytes, reps, stream,
                                             ctypes.byref(best), ctypes.byref(bms), ctypes.byref(dms)))
        return {"best": (best.cols_targets, best.rows_group), "best_ms": bms.value, "builtin_ms": dms.value}

    def keyswitch_shape(self, limbs, count, shape=None, clear=False):
        """The launch shape used for (limbs, count); sets it first when shape =
        (cols_targets, rows_group) is given, removes it with clear=True."""
        if clear:
            check(lib().shc_keyswitch_shape_set(self.handle, limbs, count, None))
        elif shape is not None:
            s = KsShape(*shape)
            check(lib().shc_keyswitch_shape_set(self.handle, limbs, count, ctypes.byref(s)))
        out = KsShape()
        check(lib().shc_keyswitch_shape_get(self.handle, limbs, count, ctypes.byref(out)))
        return (out.cols_targets, out.rows_group)

    def keyswitch_shapes_export(self):
        n = ctypes.c_size_t()
        check(lib().shc_keyswitch_shapes_export(self.handle, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        check(lib().shc_keyswitch_shapes_export(self.handle, buf, n.value, ctypes.byref(n)))
        return buf.value.decode()

    def keyswitch_shapes_import(self, text):
        check(lib().shc_keyswitch_shapes_import(self.handle, text.encode()))

    # ---- level groups: lists of (device pointer, count, limb_count)
    def ckks_multiply_groups(self, a, b, out, stream=None):
        check(lib().shc_ckks_multiply_groups(self.handle, groups(a), groups(b), groups(out), len(a), stream))

    def relinearize_groups_workspace_size(self, cts):
        n = ctypes.c_size_t()
        check(lib().shc_relinearize_groups_workspace_size(self.handle, groups(cts), len(cts), ctypes.byref(n)))
        return n.value

    def relinearize_groups(self, cts, d_key, d_ws, ws_bytes, stream=None):
        check(lib().shc_relinearize_groups(self.handle, groups(cts), len(cts), d_key, d_ws, ws_bytes, stream))

    def rescale_groups_workspace_size(self, cts):
        n = ctypes.c_size_t()
        check(lib().shc_rescale_groups_workspace_size(self.handle, groups(cts), len(cts), ctypes.byref(n)))
        return n.value

    def rescale_to_next_groups(self, cts, in_size, out, d_ws, ws_bytes, stream=None):
        check(lib().shc_rescale_to_next_groups(self.handle, groups(cts), in_size, groups(out), len(cts), d_ws,
                                               ws_bytes, stream))

    def probe_keyswitch(self, begin, end):
        """Arm the next key switch to record HipEvents around its ModUp rows pass."""
        check(lib().shc_probe_keyswitch(self.handle, begin.handle, end.handle))


class Server:
    """The library's continuous-batching request server (shc_server_*): requests with
    host numpy buffers (uint64), admitted by device-memory footprint into batches that
    run on `lanes` streams at once.  The arrays passed to submit() are kept alive here
    until the request is waited for."""

    def __init__(self, device_bytes, staging_bytes, lanes=2, max_batch=16, device=0):
        h = _vp()
        check(lib().shc_server_create(device, device_bytes, staging_bytes, lanes, max_batch, ctypes.byref(h)))
        self.handle = h
        self._held = {}
        self._keep = []

    def close(self):
        if self.handle:
            lib().shc_server_destroy(self.handle)
            self.handle = None
            self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_matvec(self, ctx, limbs, baby, giant, d_diagonals, baby_keys, baby_hoisted, giant_keys):
        t = ctypes.c_uint32()
        bk = (_vp * max(len(baby_keys), 1))(*baby_keys)
        bh = (_vp * max(len(baby_hoisted), 1))(*baby_hoisted)
        gk = (_vp * max(len(giant_keys), 1))(*giant_keys)
        check(lib().shc_server_add_matvec(self.handle, ctx.handle, limbs, baby, giant, d_diagonals, bk, bh, gk,
                                          ctypes.byref(t)))
        self._keep.append(ctx)
        return t.value

    def add_mulrelin(self, ctx, limbs, d_relin_key, rescale=False):
        t = ctypes.c_uint32()
        check(lib().shc_server_add_mulrelin(self.handle, ctx.handle, limbs, d_relin_key, int(rescale), ctypes.byref(t)))
        self._keep.append(ctx)
        return t.value

    def request_size(self, type_id):
        """(input words, output words, device bytes of a batch of one)."""
        i, o, d = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().shc_server_request_size(self.handle, type_id, ctypes.byref(i), ctypes.byref(o), ctypes.byref(d)))
        return i.value, o.value, d.value

    def submit(self, type_id, h_in, h_out):
        """h_in, h_out: C-contiguous uint64 numpy arrays of the type's input / output words."""
        rid = ctypes.c_uint64()
        check(lib().shc_server_submit(self.handle, type_id, h_in.ctypes.data, h_out.ctypes.data, ctypes.byref(rid)))
        self._held[rid.value] = h_out
        return rid.value

    def wait(self, rid, timeout=-1.0):
        """The request's RequestInfo once done, or None when `timeout` seconds passed first."""
        info = RequestInfo()
        st = lib().shc_server_wait(self.handle, rid, timeout, ctypes.byref(info))
        if st == SHC_TIMEOUT:
            return None
        self._held.pop(rid, None)
        check(st)
        return info

    def host_array(self, words):
        """A uint64 numpy array of `words` in pinned memory from the server: requests
        whose buffers are such arrays are uploaded from / downloaded into them in place
        (an input must then stay unchanged until its request is done)."""
        import numpy as np
        p = _vp()
        check(lib().shc_server_host_alloc(self.handle, words * 8, ctypes.byref(p)))
        buf = (ctypes.c_uint64 * words).from_address(p.value)
        arr = np.frombuffer(buf, dtype=np.uint64)
        self._keep.append(buf)
        return arr

    def clock(self):
        return lib().shc_server_clock(self.handle)

    def stats(self):
        s = ServerStats()
        check(lib().shc_server_get_stats(self.handle, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}
