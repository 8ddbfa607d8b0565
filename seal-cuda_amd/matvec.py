"""Diagonal matrix-vector products with baby-step giant-step rotations (Halevi-Shoup),
over the C ABI, for CKKS and BFV ciphertexts.

The reference's caller of this pattern is app/matmul.cuh:98-121 (one rotateVector and
one multiplyPlain per diagonal, summed).  Here, for a d x d matrix with d = b * g and
the vector replicated across each rotation cycle of the slots (CKKS: all N/2 slots;
BFV: each of the two rows of N/2 slots, which rotateRows turns independently):

    M v = sum_{h<g} rot_{h b}( sum_{j<b} diag'_{h,j} * rot_j(v) ),
    diag'_{h,j} = rot_{-h b}(diag_{h b + j}),  diag_k[i] = M[i mod d][(i + k) mod d],

where rot_k(x)[i] = x[i + k] (rotateVector / rotateRows(k)).  The b - 1 baby rotations
share one ModUp (shc_rotate_many), the g inner sums come from one
shc_multiply_plain_accumulate_many call over NTT-form operands (each baby rotation read
once per 8 giant steps), and the g - 1 giant rotations are ordinary key switches.  BFV
ciphertexts live in coefficient form: the b baby rotations go to NTT form once, and the
inner sums come back by one inverse transform before the giant rotations.  All
ciphertexts of a batch share the diagonals and keys.
"""
import numpy as np

import sealhip


def diagonals(M, slots, b, g, rows=1):
    """Pre-rotated diagonals diag'_{h,j} as slot vectors, [g][b][slots]: complex for CKKS
    (rows = 1), or M's dtype for BFV (rows = 2: each row of slots / 2 rotates on its own)."""
    d = M.shape[0]
    cycle = slots // rows
    assert M.shape == (d, d) and d == b * g and cycle % d == 0
    i = np.arange(cycle)
    out = np.empty((g, b, rows, cycle), dtype=np.complex128 if rows == 1 else M.dtype)
    for h in range(g):
        for j in range(b):
            k = h * b + j
            diag = M[i % d, (i + k) % d]
            out[h, j, :] = np.roll(diag, h * b)  # rot_{-hb}: x[i] <- diag[i - hb]
    return out.reshape(g, b, slots)


class BsgsMatvec:
    """One matrix, encoded once: `plains` is a device buffer [g][b][L][N] of the
    pre-rotated diagonals in NTT form (CKKS: encoded at the ciphertexts' level L; BFV:
    batch-encoded, then shc_bfv_plain_to_ntt).  baby_keys / baby_hoisted: Galois keys of
    steps 1..b-1 and their prepared forms at level L; giant_keys: Galois keys of steps
    b, 2b, ..., (g-1) b.  bfv: the ciphertexts are BFV (coefficient form)."""

    def __init__(self, ctx, n, L, b, g, plains, baby_keys, baby_hoisted, giant_keys, bfv=False):
        self.ctx, self.n, self.L, self.b, self.g = ctx, n, L, b, g
        self.bfv = bfv
        self.plains = plains
        self.baby_elts = [ctx.galois_elt_from_step(j) for j in range(1, b)]
        self.giant_elts = [ctx.galois_elt_from_step(h * b) for h in range(1, g)]
        self.baby_keys, self.baby_hoisted, self.giant_keys = baby_keys, baby_hoisted, giant_keys

    def workspace_sizes(self, count):
        ctx, L = self.ctx, self.L
        return max(ctx.rotate_sum_workspace_size(count, L), ctx.galois_workspace_size(count, L))

    def run(self, d_in, count, d_out, d_baby, d_inner, d_ws, ws_bytes, stream=None, around_sums=None):
        """d_in, d_out: [count][2][L][N]; d_baby: scratch [b][count][2][L][N];
        d_inner: scratch [g][count][2][L][N]; d_ws: workspace_sizes(count) bytes.
        around_sums: optional (before, after) callables invoked right before and after the
        inner-sum launch (e.g. recording timing events on `stream`)."""
        import ctypes
        ctx, n, L, b, g = self.ctx, self.n, self.L, self.b, self.g
        ct_bytes = count * 2 * L * n * 8
        plain_words = L * n
        # baby steps: rot_0 = v, rot_1..rot_{b-1} from one ModUp
        _memcpy_d2d(d_baby, d_in, ct_bytes, stream)
        if b > 1:
            ctx.rotate_many(d_in, count, L, self.baby_elts, self.baby_hoisted, self.baby_keys,
                            d_baby + ct_bytes, d_ws, ws_bytes, stream)
        if self.bfv:
            ctx.ntt_forward(d_baby, b * count * 2, L, stream)
        # every giant step's inner sum, inner[h] = sum_j diag'_{h,j} * rot_j(v)
        if around_sums:
            around_sums[0]()
        ctx.multiply_plain_accumulate_many(d_baby, self.plains, b, g, count * 2, L, d_inner, stream)
        if around_sums:
            around_sums[1]()
        if self.bfv:
            ctx.ntt_inverse(d_inner, g * count * 2, L, stream)
        # giant steps: rot_{h b}(inner[h]), summed
        for h in range(1, g):
            ctx.apply_galois_keyswitch(d_inner + h * ct_bytes, self.giant_keys[h - 1], self.giant_elts[h - 1], count,
                                       L, d_ws, ws_bytes, stream)
        if g == 1:
            _memcpy_d2d(d_out, d_inner, ct_bytes, stream)
        else:
            ctx.add(d_inner, d_inner + ct_bytes, d_out, count * 2, L, stream)
            for h in range(2, g):
                ctx.add(d_out, d_inner + h * ct_bytes, d_out, count * 2, L, stream)


_rt = None


def _memcpy_d2d(dst, src, nbytes, stream):
    import ctypes
    global _rt
    if _rt is None:
        sealhip.lib()
        _rt = ctypes.CDLL("libamdhip64.so")
        _rt.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                       ctypes.c_void_p]
    if _rt.hipMemcpyAsync(dst, src, nbytes, 3, stream) != 0:  # hipMemcpyDeviceToDevice
        raise sealhip.SealHipError(-5, "hipMemcpyAsync failed")
