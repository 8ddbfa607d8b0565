"""GPU parity of the batched NTT / inverse NTT against the oracle.

Mirrors the reference's NTT tests (test/utils/ntt.cpp) and the batched-NTT workload:
N = 2^16, 24 primes of 60 bits = 1 mod 2^17, 128 ciphertexts x 2 polys x 24 limbs.
Bit-exact: every output word equals the oracle's fully reduced word.
"""
import numpy as np
import pytest

import checkers as ck
from gpu_util import from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu


def check_ntt(logn, bits, polys, seed):
    n = 1 << logn
    L = len(bits)
    moduli = ck.coeff_modulus_create(n, bits)
    ctx = hip().Context(hip().SCHEME_CKKS, n, moduli)
    rng = np.random.default_rng(seed)
    a = ck.uniform_residues(rng, (polys, L, n), moduli)
    d = to_dev(a)
    ctx.ntt_forward(d.data_ptr(), polys, L, stream())
    f = from_dev(d)
    np.testing.assert_array_equal(f, ck.ntt_rns(a, moduli, logn))
    ctx.ntt_inverse(d.data_ptr(), polys, L, stream())
    np.testing.assert_array_equal(from_dev(d), a)


@pytest.mark.parametrize("logn", [1, 2, 3, 4, 5, 6, 8, 10, 11, 12, 13, 14, 15, 16, 17])
def test_ntt_all_degrees(logn):
    bits = [60, 50, 40] if logn >= 8 else [30, 30]
    check_ntt(logn, bits, polys=3, seed=logn)


def test_ntt_kat():
    # test/utils/ntt.cpp:75-99 on the GPU path
    ctx = hip().Context(hip().SCHEME_CKKS, 2, [0xFFFFFFFFFFC0001])
    d = to_dev(np.array([1, 1], dtype=np.uint64))
    ctx.ntt_forward(d.data_ptr(), 1, 1, stream())
    assert [int(x) for x in from_dev(d)] == [288794978602139553, 864126526004445282]


@pytest.mark.parametrize("logn", [10, 13, 16])
def test_ntt_edge_values(logn):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [60, 60])
    ctx = hip().Context(hip().SCHEME_CKKS, n, moduli)
    for fill in ("zero", "max"):
        a = np.zeros((1, 2, n), dtype=np.uint64)
        if fill == "max":
            for i, q in enumerate(moduli):
                a[0, i, :] = q - 1
        d = to_dev(a)
        ctx.ntt_forward(d.data_ptr(), 1, 2, stream())
        np.testing.assert_array_equal(from_dev(d), ck.ntt_rns(a, moduli, logn))
        ctx.ntt_inverse(d.data_ptr(), 1, 2, stream())
        np.testing.assert_array_equal(from_dev(d), a)


def test_ntt_rejects_bad_limb_count():
    ctx = hip().Context(hip().SCHEME_CKKS, 1024, ck.coeff_modulus_create(1024, [40, 40]))
    d = to_dev(np.zeros(3 * 1024, dtype=np.uint64))
    with pytest.raises(hip().SealHipError):
        ctx.ntt_forward(d.data_ptr(), 1, 3, stream())


def test_batched_ntt_config_shape():
    """N=2^16, 24 x 60-bit primes; words of 2 ciphertexts checked against the oracle,
    the full 128-ciphertext batch checked by the forward/inverse round trip and by
    the oracle on sampled limbs."""
    logn, n = 16, 1 << 16
    moduli = ck.coeff_modulus_create(n, [60] * 24)
    assert all(q % (1 << 17) == 1 for q in moduli)
    ctx = hip().Context(hip().SCHEME_CKKS, n, moduli)
    rng = np.random.default_rng(2)
    small = ck.uniform_residues(rng, (4, 24, n), moduli)  # 2 ct x 2 polys
    d = to_dev(small)
    ctx.ntt_forward(d.data_ptr(), 4, 24, stream())
    np.testing.assert_array_equal(from_dev(d), ck.ntt_rns(small, moduli, logn))

    import torch
    polys = 256  # 128 ciphertexts x 2 polys
    gen = torch.Generator(device="cuda").manual_seed(7)
    dev = torch.empty((polys, 24, n), dtype=torch.int64, device="cuda")
    for i, q in enumerate(moduli):
        dev[:, i, :] = torch.randint(0, q, (polys, n), generator=gen, device="cuda", dtype=torch.int64)
    orig = dev.clone()
    ctx.ntt_forward(dev.data_ptr(), polys, 24, stream())
    torch.cuda.synchronize()
    for p, limb in [(0, 0), (77, 13), (255, 23)]:
        x = orig[p, limb].cpu().numpy().view(np.uint64)
        y = dev[p, limb].cpu().numpy().view(np.uint64)
        np.testing.assert_array_equal(y, ck.ntt(x, moduli[limb], logn))
    ctx.ntt_inverse(dev.data_ptr(), polys, 24, stream())
    torch.cuda.synchronize()
    assert torch.equal(dev, orig)


def test_multi_degree_transforms_in_one_launch():
    """shc_ntt_forward_multi / shc_ntt_inverse_multi: jobs of five ring degrees (2^12 .. 2^17,
    the 2^12 job on its own launches) and different prime counts and batch sizes give every
    job the oracle's words, and the inverse returns the inputs."""
    specs = [(12, [50, 40], 3), (13, [60, 50, 40, 40], 2), (15, [60] * 5, 3), (16, [60, 50, 50], 2),
             (17, [60, 60], 1), (15, [40, 40], 1)]
    ctxs, inputs, devs, jobs = [], [], [], []
    rng = np.random.default_rng(77)
    for logn, bits, polys in specs:
        n = 1 << logn
        moduli = ck.coeff_modulus_create(n, bits)
        ctx = hip().Context(hip().SCHEME_CKKS, n, moduli)
        a = ck.uniform_residues(rng, (polys, len(bits), n), moduli)
        d = to_dev(a)
        ctxs.append((ctx, moduli, logn))
        inputs.append(a)
        devs.append(d)
        jobs.append((ctx, d.data_ptr(), polys, len(bits)))
    hip().ntt_forward_multi(jobs, stream())
    outs = [from_dev(d) for d in devs]
    for (ctx, moduli, logn), a, f in zip(ctxs, inputs, outs):
        np.testing.assert_array_equal(f, ck.ntt_rns(a, moduli, logn), err_msg=f"forward, N=2^{logn}")
    hip().ntt_inverse_multi(jobs, stream())
    for (ctx, moduli, logn), a, d in zip(ctxs, inputs, devs):
        np.testing.assert_array_equal(from_dev(d), a, err_msg=f"round trip, N=2^{logn}")
    # errors: a level beyond the context, too many jobs
    ctx, moduli, logn = ctxs[2]
    with pytest.raises(hip().SealHipError):
        hip().ntt_forward_multi([(ctx, devs[2].data_ptr(), 1, 9)], stream())
    with pytest.raises(hip().SealHipError, match="16"):
        hip().ntt_forward_multi([(ctx, devs[2].data_ptr(), 1, 1)] * 17, stream())
