// End-to-end test of the troy_cuda-style C++ API (include/troyhip.hpp) on the GPU.
//
// Mirrors the reference's evaluator tests (test/evaluator_cuda.cu): BFV and CKKS
// pipelines through KeyGenerator, Encryptor, Evaluator, Decryptor and the encoders,
// each result checked against the plaintext arithmetic it encrypts (exact mod t for
// BFV; within CKKS noise for CKKS), plus the exception types of invalid calls.
// Prints one "ok <name>" line per check and "facade ok" at the end; exits 1 on the
// first failure.
#include "troyhip.hpp"

#include <cstdio>
#include <cstdlib>
#include <random>
#include <sstream>

using namespace troyhip;

static int failures = 0;
#define EXPECT(cond, name)                                        \
    do {                                                          \
        if (cond) {                                               \
            std::printf("ok %s\n", name);                         \
        } else {                                                  \
            std::printf("FAIL %s (%s:%d)\n", name, __FILE__, __LINE__); \
            std::fflush(stdout);                                  \
            std::exit(1);                                         \
        }                                                         \
    } while (0)

template <class F>
static bool throws_invalid(F f) {
    try {
        f();
    } catch (const std::invalid_argument &) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

static void bfv() {
    const size_t n = 8192;
    EncryptionParameters parms(SchemeType::bfv);
    parms.setPolyModulusDegree(n);
    parms.setCoeffModulus(CoeffModulus::Create(n, {43, 43, 44, 44, 44}));
    parms.setPlainModulus(PlainModulus::Batching(n, 20));
    SEALContext ctx(parms);
    EXPECT(ctx.parametersSet(), "bfv context");
    const uint64_t t = parms.plainModulus().value();
    KeyGenerator keygen(ctx);
    PublicKey pk = keygen.createPublicKey();
    RelinKeys rk = keygen.createRelinKeys();
    GaloisKeys gk = keygen.createGaloisKeys(std::vector<int>{1, -1, 4});
    keygen.createGaloisKeys(std::vector<uint32_t>{(uint32_t)(2 * n - 1)}, gk);
    Encryptor enc(ctx, pk, keygen.secretKey());
    Decryptor dec(ctx, keygen.secretKey());
    Evaluator ev(ctx);
    BatchEncoder be(ctx);

    std::mt19937_64 rng(1);
    std::vector<uint64_t> x(n), y(n);
    for (auto &v : x) v = rng() % t;
    for (auto &v : y) v = rng() % t;
    Plaintext px, py, pout;
    be.encode(x, px);
    be.encode(y, py);
    Ciphertext cx, cy, c;
    enc.encrypt(px, cx);
    enc.encryptSymmetric(py, cy);
    std::vector<uint64_t> out;

    auto decrypt_eq = [&](const Ciphertext &ct, const std::vector<uint64_t> &want) {
        dec.decrypt(ct, pout);
        be.decode(pout, out);
        return out == want;
    };
    std::vector<uint64_t> w(n);
    for (size_t i = 0; i < n; i++) w[i] = (x[i] + y[i]) % t;
    ev.add(cx, cy, c);
    EXPECT(decrypt_eq(c, w), "bfv add");
    for (size_t i = 0; i < n; i++) w[i] = (x[i] + t - y[i]) % t;
    ev.sub(cx, cy, c);
    EXPECT(decrypt_eq(c, w), "bfv sub");
    for (size_t i = 0; i < n; i++) w[i] = (t - x[i]) % t;
    ev.negate(cx, c);
    EXPECT(decrypt_eq(c, w), "bfv negate");
    for (size_t i = 0; i < n; i++) w[i] = (x[i] + y[i]) % t;
    ev.addPlain(cx, py, c);
    EXPECT(decrypt_eq(c, w), "bfv add_plain");
    for (size_t i = 0; i < n; i++) w[i] = (x[i] + t - y[i]) % t;
    ev.subPlain(cx, py, c);
    EXPECT(decrypt_eq(c, w), "bfv sub_plain");
    for (size_t i = 0; i < n; i++) w[i] = (unsigned __int128)x[i] * y[i] % t;
    ev.multiplyPlain(cx, py, c);
    EXPECT(decrypt_eq(c, w), "bfv multiply_plain");
    {
        Ciphertext cn;
        ev.transformToNtt(cx, cn);
        Plaintext pn = py;
        ev.transformToNttInplace(pn, cn.parmsID());
        ev.multiplyPlainInplace(cn, pn);
        ev.transformFromNttInplace(cn);
        EXPECT(decrypt_eq(cn, w), "bfv multiply_plain ntt");
    }
    ev.multiply(cx, cy, c);
    EXPECT(c.size() == 3 && decrypt_eq(c, w), "bfv multiply (size 3)");
    ev.relinearizeInplace(c, rk);
    EXPECT(c.size() == 2 && decrypt_eq(c, w), "bfv relinearize");
    ev.modSwitchToNextInplace(c);
    EXPECT(c.coeffModulusSize() == 3 && decrypt_eq(c, w), "bfv mod_switch_to_next");
    // rows are two halves of n/2 slots; rotating by s moves slot i+s to i in each row
    const size_t h = n / 2;
    auto rot = [&](const std::vector<uint64_t> &v, int s) {
        std::vector<uint64_t> r(n);
        for (size_t row = 0; row < 2; row++)
            for (size_t i = 0; i < h; i++) r[row * h + i] = v[row * h + (i + h + s) % h];
        return r;
    };
    ev.rotateRows(cx, 1, gk, c);
    EXPECT(decrypt_eq(c, rot(x, 1)), "bfv rotate_rows 1");
    ev.rotateRows(cx, 3, gk, c);   // NAF: 4 - 1
    EXPECT(decrypt_eq(c, rot(x, 3)), "bfv rotate_rows 3 (NAF)");
    ev.rotateColumns(cx, gk, c);
    std::vector<uint64_t> sw(n);
    for (size_t i = 0; i < n; i++) sw[i] = x[(i + h) % n];
    EXPECT(decrypt_eq(c, sw), "bfv rotate_columns");
    EXPECT(throws_invalid([&] { ev.rotateRows(cx, 2, gk, c); }), "bfv missing galois key throws");
    Ciphertext low;
    ev.modSwitchToNext(cx, low);
    EXPECT(throws_invalid([&] { ev.add(cx, low, c); }), "bfv level mismatch throws");
}

static void ckks() {
    const size_t n = 8192;
    EncryptionParameters parms(SchemeType::ckks);
    parms.setPolyModulusDegree(n);
    parms.setCoeffModulus(CoeffModulus::Create(n, {60, 40, 40, 60}));
    SEALContext ctx(parms);
    EXPECT(ctx.parametersSet(), "ckks context");
    const double scale = std::pow(2.0, 40);
    KeyGenerator keygen(ctx);
    PublicKey pk = keygen.createPublicKey();
    RelinKeys rk = keygen.createRelinKeys();
    GaloisKeys gk = keygen.createGaloisKeys();
    Encryptor enc(ctx, pk);
    Decryptor dec(ctx, keygen.secretKey());
    Evaluator ev(ctx);
    CKKSEncoder ce(ctx);
    const size_t slots = ce.slotCount();
    std::mt19937_64 rng(2);
    std::uniform_real_distribution<double> u(-1, 1);
    std::vector<std::complex<double>> x(slots), y(slots), out;
    for (auto &v : x) v = {u(rng), u(rng)};
    for (auto &v : y) v = {u(rng), u(rng)};
    Plaintext px, py, p;
    ce.encode(x, scale, px);
    ce.encode(y, scale, py);
    Ciphertext cx, cy, c;
    enc.encrypt(px, cx);
    enc.encrypt(py, cy);
    auto close = [&](const Ciphertext &ct, const std::vector<std::complex<double>> &want, double tol) {
        dec.decrypt(ct, p);
        ce.decode(p, out);
        double err = 0;
        for (size_t i = 0; i < slots; i++) err = std::max(err, std::abs(out[i] - want[i]));
        return err < tol;
    };
    std::vector<std::complex<double>> w(slots);
    EXPECT(close(cx, x, 1e-5), "ckks encrypt/decrypt");
    for (size_t i = 0; i < slots; i++) w[i] = x[i] + y[i];
    ev.add(cx, cy, c);
    EXPECT(close(c, w, 1e-5), "ckks add");
    ev.addPlain(cx, py, c);
    EXPECT(close(c, w, 1e-5), "ckks add_plain");
    for (size_t i = 0; i < slots; i++) w[i] = x[i] - y[i];
    ev.sub(cx, cy, c);
    EXPECT(close(c, w, 1e-5), "ckks sub");
    for (size_t i = 0; i < slots; i++) w[i] = x[i] * y[i];
    ev.multiplyPlain(cx, py, c);
    ev.rescaleToNextInplace(c);
    EXPECT(close(c, w, 1e-4), "ckks multiply_plain + rescale");
    ev.multiply(cx, cy, c);
    ev.relinearizeInplace(c, rk);
    ev.rescaleToNextInplace(c);
    EXPECT(c.coeffModulusSize() == 2 && close(c, w, 1e-4), "ckks multiply + relinearize + rescale");
    for (size_t i = 0; i < slots; i++) w[i] = x[(i + 1) % slots];
    ev.rotateVector(cx, 1, gk, c);
    EXPECT(close(c, w, 1e-4), "ckks rotate_vector 1");
    for (size_t i = 0; i < slots; i++) w[i] = x[(i + slots - 5) % slots];
    ev.rotateVector(cx, -5, gk, c);
    EXPECT(close(c, w, 1e-4), "ckks rotate_vector -5 (NAF)");
    for (size_t i = 0; i < slots; i++) w[i] = std::conj(x[i]);
    ev.complexConjugate(cx, gk, c);
    EXPECT(close(c, w, 1e-4), "ckks complex_conjugate");
    ev.modSwitchToNext(cx, c);
    EXPECT(c.coeffModulusSize() == 2 && close(c, x, 1e-5), "ckks mod_switch_to_next (drop)");
    Ciphertext d;
    EXPECT(throws_invalid([&] { ev.add(cx, c, d); }), "ckks level mismatch throws");
    ev.add(cx, cy, cy);   // destination aliasing an operand
    for (size_t i = 0; i < slots; i++) w[i] = x[i] + y[i];
    EXPECT(close(cy, w, 1e-5), "ckks add into an operand");
    Ciphertext sq;
    ev.multiply(cx, cx, sq);
    EXPECT(throws_invalid([&] { Ciphertext t2; ev.multiply(sq, cx, t2); }), "ckks size-3 multiply throws");

    // serialization in the reference's layout: round trip, header fields, terms
    {
        std::stringstream ss;
        cx.save(ss);
        const std::string bytes = ss.str();
        uint64_t size_field = 0, n_field = 0, l_field = 0;
        std::memcpy(&size_field, bytes.data() + 33, 8);
        std::memcpy(&n_field, bytes.data() + 41, 8);
        std::memcpy(&l_field, bytes.data() + 49, 8);
        EXPECT(bytes.size() == 32 + 1 + 8 * 3 + 8 * 3 + 1 + 8 + cx.size() * cx.polyWords() * 8 && size_field == 2 &&
                   n_field == n && l_field == 3 && bytes[32] == 1,
               "ckks save layout");
        Ciphertext back;
        back.load(ss, ctx);
        EXPECT(back.toHost() == cx.toHost() && back.scale() == cx.scale() && back.isNttForm() &&
                   back.parmsID() == cx.parmsID() && back.parmsID() == ctx.idOf(3),
               "ckks save/load round trip");
        std::vector<size_t> terms{0, 1, 7, 100};
        std::stringstream st;
        cx.saveTerms(st, ev, terms);
        Ciphertext part;
        part.loadTerms(st, ev, ctx, terms);
        Ciphertext full_coeff = cx, part_coeff = part;
        ev.transformFromNttInplace(full_coeff);
        ev.transformFromNttInplace(part_coeff);
        const auto f = full_coeff.toHost().data(), g = part_coeff.toHost().data();
        bool ok = true;
        for (size_t j = 0; j < 3; j++)
            for (size_t c : terms) ok = ok && f[j * n + c] == g[j * n + c];
        for (size_t i = 3 * n; i < f.size(); i++) ok = ok && f[i] == g[i];
        EXPECT(ok, "ckks saveTerms/loadTerms keep the listed terms and c1");
    }

    EncryptionParameters bad(SchemeType::ckks);
    bad.setPolyModulusDegree(8192);
    bad.setCoeffModulus(CoeffModulus::Create(8192, {60, 60, 60, 60}));
    SEALContext insecure(bad);
    EXPECT(!insecure.parametersSet(), "insecure parameters rejected at tc128");
    EXPECT(throws_invalid([&] { Evaluator e2(insecure); }), "unset context throws");
    SEALContext none(bad, true, SecurityLevel::none);
    EXPECT(none.parametersSet(), "SecurityLevel::none accepts them");
}

int main() {
    KernelProvider::initialize();
    bfv();
    ckks();
    std::printf("facade ok\n");
    return failures ? 1 : 0;
}
