// The reference's own C++ tests, compiled against this build through the troyn:: alias.
//
// A program written for the reference includes troy_cuda.cuh and uses namespace troyn;
// here the same code includes troyhip.hpp and aliases troyn to troyhip.  The cases
// follow test/evaluator_cuda.cu:2459-2534 (BFVRelinearize: decrypt KATs of hexadecimal
// plaintexts through square, relinearize and modulus switching at N = 128, t = 64),
// plus the reference containers' behaviour: hexadecimal plaintext strings
// (src/plaintext.h:99-130, 470-498), PlaintextCuda::save / load
// (src/plaintext_cuda.cu:6-27), parms_id hash blocks (src/encryptionparams.h:36),
// cpu() / conversion constructors (src/ciphertext_cuda.cuh:20-49,
// src/plaintext_cuda.cuh:27-42), KSwitchKeys (src/kswitchkeys_cuda.cuh:31-170), key
// save / load in the reference's layouts (src/kswitchkeys_cuda.cuh:330-355), fresh
// key material per KeyGenerator, and a memory pool that serves steady-state operations
// without new device allocations.  Prints "ok <name>" per check, "troyn kat ok" at the
// end; exits 1 on the first failure.
#include "troyhip.hpp"

#include <cstdio>
#include <cstdlib>
#include <sstream>

namespace troyn = troyhip;
using namespace troyn;

#define EXPECT(cond, name)                                               \
    do {                                                                 \
        if (cond) {                                                      \
            std::printf("ok %s\n", name);                                \
        } else {                                                         \
            std::printf("FAIL %s (%s:%d)\n", name, __FILE__, __LINE__); \
            std::fflush(stdout);                                         \
            std::exit(1);                                                \
        }                                                                \
    } while (0)

template <class E, class F>
static bool throws(F f) {
    try {
        f();
    } catch (const E &) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

// EvaluatorCudaTest.BFVRelinearize (test/evaluator_cuda.cu:2459-2534), statement for
// statement.
static void bfv_relinearize() {
    EncryptionParameters parms(SchemeType::bfv);
    Modulus plain_modulus(1 << 6);
    parms.setPolyModulusDegree(128);
    parms.setPlainModulus(plain_modulus);
    parms.setCoeffModulus(CoeffModulus::Create(128, {40, 40, 40, 40}));

    SEALContext context(parms, true, SecurityLevel::none);
    KeyGenerator keygen(context);
    PublicKey pk;
    keygen.createPublicKey(pk);
    RelinKeys rlk;
    keygen.createRelinKeys(rlk);

    Encryptor encryptor(context, pk);
    Evaluator evaluator(context);
    Decryptor decryptor(context, keygen.secretKey());

    Ciphertext encrypted(context);
    Ciphertext encrypted2(context);

    Plaintext plain;
    Plaintext plain2;

    plain = 0;
    encryptor.encrypt(plain, encrypted);
    evaluator.squareInplace(encrypted);
    evaluator.relinearizeInplace(encrypted, rlk);
    decryptor.decrypt(encrypted, plain2);
    EXPECT(plain == plain2, "BFVRelinearize: 0 squared");

    encryptor.encrypt(plain, encrypted);
    evaluator.squareInplace(encrypted);
    evaluator.relinearizeInplace(encrypted, rlk);
    evaluator.squareInplace(encrypted);
    evaluator.relinearizeInplace(encrypted, rlk);
    decryptor.decrypt(encrypted, plain2);
    EXPECT(plain == plain2, "BFVRelinearize: 0 to the fourth");

    plain = "1x^10 + 2";
    encryptor.encrypt(plain, encrypted);
    evaluator.squareInplace(encrypted);
    evaluator.relinearizeInplace(encrypted, rlk);
    decryptor.decrypt(encrypted, plain2);
    EXPECT(plain2.to_string() == "1x^20 + 4x^10 + 4", "BFVRelinearize: (1x^10 + 2)^2");

    encryptor.encrypt(plain, encrypted);
    evaluator.squareInplace(encrypted);
    evaluator.relinearizeInplace(encrypted, rlk);
    evaluator.squareInplace(encrypted);
    evaluator.relinearizeInplace(encrypted, rlk);
    decryptor.decrypt(encrypted, plain2);
    EXPECT(plain2.to_string() == "1x^40 + 8x^30 + 18x^20 + 20x^10 + 10", "BFVRelinearize: (1x^10 + 2)^4");

    // Relinearization with modulus switching
    plain = "1x^10 + 2";
    encryptor.encrypt(plain, encrypted);
    evaluator.squareInplace(encrypted);
    evaluator.relinearizeInplace(encrypted, rlk);
    evaluator.modSwitchToNextInplace(encrypted);
    decryptor.decrypt(encrypted, plain2);
    EXPECT(plain2.to_string() == "1x^20 + 4x^10 + 4", "BFVRelinearize: square, modswitch");

    encryptor.encrypt(plain, encrypted);
    evaluator.squareInplace(encrypted);
    evaluator.relinearizeInplace(encrypted, rlk);
    evaluator.modSwitchToNextInplace(encrypted);
    evaluator.squareInplace(encrypted);
    evaluator.relinearizeInplace(encrypted, rlk);
    evaluator.modSwitchToNextInplace(encrypted);
    decryptor.decrypt(encrypted, plain2);
    EXPECT(plain2.to_string() == "1x^40 + 8x^30 + 18x^20 + 20x^10 + 10", "BFVRelinearize: twice with modswitch");
}

static void plaintext_strings_and_files() {
    Plaintext p("7FFx^3 + 1x^1 + 3");
    EXPECT(p.coeffCount() == 4 && p[3] == 0x7FF && p[2] == 0 && p[1] == 1 && p[0] == 3, "hex plaintext parse");
    EXPECT(p.to_string() == "7FFx^3 + 1x^1 + 3", "hex plaintext round trip");
    Plaintext lower("abcx^2 + f");
    EXPECT(lower.to_string() == "ABCx^2 + F", "lower-case hex, upper-case output");
    Plaintext zero("0");
    EXPECT(zero.to_string() == "0" && zero.isZero(), "zero polynomial");
    EXPECT(Plaintext("0x^5 + 1").to_string() == "1", "zero terms are omitted");
    for (const char *bad : {"1x^2 + 1x^3", "1x^1 +1", "x^2", "1x2", "1 + 1", "-1",
                            "10000000000000000x^1"})
        EXPECT(throws<std::invalid_argument>([&] { Plaintext q(bad); }), bad);
    Plaintext c;
    c = 0x1234;
    EXPECT(c.coeffCount() == 1 && c.to_string() == "1234", "constant assignment");
    // the same polynomial with extra zero coefficients is equal
    Plaintext padded("2x^1 + 1");
    padded.resize(100);
    EXPECT(padded == Plaintext("2x^1 + 1"), "equality ignores leading zero coefficients");
    EXPECT(!(padded == Plaintext("2x^1 + 2")), "equality sees a different coefficient");
    // PlaintextCuda::save / load layout: parms_id, coeff count, scale, word count, words
    std::stringstream ss;
    p.save(ss);
    const std::string bytes = ss.str();
    uint64_t count = 0, words = 0;
    double scale = 0;
    std::memcpy(&count, bytes.data() + 32, 8);
    std::memcpy(&scale, bytes.data() + 40, 8);
    std::memcpy(&words, bytes.data() + 48, 8);
    EXPECT(bytes.size() == 56 + 4 * 8 && count == 4 && words == 4 && scale == 1.0, "plaintext save layout");
    Plaintext back;
    back.load(ss);
    EXPECT(back == p && back.to_string() == p.to_string(), "plaintext save/load round trip");
    std::stringstream cut(bytes.substr(0, bytes.size() - 3));
    EXPECT(throws<std::invalid_argument>([&] { Plaintext q; q.load(cut); }), "truncated plaintext throws");
}

static void containers_and_keys() {
    const size_t n = 8192;  // 60 + 40 + 40 + 60 = 200 bits: within the 128-bit security bound (218)
    EncryptionParameters parms(SchemeType::ckks);
    parms.setPolyModulusDegree(n);
    parms.setCoeffModulus(CoeffModulus::Create(n, {60, 40, 40, 60}));
    SEALContext context(parms);
    // parms_id is the 4-word hash block of each level, distinct per level
    const ParmsID key = context.keyParmsID(), first = context.firstParmsID(), last = context.lastParmsID();
    EXPECT(key != first && first != last && key != parmsIDZero && first != parmsIDZero, "parms ids are hash blocks");
    EXPECT(context.getContextData(first)->nextParmsID() == context.idOf(2) &&
               context.getContextData(last)->nextParmsID() == parmsIDZero,
           "chain links by parms_id");
    ParmsID bogus = first;
    bogus[0] ^= 1;
    EXPECT(!context.getContextData(bogus), "unknown parms_id has no context data");
    // ContextDataCuda neighbours, chain indices, per-level parms and qualifiers
    // (src/context_cuda.cuh:64-135; the reference's tests walk the chain as
    // context.firstContextData()->nextContextData()->parmsID())
    auto fcd = context.firstContextData(), kcd = context.keyContextData(), lcd = context.lastContextData();
    EXPECT(fcd->nextContextData()->parmsID() == context.idOf(2) && !lcd->nextContextData() &&
               kcd->nextContextData()->parmsID() == first && fcd->prevContextData()->parmsID() == key &&
               !kcd->prevContextData() && lcd->prevContextData()->parmsID() == context.idOf(2),
           "context data next / prev");
    EXPECT(lcd->chainIndex() == 0 && fcd->chainIndex() == 2 && kcd->chainIndex() == 3, "chain indices");
    EXPECT(fcd->parms().coeffModulus().size() == 3 && fcd->parms().polyModulusDegree() == n &&
               fcd->qualifiers().using_batching && fcd->qualifiers().using_ntt && fcd->qualifiers().parametersSet() &&
               fcd->qualifiers().sec_level == SecurityLevel::tc128 && context.using_keyswitching(),
           "per-level parms and qualifiers");
    {
        const std::vector<uint64_t> w = fcd->totalCoeffModulus();
        bool divides = w.size() == 3 && w[2] != 0;
        for (const Modulus &m : fcd->coeffModulus()) {
            unsigned __int128 r = 0;
            for (size_t k = w.size(); k-- > 0;) r = ((r << 64) | w[k]) % m.value();
            divides &= r == 0;
        }
        EXPECT(divides, "totalCoeffModulus is the product of the level's primes");
    }

    KeyGenerator keygen(context), keygen2(context);
    EXPECT(keygen.secretKey().data().toHost().data() != keygen2.secretKey().data().toHost().data(),
           "two key generators draw different secret keys");
    RelinKeys rlk = keygen.createRelinKeys();
    KSwitchKeys &as_base = rlk;
    EXPECT(rlk.size() == 1 && as_base.size() == 1 && rlk.hasKey(2) && !rlk.hasKey(3) && rlk.parmsID() == key,
           "relin keys as KSwitchKeys");
    GaloisKeys gk = keygen.createGaloisKeys(std::vector<int>{1, -1});
    const uint32_t e1 = 3;
    EXPECT(gk.size() == 2 && gk.hasKey(e1) && GaloisKeys::getIndex(e1) == 1 && !gk.hasKey(5), "galois keys index");
    EXPECT(throws<std::invalid_argument>([&] { (void)gk.data(2); }), "absent key throws");

    CKKSEncoder encoder(context);
    Encryptor encryptor(context, keygen.createPublicKey());
    Decryptor decryptor(context, keygen.secretKey());
    Evaluator evaluator(context);
    std::vector<std::complex<double>> v(encoder.slotCount());
    for (size_t i = 0; i < v.size(); i++) v[i] = {std::sin(0.1 * i), std::cos(0.3 * i)};
    Plaintext p;
    encoder.encode(v, std::pow(2.0, 40), p);
    Ciphertext ct = encryptor.encrypt(p);
    // cpu() and the conversion constructor carry every field and word
    troyhip::host::Ciphertext h = ct.cpu();
    Ciphertext again(h);
    EXPECT(h.size() == 2 && h.coeffModulusSize() == 3 && h.isNttForm() && h.parmsID() == first &&
               again.toHost() == h && again.parmsID() == ct.parmsID() && again.scale() == ct.scale(),
           "ciphertext cpu() round trip");
    troyhip::host::Plaintext hp = p.cpu();
    Plaintext p2(hp);
    EXPECT(p2 == p && p2.isNttForm() && p2.parmsID() == first, "plaintext cpu() round trip");
    // a ciphertext loaded from a stream names its level by the hash block
    std::stringstream ss;
    ct.save(ss);
    Ciphertext loaded;
    loaded.load(ss, context);
    EXPECT(loaded.parmsID() == first && loaded.coeffModulusSize() == 3, "loaded ciphertext level");

    // reserve / capacity (src/ciphertext_cuda.cuh:125-215, src/plaintext_cuda.cuh:59-72)
    Plaintext pr("3x^2 + 1");
    pr.reserve(5000);
    EXPECT(pr.capacity() >= 5000 && pr.to_string() == "3x^2 + 1", "plaintext reserve keeps its coefficients");
    pr.shrinkToFit();  // (pool blocks are whole 4 KiB pages)
    EXPECT(pr.capacity() < 5000 && pr.to_string() == "3x^2 + 1", "plaintext shrinkToFit keeps its coefficients");
    EXPECT(throws<std::logic_error>([&] { Plaintext q = p; q.reserve(10); }), "no reserve for an NTT-form plaintext");
    Ciphertext cr = ct;
    const std::vector<uint64_t> words = cr.toHost().data();
    cr.reserve(5);
    EXPECT(cr.sizeCapacity() >= 5 && cr.size() == 2 && cr.toHost().data() == words, "ciphertext reserve keeps its words");
    cr.resize(3);
    EXPECT(cr.sizeCapacity() >= 5 && cr.size() == 3, "a resize within the capacity keeps the block");
    EXPECT(throws<std::invalid_argument>([&] { cr.reserve(17); }), "size_capacity above 16 throws");

    // steady state: a mul + relin + rescale loop reuses pooled blocks
    RelinKeys rk = keygen.createRelinKeys();
    auto step = [&] {
        Ciphertext a = ct;
        evaluator.multiplyInplace(a, ct);
        evaluator.relinearizeInplace(a, rk);
        evaluator.rescaleToNextInplace(a);
        return a;
    };
    // warm up with the loop's own pattern (the previous result is alive while the next
    // one is computed), then a steady loop takes every block from the pool
    Ciphertext r;
    for (int i = 0; i < 3; i++) r = step();
    const size_t allocs = MemoryPool::deviceAllocations();
    for (int i = 0; i < 10; i++) r = step();
    EXPECT(MemoryPool::deviceAllocations() == allocs, "pooled steady state: no device allocation");
    Plaintext out;
    decryptor.decrypt(r, out);
    std::vector<std::complex<double>> got;
    encoder.decode(out, got);
    double err = 0;
    for (size_t i = 0; i < v.size(); i++) err = std::max(err, std::abs(got[i] - v[i] * v[i]));
    EXPECT(err < 1e-4, "pooled result decrypts to the square");
    const size_t cached = MemoryPool::cachedBytes();
    EXPECT(MemoryPool::release() == cached && MemoryPool::cachedBytes() == 0, "pool release");
}

// SecretKeyCuda / PublicKeyCuda / KSwitchKeysCuda save / load (src/secretkey_cuda.cuh:292-297,
// src/publickey_cuda.cuh:252-257, src/kswitchkeys_cuda.cuh:330-355): the reference's byte
// layouts, and loaded keys that decrypt, encrypt, relinearize and rotate as the originals.
static void key_files() {
    const size_t n = 8192;
    EncryptionParameters parms(SchemeType::ckks);
    parms.setPolyModulusDegree(n);
    parms.setCoeffModulus(CoeffModulus::Create(n, {60, 40, 40, 60}));
    SEALContext context(parms);
    const size_t K = 4;
    const ParmsID key = context.keyParmsID();
    KeyGenerator keygen(context);
    RelinKeys rlk = keygen.createRelinKeys();
    GaloisKeys gk = keygen.createGaloisKeys(std::vector<int>{1});  // one key among N slots

    // relin keys: parms_id, slot count, then per slot the public-key count and each
    // public key as a size-2 ciphertext of the key level
    std::stringstream ks;
    rlk.save(ks);
    const std::string kb = ks.str();
    const size_t pk_bytes = 82 + 8 + 2 * K * n * 8;  // header, word count, words
    uint64_t slots = 0, m = 0, size = 0, nn = 0, L = 0, words = 0;
    std::memcpy(&slots, kb.data() + 32, 8);
    std::memcpy(&m, kb.data() + 40, 8);
    const char *pk0 = kb.data() + 48;
    ParmsID pid{};
    std::memcpy(pid.data(), pk0, 32);
    std::memcpy(&size, pk0 + 33, 8);
    std::memcpy(&nn, pk0 + 41, 8);
    std::memcpy(&L, pk0 + 49, 8);
    std::memcpy(&words, pk0 + 82, 8);
    double scale = 0;
    std::memcpy(&scale, pk0 + 57, 8);
    EXPECT(kb.size() == 48 + (K - 1) * pk_bytes && slots == 1 && m == K - 1 && pid == key && pk0[32] == 1 &&
               size == 2 && nn == n && L == K && scale == 1.0 && words == 2 * K * n,
           "relin keys save layout");
    const std::vector<uint64_t> rw = rlk.data()[0].toHost();
    EXPECT(std::memcmp(pk0 + 90, rw.data(), 2 * K * n * 8) == 0 &&
               std::memcmp(kb.data() + 48 + pk_bytes + 90, rw.data() + 2 * K * n, 2 * K * n * 8) == 0,
           "relin key words in public-key order");
    RelinKeys rlk2;
    rlk2.load(ks);
    EXPECT(rlk2.parmsID() == key && rlk2.hasKey(2) && rlk2.data()[0].toHost() == rw &&
               rlk2.polyModulusDegree() == n && rlk2.keyModulusSize() == K,
           "relin keys load round trip");

    // galois keys: absent slots are saved as empty
    std::stringstream gs;
    gk.save(gs);
    GaloisKeys gk2;
    gk2.load(gs);
    const uint32_t e1 = 3;
    EXPECT(gk2.size() == 1 && gk2.hasKey(e1) && !gk2.hasKey(5) && gk2.data().size() == gk.data().size() &&
               gk2.key(e1).toHost() == gk.key(e1).toHost(),
           "galois keys load round trip");

    // secret and public keys: plaintext and ciphertext layouts
    std::stringstream ss, ps;
    keygen.secretKey().save(ss);
    const PublicKey pk = keygen.createPublicKey();
    pk.save(ps);
    SecretKey sk2;
    sk2.load(ss);
    PublicKey pk2;
    pk2.load(ps);
    EXPECT(sk2.parmsID() == key && sk2.data() == keygen.secretKey().data(), "secret key load round trip");
    EXPECT(pk2.parmsID() == key && pk2.data().size() == 2 && pk2.data().isNttForm() &&
               pk2.data().toHost().data() == pk.data().toHost().data(),
           "public key load round trip");

    // the loaded keys at work: encrypt with pk2, square, relinearize with rlk2, rotate
    // with gk2, decrypt with sk2
    CKKSEncoder encoder(context);
    Encryptor encryptor(context, pk2);
    Decryptor decryptor(context, sk2);
    Evaluator evaluator(context);
    std::vector<std::complex<double>> v(encoder.slotCount());
    for (size_t i = 0; i < v.size(); i++) v[i] = {std::sin(0.1 * i), std::cos(0.3 * i)};
    Plaintext p;
    encoder.encode(v, std::pow(2.0, 40), p);
    Ciphertext ct = encryptor.encrypt(p);
    evaluator.squareInplace(ct);
    evaluator.relinearizeInplace(ct, rlk2);
    evaluator.rescaleToNextInplace(ct);
    evaluator.rotateVectorInplace(ct, 1, gk2);
    Plaintext out;
    decryptor.decrypt(ct, out);
    std::vector<std::complex<double>> got;
    encoder.decode(out, got);
    double err = 0;
    for (size_t i = 0; i < v.size(); i++) {
        const std::complex<double> w = v[(i + 1) % v.size()];
        err = std::max(err, std::abs(got[i] - w * w));
    }
    EXPECT(err < 1e-4, "loaded keys: square, relinearize, rotate, decrypt");

    // a truncated key stream and a key of another shape are rejected
    std::stringstream cut(kb.substr(0, kb.size() - 5));
    EXPECT(throws<std::invalid_argument>([&] { RelinKeys r; r.load(cut); }), "truncated keys throw");
    std::string bad = kb;
    uint64_t wrong_m = K;
    std::memcpy(&bad[40], &wrong_m, 8);
    std::stringstream bs(bad);
    EXPECT(throws<std::invalid_argument>([&] { RelinKeys r; r.load(bs); }), "public-key count must be K - 1");
}

// EvaluatorCudaTest.BFVEncryptMultiplyManyDecrypt and BFVEncryptExponentiateDecrypt
// (test/evaluator_cuda.cu:4575-4732), statement for statement (ASSERT_EQ / ASSERT_TRUE as
// EXPECT), and the first two cases of BFVEncryptAddManyDecrypt (:4734-4790) in the same
// context.
static void bfv_many_and_exponentiate() {
    EncryptionParameters parms(SchemeType::bfv);
    Modulus plain_modulus(1 << 6);
    parms.setPolyModulusDegree(128);
    parms.setPlainModulus(plain_modulus);
    parms.setCoeffModulus(CoeffModulus::Create(128, {40, 40, 40}));

    SEALContext context(parms, false, SecurityLevel::none);
    // a chain that is not expanded: the key level and the first level only
    EXPECT(context.firstContextData()->chainIndex() == 0 && context.keyContextData()->chainIndex() == 1 &&
               !context.firstContextData()->nextContextData() && context.lastParmsID() == context.firstParmsID() &&
               !context.firstContextData()->qualifiers().using_batching &&
               context.firstContextData()->qualifiers().using_fast_plain_lift &&
               context.firstContextData()->qualifiers().sec_level == SecurityLevel::none,
           "unexpanded chain: indices, neighbours, BFV qualifiers (t = 64: no batching)");
    KeyGenerator keygen(context);
    PublicKey pk;
    keygen.createPublicKey(pk);
    RelinKeys rlk;
    keygen.createRelinKeys(rlk);

    Encryptor encryptor(context, pk);
    Evaluator evaluator(context);
    Decryptor decryptor(context, keygen.secretKey());

    Ciphertext encrypted1, encrypted2, encrypted3, encrypted4, product;
    Plaintext plain, plain1, plain2, plain3, plain4;
    auto same_level = [&](const std::vector<Ciphertext> &v, const Ciphertext &r) {
        for (const Ciphertext &c : v)
            if (!(c.parmsID() == r.parmsID())) return false;
        return r.parmsID() == context.firstParmsID();
    };

    plain1 = "1x^2 + 1";
    plain2 = "1x^2 + 1x^1";
    plain3 = "1x^2 + 1x^1 + 1";
    encryptor.encrypt(plain1, encrypted1);
    encryptor.encrypt(plain2, encrypted2);
    encryptor.encrypt(plain3, encrypted3);
    std::vector<Ciphertext> encrypteds{encrypted1, encrypted2, encrypted3};
    evaluator.multiplyMany(encrypteds, rlk, product);
    decryptor.decrypt(product, plain);
    EXPECT(encrypteds.size() == 3 && plain.to_string() == "1x^6 + 2x^5 + 3x^4 + 3x^3 + 2x^2 + 1x^1" &&
               same_level(encrypteds, product),
           "BFVEncryptMultiplyManyDecrypt: three");

    plain1 = "3Fx^3 + 3F";
    plain2 = "3Fx^4 + 3F";
    encryptor.encrypt(plain1, encrypted1);
    encryptor.encrypt(plain2, encrypted2);
    encrypteds = {encrypted1, encrypted2};
    evaluator.multiplyMany(encrypteds, rlk, product);
    decryptor.decrypt(product, plain);
    EXPECT(encrypteds.size() == 2 && plain.to_string() == "1x^7 + 1x^4 + 1x^3 + 1" && same_level(encrypteds, product),
           "BFVEncryptMultiplyManyDecrypt: two");

    plain1 = "1x^1";
    plain2 = "3Fx^4 + 3Fx^3 + 3Fx^2 + 3Fx^1 + 3F";
    plain3 = "1x^2 + 1x^1 + 1";
    encryptor.encrypt(plain1, encrypted1);
    encryptor.encrypt(plain2, encrypted2);
    encryptor.encrypt(plain3, encrypted3);
    encrypteds = {encrypted1, encrypted2, encrypted3};
    evaluator.multiplyMany(encrypteds, rlk, product);
    decryptor.decrypt(product, plain);
    EXPECT(encrypteds.size() == 3 && plain.to_string() == "3Fx^7 + 3Ex^6 + 3Dx^5 + 3Dx^4 + 3Dx^3 + 3Ex^2 + 3Fx^1" &&
               same_level(encrypteds, product),
           "BFVEncryptMultiplyManyDecrypt: three, negative coefficients");

    plain1 = "1";
    plain2 = "3F";
    plain3 = "1";
    plain4 = "3F";
    encryptor.encrypt(plain1, encrypted1);
    encryptor.encrypt(plain2, encrypted2);
    encryptor.encrypt(plain3, encrypted3);
    encryptor.encrypt(plain4, encrypted4);
    encrypteds = {encrypted1, encrypted2, encrypted3, encrypted4};
    evaluator.multiplyMany(encrypteds, rlk, product);
    decryptor.decrypt(product, plain);
    EXPECT(encrypteds.size() == 4 && plain.to_string() == "1" && same_level(encrypteds, product),
           "BFVEncryptMultiplyManyDecrypt: four constants");

    plain1 = "1x^16 + 1x^15 + 1x^8 + 1x^7 + 1x^6 + 1x^3 + 1x^2 + 1";
    plain2 = "0";
    plain3 = "1x^13 + 1x^12 + 1x^5 + 1x^4 + 1x^3 + 1";
    plain4 = "1x^15 + 1x^10 + 1x^9 + 1x^8 + 1x^2 + 1x^1 + 1";
    encryptor.encrypt(plain1, encrypted1);
    encryptor.encrypt(plain2, encrypted2);
    encryptor.encrypt(plain3, encrypted3);
    encryptor.encrypt(plain4, encrypted4);
    encrypteds = {encrypted1, encrypted2, encrypted3, encrypted4};
    evaluator.multiplyMany(encrypteds, rlk, product);
    decryptor.decrypt(product, plain);
    EXPECT(encrypteds.size() == 4 && plain.to_string() == "0" && same_level(encrypteds, product),
           "BFVEncryptMultiplyManyDecrypt: a zero factor");

    // the product tree: ((c1 c2) (c3 c4)) with a relinearization after each product,
    // word for word
    Ciphertext p12, p34, tree;
    evaluator.multiply(encrypted1, encrypted2, p12);
    evaluator.relinearizeInplace(p12, rlk);
    evaluator.multiply(encrypted3, encrypted4, p34);
    evaluator.relinearizeInplace(p34, rlk);
    evaluator.multiply(p12, p34, tree);
    evaluator.relinearizeInplace(tree, rlk);
    EXPECT(tree.toHost().data() == product.toHost().data(), "multiplyMany follows the reference's product tree");
    EXPECT(throws<std::invalid_argument>([&] { evaluator.multiplyMany({}, rlk, product); }), "empty list throws");

    Ciphertext encrypted;
    plain = "1x^2 + 1";
    encryptor.encrypt(plain, encrypted);
    evaluator.exponentiateInplace(encrypted, 1, rlk);
    decryptor.decrypt(encrypted, plain);
    EXPECT(plain.to_string() == "1x^2 + 1" && encrypted.parmsID() == context.firstParmsID(),
           "BFVEncryptExponentiateDecrypt: 1");

    plain = "1x^2 + 1x^1 + 1";
    encryptor.encrypt(plain, encrypted);
    evaluator.exponentiateInplace(encrypted, 2, rlk);
    decryptor.decrypt(encrypted, plain);
    EXPECT(plain.to_string() == "1x^4 + 2x^3 + 3x^2 + 2x^1 + 1" && encrypted.parmsID() == context.firstParmsID(),
           "BFVEncryptExponentiateDecrypt: 2");

    plain = "3Fx^2 + 3Fx^1 + 3F";
    encryptor.encrypt(plain, encrypted);
    evaluator.exponentiateInplace(encrypted, 3, rlk);
    decryptor.decrypt(encrypted, plain);
    EXPECT(plain.to_string() == "3Fx^6 + 3Dx^5 + 3Ax^4 + 39x^3 + 3Ax^2 + 3Dx^1 + 3F" &&
               encrypted.parmsID() == context.firstParmsID(),
           "BFVEncryptExponentiateDecrypt: 3");

    plain = "1x^8";
    encryptor.encrypt(plain, encrypted);
    evaluator.exponentiateInplace(encrypted, 4, rlk);
    decryptor.decrypt(encrypted, plain);
    EXPECT(plain.to_string() == "1x^32" && encrypted.parmsID() == context.firstParmsID(),
           "BFVEncryptExponentiateDecrypt: 4");
    EXPECT(throws<std::invalid_argument>([&] { evaluator.exponentiateInplace(encrypted, 0, rlk); }),
           "exponent 0 throws");

    Ciphertext sum;
    plain1 = "1x^2 + 1";
    plain2 = "1x^2 + 1x^1";
    plain3 = "1x^2 + 1x^1 + 1";
    encryptor.encrypt(plain1, encrypted1);
    encryptor.encrypt(plain2, encrypted2);
    encryptor.encrypt(plain3, encrypted3);
    encrypteds = {encrypted1, encrypted2, encrypted3};
    evaluator.addMany(encrypteds, sum);
    decryptor.decrypt(sum, plain);
    EXPECT(plain.to_string() == "3x^2 + 2x^1 + 2" && same_level(encrypteds, sum), "BFVEncryptAddManyDecrypt: three");

    plain1 = "3Fx^3 + 3F";
    plain2 = "3Fx^4 + 3F";
    encryptor.encrypt(plain1, encrypted1);
    encryptor.encrypt(plain2, encrypted2);
    encrypteds = {encrypted1, encrypted2};
    evaluator.addMany(encrypteds, sum);
    decryptor.decrypt(sum, plain);
    EXPECT(plain.to_string() == "3Fx^4 + 3Fx^3 + 3E" && same_level(encrypteds, sum), "BFVEncryptAddManyDecrypt: two");
}

// BatchEncoderCuda / CKKSEncoderCuda encodePolynomial / decodePolynomial
// (src/batchencoder_cuda.cu:124-170, 267-285; src/ckks_cuda.cu:455-572, 983-1053): the
// values are coefficients; products of encrypted polynomials decode to the polynomial
// products (the coefficient-encoded callers, app/LinearHelperCKKS.cuh:139-670).
static void polynomial_encoders() {
    {
        EncryptionParameters parms(SchemeType::bfv);
        parms.setPolyModulusDegree(4096);
        parms.setPlainModulus(PlainModulus::Batching(4096, 17));
        parms.setCoeffModulus(CoeffModulus::Create(4096, {36, 36, 37}));
        SEALContext context(parms);
        const uint64_t t = parms.plainModulus().value();
        EXPECT(context.firstContextData()->qualifiers().using_batching &&
                   context.firstContextData()->plainUpperHalfThreshold() == (t + 1) / 2,
               "BFV qualifiers: batching with a prime t = 1 mod 2N");
        BatchEncoder encoder(context);
        Plaintext p;
        encoder.encodePolynomial(std::vector<uint64_t>{1, 2, t + 4463}, p);
        EXPECT(p.coeffCount() == 3 && p.to_string() == "116Fx^2 + 2x^1 + 1" && !p.isNttForm(),
               "BatchEncoder::encodePolynomial (unsigned): coefficients mod t, values.size() of them");
        Plaintext q;
        encoder.encodePolynomial(std::vector<int64_t>{-1, 5, -(int64_t)t, -3}, q);
        std::vector<uint64_t> qu;
        encoder.decodePolynomial(q, qu);
        EXPECT(q.coeffCount() == 4096 && qu.size() == 4096 && qu[0] == t - 1 && qu[1] == 5 && qu[2] == t &&
                   qu[3] == t - 3 && qu[4] == 0,
               "BatchEncoder::encodePolynomial (signed): t - (|v| mod t), padded to N");
        std::vector<int64_t> qs;
        encoder.decodePolynomial(q, qs);
        EXPECT(qs[0] == -1 && qs[1] == 5 && qs[3] == -3 && qs[4] == 0, "BatchEncoder::decodePolynomial (signed)");
        std::vector<uint64_t> pu;
        encoder.decodePolynomial(p, pu);
        EXPECT(pu.size() == 4096 && pu[0] == 1 && pu[1] == 2 && pu[2] == 4463 && pu[3] == 0,
               "decodePolynomial of a short plaintext pads with zeros");

        KeyGenerator keygen(context);
        Encryptor encryptor(context, keygen.createPublicKey());
        Decryptor decryptor(context, keygen.secretKey());
        Evaluator evaluator(context);
        RelinKeys rk = keygen.createRelinKeys();
        Plaintext a, b, out;
        encoder.encodePolynomial(std::vector<int64_t>{2, 1}, a);  // x + 2
        encoder.encodePolynomial(std::vector<int64_t>{3, -1}, b);  // 3 - x
        Ciphertext ca = encryptor.encrypt(a), cb = encryptor.encrypt(b);
        evaluator.multiplyInplace(ca, cb);
        evaluator.relinearizeInplace(ca, rk);
        decryptor.decrypt(ca, out);
        std::vector<int64_t> prod;
        encoder.decodePolynomial(out, prod);
        EXPECT(prod[0] == 6 && prod[1] == 1 && prod[2] == -1 && prod[3] == 0,
               "BFV: (x + 2)(3 - x) through encodePolynomial, multiply, relinearize, decodePolynomial");

        // EncryptorCuda::encryptZero at a lower level, encryptZeroSymmetric, setPublicKey
        // (src/encryptor_cuda.cuh:87-101, 170-388)
        const ParmsID second = context.getContextData(context.firstParmsID())->nextParmsID();
        Ciphertext z = encryptor.encryptZero(second);
        decryptor.decrypt(z, out);
        EXPECT(z.parmsID() == second && z.coeffModulusSize() == 1 && !z.isNttForm() && out.to_string() == "0",
               "encryptZero at the second level decrypts to 0");
        Encryptor sym(context, keygen.secretKey());
        Ciphertext zs = sym.encryptZeroSymmetric();
        decryptor.decrypt(zs, out);
        EXPECT(zs.parmsID() == context.firstParmsID() && out.to_string() == "0" &&
                   throws<std::logic_error>([&] { (void)sym.encryptZero(); }),
               "encryptZeroSymmetric decrypts to 0; no public key, no encryptZero");
        sym.setPublicKey(keygen.createPublicKey());
        Ciphertext ca2 = sym.encrypt(a);
        decryptor.decrypt(ca2, out);
        std::vector<int64_t> av;
        encoder.decodePolynomial(out, av);
        EXPECT(av[0] == 2 && av[1] == 1 && av[2] == 0, "setPublicKey enables public-key encryption");
    }
    {
        EncryptionParameters parms(SchemeType::ckks);
        parms.setPolyModulusDegree(8192);
        parms.setCoeffModulus(CoeffModulus::Create(8192, {60, 40, 40, 60}));
        SEALContext context(parms);
        CKKSEncoder encoder(context);
        const double scale = std::pow(2.0, 40);
        Plaintext p;
        encoder.encodePolynomial(std::vector<double>{0.5, 0.25}, scale, p);
        std::vector<double> back;
        encoder.decodePolynomial(p, back);
        EXPECT(p.isNttForm() && p.parmsID() == context.firstParmsID() && p.scale() == scale && back.size() == 8192 &&
                   back[0] == 0.5 && back[1] == 0.25 && back[2] == 0.0,
               "CKKSEncoder::encodePolynomial / decodePolynomial round trip");
        KeyGenerator keygen(context);
        Encryptor encryptor(context, keygen.createPublicKey());
        Decryptor decryptor(context, keygen.secretKey());
        Evaluator evaluator(context);
        Plaintext m, out;
        encoder.encodePolynomial(std::vector<double>{2.0, -1.0}, scale, m);  // 2 - x
        Ciphertext c = encryptor.encrypt(p);                                   // 0.5 + 0.25 x
        evaluator.multiplyPlainInplace(c, m);
        evaluator.rescaleToNextInplace(c);
        decryptor.decrypt(c, out);
        std::vector<double> r;
        encoder.decodePolynomial(out, r);
        double err = std::fabs(r[0] - 1.0) + std::fabs(r[1]) + std::fabs(r[2] + 0.25);
        for (size_t i = 3; i < r.size(); i++) err = std::max(err, std::fabs(r[i]));
        EXPECT(err < 1e-5, "CKKS: (0.5 + 0.25x)(2 - x) = 1 - 0.25x^2 through encodePolynomial and multiplyPlain");
    }
}

// CoeffModulus::BFVDefault (src/modulus.cpp:53-78): every row of the three tables holds
// distinct primes = 1 mod 2N whose bits add up to at most the security bound (to it, but
// for two 192-bit rows), and N = 8192 at
// 128 bits is the BFV plain-path workload's list; that workload (N = 8192, t = 65537,
// two batched plaintexts, multiply, relinearize, decrypt) runs through the troyn:: API.
static void bfv_default_moduli() {
    bool ok = true;
    for (SecurityLevel sec : {SecurityLevel::tc128, SecurityLevel::tc192, SecurityLevel::tc256})
        for (size_t n = 1024; n <= 32768; n *= 2) {
            const std::vector<Modulus> q = CoeffModulus::BFVDefault(n, sec);
            int bits = 0;
            for (size_t i = 0; i < q.size(); i++) {
                bits += q[i].bitCount();
                ok &= troyhip::detail::is_prime64(q[i].value()) && q[i].value() % (2 * n) == 1;
                for (size_t j = 0; j < i; j++) ok &= q[i] != q[j];
            }
            ok &= bits <= CoeffModulus::MaxBitCount(n, sec) && bits + 12 > CoeffModulus::MaxBitCount(n, sec);
        }
    EXPECT(ok, "BFVDefault: primes = 1 mod 2N within the security bound, for every degree and level");
    const std::vector<Modulus> d = CoeffModulus::BFVDefault(8192);
    const std::vector<uint64_t> want{0x7fffffd8001, 0x7fffffc8001, 0xfffffffc001, 0xffffff6c001, 0xfffffebc001};
    bool same = d.size() == want.size();
    for (size_t i = 0; same && i < d.size(); i++) same = d[i].value() == want[i];
    EXPECT(same, "BFVDefault(8192) is the BFV plain-path workload's list");
    EXPECT(throws<std::invalid_argument>([] { (void)CoeffModulus::BFVDefault(65536); }) &&
               throws<std::invalid_argument>([] { (void)CoeffModulus::BFVDefault(8192, SecurityLevel::none); }),
           "BFVDefault rejects other degrees and no security level");

    EncryptionParameters parms(SchemeType::bfv);
    parms.setPolyModulusDegree(8192);
    parms.setCoeffModulus(CoeffModulus::BFVDefault(8192));
    parms.setPlainModulus(65537);
    SEALContext context(parms);
    KeyGenerator keygen(context);
    Encryptor encryptor(context, keygen.createPublicKey());
    Decryptor decryptor(context, keygen.secretKey());
    Evaluator evaluator(context);
    BatchEncoder encoder(context);
    RelinKeys rk = keygen.createRelinKeys();
    std::vector<uint64_t> x(8192), y(8192);
    uint64_t st = 1;
    auto next = [&st] {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (st >> 33) % 65537;
    };
    for (size_t i = 0; i < 8192; i++) {
        x[i] = next();
        y[i] = next();
    }
    Plaintext px, py, out;
    encoder.encode(x, px);
    encoder.encode(y, py);
    Ciphertext cx = encryptor.encrypt(px), cy = encryptor.encrypt(py);
    evaluator.multiplyInplace(cx, cy);
    evaluator.relinearizeInplace(cx, rk);
    decryptor.decrypt(cx, out);
    std::vector<uint64_t> z;
    encoder.decode(out, z);
    bool exact = z.size() == 8192;
    for (size_t i = 0; exact && i < 8192; i++) exact = z[i] == x[i] * y[i] % 65537;
    EXPECT(exact, "BFV plain path at N = 8192, BFVDefault, t = 65537: slots of the product exact");
}

int main() {
    KernelProvider::initialize();
    bfv_relinearize();
    plaintext_strings_and_files();
    containers_and_keys();
    key_files();
    bfv_many_and_exponentiate();
    polynomial_encoders();
    bfv_default_moduli();
    std::printf("troyn kat ok\n");
    return 0;
}
