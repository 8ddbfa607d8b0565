// extern "C" shim over the reference's CPU API (troy_cpu.h), compiled only into
// oracle/_ref/libtroyref.so by oracle/ref.mk.  It lets the Python tests feed the
// same inputs to the reference CPU path and to the oracle restatement, and lets
// bench.py time the reference CPU path.  Each function names the reference code it
// drives.  Keys and ciphertexts are injected as raw words through the reference's
// mutable data() accessors, so no reference PRNG stream is involved.

#include "troy_cpu.h"
#include "utils/ntt.h"
#include "utils/numth.h"
#include "utils/galois.h"
#include "utils/rns.h"
#include <cstdio>
#include <cstring>
#include <vector>
#include <memory>

using namespace troy;
using namespace troy::util;

extern "C" {

// CoeffModulus::Create (modulus.cpp:80-121)
int ref_coeff_modulus_create(uint64_t n, const int* bits, uint64_t count, uint64_t* out) {
    try {
        std::vector<int> b(bits, bits + count);
        auto mods = CoeffModulus::Create(n, b);
        for (size_t i = 0; i < mods.size(); i++) out[i] = mods[i].value();
        return 0;
    } catch (...) { return -1; }
}

// getPrimes (utils/numth.cpp:261-283)
int ref_get_primes(uint64_t factor, int bit_size, uint64_t count, uint64_t* out) {
    try {
        auto p = getPrimes(factor, bit_size, count);
        for (size_t i = 0; i < p.size(); i++) out[i] = p[i].value();
        return 0;
    } catch (...) { return -1; }
}

// NTTTables::initialize (utils/ntt.cpp:17-66): operand and Shoup quotient of every
// root power, plus n^-1.
int ref_ntt_tables(int logn, uint64_t q, uint64_t* rp, uint64_t* rq, uint64_t* irp, uint64_t* irq,
                   uint64_t* inv_n) {
    try {
        NTTTables t(logn, Modulus(q));
        size_t n = size_t(1) << logn;
        for (size_t i = 0; i < n; i++) {
            rp[i] = t.getFromRootPowers(i).operand;
            rq[i] = t.getFromRootPowers(i).quotient;
            irp[i] = t.getFromInvRootPowers(i).operand;
            irq[i] = t.getFromInvRootPowers(i).quotient;
        }
        inv_n[0] = t.invDegreeModulo().operand;
        inv_n[1] = t.invDegreeModulo().quotient;
        return 0;
    } catch (...) { return -1; }
}

// nttNegacyclicHarvey / inverseNttNegacyclicHarvey (utils/ntt.cpp:158-213), fully reduced,
// applied to `count` consecutive length-2^logn polynomials all modulo q.
int ref_ntt(int logn, uint64_t q, uint64_t* data, uint64_t count, int inverse) {
    try {
        NTTTables t(logn, Modulus(q));
        size_t n = size_t(1) << logn;
        for (size_t c = 0; c < count; c++) {
            if (inverse) inverseNttNegacyclicHarvey(HostPointer<uint64_t>(data + c * n), t);
            else nttNegacyclicHarvey(HostPointer<uint64_t>(data + c * n), t);
        }
        return 0;
    } catch (...) { return -1; }
}

// GaloisTool::getEltFromStep (utils/galois.cpp:44-86)
int64_t ref_galois_elt_from_step(int logn, int step) {
    try {
        GaloisTool g(logn);
        return g.getEltFromStep(step);
    } catch (...) { return -1; }
}

// GaloisTool::applyGalois / applyGaloisNtt (utils/galois.cpp:143-175)
int ref_apply_galois(int logn, uint32_t elt, uint64_t q, const uint64_t* in, uint64_t* out, int ntt) {
    try {
        GaloisTool g(logn);
        if (ntt) g.applyGaloisNtt(ConstHostPointer<uint64_t>(in), elt, HostPointer<uint64_t>(out));
        else g.applyGalois(ConstHostPointer<uint64_t>(in), elt, Modulus(q), HostPointer<uint64_t>(out));
        return 0;
    } catch (...) { return -1; }
}

// ---------------------------------------------------------------- evaluator with injected words

struct RefSession {
    std::unique_ptr<SEALContext> context;
    std::unique_ptr<Evaluator> evaluator;
    SchemeType scheme;
    size_t n;
    size_t K;
};

// EncryptionParameters + SEALContext(parms, expand_mod_chain = true, SecurityLevel::none)
// (src/context.cpp:458-522).  scheme: 1 = BFV, 2 = CKKS.
void *ref_session_create(int scheme, uint64_t n, const uint64_t *moduli, uint64_t K, uint64_t plain) {
    try {
        auto s = new RefSession();
        s->scheme = scheme == 1 ? SchemeType::bfv : SchemeType::ckks;
        EncryptionParameters parms(s->scheme);
        parms.setPolyModulusDegree(n);
        std::vector<Modulus> mods;
        for (size_t i = 0; i < K; i++) mods.emplace_back(moduli[i]);
        parms.setCoeffModulus(mods);
        if (scheme == 1) parms.setPlainModulus(plain);
        s->context.reset(new SEALContext(parms, true, SecurityLevel::none));
        if (!s->context->parametersSet()) { delete s; return nullptr; }
        s->evaluator.reset(new Evaluator(*s->context));
        s->n = n;
        s->K = K;
        return s;
    } catch (...) { return nullptr; }
}

void ref_session_destroy(void *h) { delete static_cast<RefSession *>(h); }

static ParmsID parms_for_level(const RefSession &s, size_t L) {
    auto cd = s.context->firstContextData();
    while (cd && cd->parms().coeffModulus().size() != L) cd = cd->nextContextData();
    if (!cd) throw std::invalid_argument("no level with that many primes");
    return cd->parmsID();
}

static void load_ct(const RefSession &s, Ciphertext &ct, size_t L, size_t size, const uint64_t *data, double scale) {
    ct.resize(*s.context, parms_for_level(s, L), size);
    ct.isNttForm() = (s.scheme == SchemeType::ckks);
    ct.scale() = (s.scheme == SchemeType::ckks) ? scale : 1.0;
    std::memcpy(ct.data(), data, size * L * s.n * 8);
}

// KSwitchKeys with one key vector of K-1 size-2 key-level ciphertexts (NTT form),
// filled from words laid out [K-1][2][K][N] (KeyGenerator::generateOneKswitchKey layout,
// src/keygenerator.cpp:294-329).
static void load_kswitch(const RefSession &s, KSwitchKeys &ks, size_t index, const uint64_t *words) {
    auto key_parms = s.context->keyParmsID();
    ks.parmsID() = key_parms;
    ks.data().resize(index + 1);
    auto &vec = ks.data()[index];
    vec.clear();
    const size_t dnum = s.K - 1, per = 2 * s.K * s.n;
    for (size_t j = 0; j < dnum; j++) {
        PublicKey pk;
        pk.data().resize(*s.context, key_parms, 2);
        pk.data().isNttForm() = true;
        std::memcpy(pk.data().data(), words + j * per, per * 8);
        pk.parmsID() = key_parms;
        vec.push_back(std::move(pk));
    }
}

// Evaluator::multiplyInplace (src/evaluator.cpp:355-383): size-2 x size-2 -> size 3.
int ref_multiply(void *h, uint64_t L, const uint64_t *a, const uint64_t *b, uint64_t *out) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        Ciphertext x, y;
        load_ct(s, x, L, 2, a, 1024.0);
        load_ct(s, y, L, 2, b, 1024.0);
        s.evaluator->multiplyInplace(x, y);
        std::memcpy(out, x.data(), 3 * L * s.n * 8);
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_multiply: %s\n", e.what()); return -1; }
}

// Evaluator::relinearizeInplace (src/evaluator.cpp:1113-1163): size 3 -> size 2.
int ref_relinearize(void *h, uint64_t L, const uint64_t *ct3, const uint64_t *key, uint64_t *out) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        Ciphertext x;
        load_ct(s, x, L, 3, ct3, 1024.0);
        RelinKeys rk;
        load_kswitch(s, rk, 0, key);
        s.evaluator->relinearizeInplace(x, rk);
        std::memcpy(out, x.data(), 2 * L * s.n * 8);
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_relinearize: %s\n", e.what()); return -1; }
}

// Evaluator::applyGaloisInplace (src/evaluator.cpp:2210-2250) with one Galois key
// stored at GaloisKeys::getIndex(elt) = (elt - 1) / 2.
int ref_apply_galois_keyswitch(void *h, uint64_t L, const uint64_t *ct, uint32_t elt, const uint64_t *key,
                               uint64_t *out) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        Ciphertext x;
        load_ct(s, x, L, 2, ct, 1024.0);
        GaloisKeys gk;
        load_kswitch(s, gk, (elt - 1) >> 1, key);
        s.evaluator->applyGaloisInplace(x, elt, gk);
        std::memcpy(out, x.data(), 2 * L * s.n * 8);
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_apply_galois: %s\n", e.what()); return -1; }
}

static SecretKey load_sk(const RefSession &s, const uint64_t *sk) {
    SecretKey key;
    key.data().parmsID() = parmsIDZero;
    key.data().resize(s.K * s.n);
    std::memcpy(key.data().data(), sk, s.K * s.n * 8);
    key.parmsID() = s.context->keyParmsID();
    return key;
}

// Decryptor::decrypt (src/decryptor.cpp:79-153) with an injected key-level secret key
// (NTT form, [K][N]).  CKKS: out = NTT-form plaintext [L][N]; BFV: out = [N] mod t.
int ref_decrypt(void *h, uint64_t L, const uint64_t *ct, const uint64_t *sk, uint64_t *out) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        Ciphertext x;
        load_ct(s, x, L, 2, ct, 1024.0);
        SecretKey key = load_sk(s, sk);
        Decryptor dec(*s.context, key);
        Plaintext p;
        dec.decrypt(x, p);
        if (s.scheme == SchemeType::ckks) {
            std::memcpy(out, p.data(), L * s.n * 8);
        } else {
            std::memset(out, 0, s.n * 8);
            std::memcpy(out, p.data(), p.coeffCount() * 8);
        }
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_decrypt: %s\n", e.what()); return -1; }
}

// Encryptor::encrypt (src/encryptor.cpp:172-233) with an injected key-level public key
// ([2][K][N], NTT form).  plain: CKKS [L][N] NTT form (scale 2^40); BFV [N] mod t, at
// the first data level (L = K - 1).  out: [2][L][N].  The reference draws u, e0, e1
// from its own generator, so this call is timed, never compared word for word.
int ref_encrypt(void *h, uint64_t L, const uint64_t *pk_words, const uint64_t *plain, uint64_t *out) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        auto key_parms = s.context->keyParmsID();
        PublicKey pk;
        pk.data().resize(*s.context, key_parms, 2);
        pk.data().isNttForm() = true;
        std::memcpy(pk.data().data(), pk_words, 2 * s.K * s.n * 8);
        pk.parmsID() = key_parms;
        Encryptor enc(*s.context, pk);
        Plaintext p;
        if (s.scheme == SchemeType::ckks) {
            p.parmsID() = parmsIDZero;
            p.resize(L * s.n);
            std::memcpy(p.data(), plain, L * s.n * 8);
            p.parmsID() = parms_for_level(s, L);
            p.scale() = 1099511627776.0;  // 2^40
        } else {
            if (L + 1 != s.K) throw std::invalid_argument("BFV encrypts at the first data level");
            p.resize(s.n);
            std::memcpy(p.data(), plain, s.n * 8);
        }
        Ciphertext c;
        enc.encrypt(p, c);
        std::memcpy(out, c.data(), 2 * L * s.n * 8);
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_encrypt: %s\n", e.what()); return -1; }
}

// KeyGenerator(context) + createRelinKeys (src/keygenerator.cpp:36-150): a fresh secret
// key, then its relinearization key, from the reference's own generator -- timed by the
// benchmark's CPU baseline; the key is discarded.
int ref_keygen_relin(void *h) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        KeyGenerator kg(*s.context);
        RelinKeys rk = kg.createRelinKeys();
        return rk.data().empty() ? -1 : 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_keygen_relin: %s\n", e.what()); return -1; }
}

// BatchEncoder::encode / decode (src/batchencoder.cpp:104-200)
int ref_batch_encode(void *h, const uint64_t *values, uint64_t count, uint64_t *plain) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        BatchEncoder enc(*s.context);
        Plaintext p;
        enc.encode(std::vector<uint64_t>(values, values + count), p);
        std::memset(plain, 0, s.n * 8);
        std::memcpy(plain, p.data(), p.coeffCount() * 8);
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_batch_encode: %s\n", e.what()); return -1; }
}

int ref_batch_decode(void *h, const uint64_t *plain, uint64_t *values) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        BatchEncoder enc(*s.context);
        Plaintext p;
        p.resize(s.n);
        std::memcpy(p.data(), plain, s.n * 8);
        std::vector<uint64_t> v;
        enc.decode(p, v);
        std::memcpy(values, v.data(), s.n * 8);
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_batch_decode: %s\n", e.what()); return -1; }
}

// CKKSEncoder::encode (src/ckks.cpp:98-277) of `count` complex values (re, im pairs)
// at the level with L primes; out: NTT-form plaintext [L][N].
int ref_ckks_encode(void *h, uint64_t L, const double *values, uint64_t count, double scale, uint64_t *out) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        CKKSEncoder enc(*s.context);
        std::vector<std::complex<double>> v(count);
        for (size_t i = 0; i < count; i++) v[i] = {values[2 * i], values[2 * i + 1]};
        Plaintext p;
        enc.encode(v, parms_for_level(s, L), scale, p);
        std::memcpy(out, p.data(), L * s.n * 8);
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_ckks_encode: %s\n", e.what()); return -1; }
}

// CKKSEncoder::decode (src/ckks.cpp:474-580): NTT-form plaintext [L][N] -> N/2 complex
int ref_ckks_decode(void *h, uint64_t L, const uint64_t *plain, double scale, double *values) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        CKKSEncoder enc(*s.context);
        Plaintext p;
        p.parmsID() = parmsIDZero;
        p.resize(L * s.n);
        std::memcpy(p.data(), plain, L * s.n * 8);
        p.parmsID() = parms_for_level(s, L);
        p.scale() = scale;
        std::vector<std::complex<double>> v;
        enc.decode(p, v);
        for (size_t i = 0; i < v.size(); i++) {
            values[2 * i] = v[i].real();
            values[2 * i + 1] = v[i].imag();
        }
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_ckks_decode: %s\n", e.what()); return -1; }
}

// Evaluator::rescaleToNext (CKKS) / modSwitchToNext (BFV) (src/evaluator.cpp:1165-1240).
int ref_rescale(void *h, uint64_t L, uint64_t size, const uint64_t *in, uint64_t *out) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        Ciphertext x, y;
        load_ct(s, x, L, size, in, 1024.0 * 1024.0 * 1024.0);
        if (s.scheme == SchemeType::ckks) s.evaluator->rescaleToNext(x, y);
        else s.evaluator->modSwitchToNext(x, y);
        std::memcpy(out, y.data(), size * (L - 1) * s.n * 8);
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_rescale: %s\n", e.what()); return -1; }
}

// Evaluator::addInplace / subInplace / negateInplace (src/evaluator.cpp:149-357) on
// ciphertexts of `size` polys at the level with L primes: op 0 = a + b, 1 = a - b,
// 2 = -a.
int ref_ct_binary(void *h, uint64_t L, uint64_t size, int op, const uint64_t *a, const uint64_t *b, uint64_t *out) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        Ciphertext x, y;
        load_ct(s, x, L, size, a, 1024.0);
        if (op == 2) {
            s.evaluator->negateInplace(x);
        } else {
            load_ct(s, y, L, size, b, 1024.0);
            if (op == 0) s.evaluator->addInplace(x, y);
            else s.evaluator->subInplace(x, y);
        }
        std::memcpy(out, x.data(), size * L * s.n * 8);
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_ct_binary: %s\n", e.what()); return -1; }
}

// Evaluator::addPlainInplace / subPlainInplace / multiplyPlainInplace (src/evaluator.cpp:
// 1603-1955): op 0 = add, 1 = sub, 2 = multiply.  BFV plain: [N] mod t, ciphertext in
// coefficient form; CKKS plain: NTT form [L][N] at the ciphertext's level.
int ref_plain_op(void *h, uint64_t L, uint64_t size, int op, const uint64_t *ct, const uint64_t *plain,
                 uint64_t *out) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        Ciphertext x;
        load_ct(s, x, L, size, ct, 1024.0);
        Plaintext p;
        if (s.scheme == SchemeType::ckks) {
            p.parmsID() = parmsIDZero;
            p.resize(L * s.n);
            std::memcpy(p.data(), plain, L * s.n * 8);
            p.parmsID() = parms_for_level(s, L);
            p.scale() = 1024.0;
        } else {
            p.resize(s.n);
            std::memcpy(p.data(), plain, s.n * 8);
        }
        if (op == 0) s.evaluator->addPlainInplace(x, p);
        else if (op == 1) s.evaluator->subPlainInplace(x, p);
        else s.evaluator->multiplyPlainInplace(x, p);
        std::memcpy(out, x.data(), size * L * s.n * 8);
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_plain_op: %s\n", e.what()); return -1; }
}

// Plaintext(const std::string &) and Plaintext::to_string (src/plaintext.cpp:91-200,
// src/plaintext.h:491-498): parses `hex`, writes its coefficients (at most *count) and
// their number to *count, and its to_string() into str (at most str_cap bytes, NUL
// terminated).  Returns -1 when the reference rejects the string (std::invalid_argument).
int ref_plaintext_hex(const char *hex, uint64_t *coeffs, uint64_t *count, char *str, uint64_t str_cap) {
    try {
        Plaintext p(hex);
        const uint64_t cap = *count;
        *count = p.coeffCount();
        for (uint64_t i = 0; i < *count && i < cap; i++) coeffs[i] = p[i];
        const std::string s = p.to_string();
        std::snprintf(str, str_cap, "%s", s.c_str());
        return 0;
    } catch (const std::invalid_argument &) { return -1; }
    catch (...) { return -2; }
}

// EncryptionParameters::parmsID (src/encryptionparams.cpp:118-156) of the chain level
// with L primes; out: 4 words.
int ref_parms_id(void *h, uint64_t L, uint64_t *out) {
    try {
        auto &s = *static_cast<RefSession *>(h);
        ParmsID id = L == s.K ? s.context->keyParmsID() : parms_for_level(s, L);
        for (int i = 0; i < 4; i++) out[i] = id[i];
        return 0;
    } catch (const std::exception &e) { fprintf(stderr, "ref_parms_id: %s\n", e.what()); return -1; }
}

}  // extern "C"
