// bench_facade.cpp -- times the plain drop-in path: the troyn:: class API (include/troyhip.hpp)
// used one ciphertext at a time, as a reference program calls EvaluatorCuda
// (evaluator.multiply(a, b, out); evaluator.relinearizeInplace(out, rk)), on the headline
// shape: CKKS, N=2^16, CoeffModulus::Create(65536, {60 x 25}), device-generated keys,
// encryptions of encoded random slots.  Prints one JSON line.
//   bench_facade [ops] [warmup]
//   bench_facade io [reps]   -- the wire formats and key generation on the same shape:
//     Ciphertext::save / load of one level-24 ciphertext (2 x 24 limbs) through a
//     std::stringstream, RelinKeys::save / load, KeyGenerator::createRelinKeys and
//     createGaloisKeys (one step); each loaded object is saved again and must give the
//     same bytes.
#include "troyhip.hpp"
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <sstream>

namespace troyn = troyhip;

namespace {

// A stream buffer over one preallocated block: a save appends to it, a load reads back
// what was written -- the timed calls measure the library, not a stream growing or a
// string being copied.
struct BlockBuf : std::streambuf {
    std::vector<char> mem;
    explicit BlockBuf(size_t cap) : mem(cap) { for_write(); }
    void for_write() { setp(mem.data(), mem.data() + mem.size()); }
    size_t written() const { return (size_t)(pptr() - mem.data()); }
    void for_read(size_t len) { setg(mem.data(), mem.data(), mem.data() + len); }
    std::string bytes(size_t len) const { return std::string(mem.data(), len); }
};

template <class F>
double ms_per_call(F &&f, int reps) {
    f();
    (void)hipDeviceSynchronize();
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < reps; i++) f();
    (void)hipDeviceSynchronize();
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
}

int io_mode(int reps) {
    const size_t n = 65536;
    troyn::KernelProvider::initialize();
    troyn::EncryptionParameters parms(troyn::SchemeType::ckks);
    parms.setPolyModulusDegree(n);
    parms.setCoeffModulus(troyn::CoeffModulus::Create(n, std::vector<int>(25, 60)));
    troyn::SEALContext context(parms, true, troyn::SecurityLevel::none);
    troyn::KeyGenerator keygen(context);
    troyn::Encryptor encryptor(context, keygen.createPublicKey());
    troyn::CKKSEncoder encoder(context);
    std::vector<std::complex<double>> v(encoder.slotCount(), {0.25, -0.5});
    troyn::Plaintext p;
    encoder.encode(v, std::pow(2.0, 40), p);
    troyn::Ciphertext ct;
    encryptor.encrypt(p, ct);
    BlockBuf cb((size_t)ct.size() * ct.coeffModulusSize() * n * 8 + 4096);
    std::ostream cos(&cb);
    const double save_ms = ms_per_call([&] {
        cb.for_write();
        ct.save(cos);
    }, reps);
    const size_t ct_len = cb.written();
    const std::string blob = cb.bytes(ct_len);
    troyn::Ciphertext back;
    std::istream cis(&cb);
    const double load_ms = ms_per_call([&] {
        cb.for_read(ct_len);
        cis.clear();
        back.load(cis, context);
    }, reps);
    std::ostringstream again;
    back.save(again);
    if (again.str() != blob) {
        std::fprintf(stderr, "ciphertext save -> load -> save changed the bytes\n");
        return 1;
    }
    troyn::RelinKeys rk;
    const int kreps = reps > 4 ? reps / 4 : 1;
    const double rk_ms = ms_per_call([&] { keygen.createRelinKeys(rk); }, kreps);
    troyn::GaloisKeys gk;
    const double gk_ms = ms_per_call([&] { keygen.createGaloisKeys(std::vector<int>{1}, gk); }, kreps);
    const size_t K = 25;
    BlockBuf kb((K - 1) * (2 * K * n * 8 + 4096) + 4096);
    std::ostream kos(&kb);
    const double ksave_ms = ms_per_call([&] {
        kb.for_write();
        rk.save(kos);
    }, kreps);
    const size_t k_len = kb.written();
    const std::string kblob = kb.bytes(k_len);
    troyn::RelinKeys rk2;
    std::istream kis(&kb);
    const double kload_ms = ms_per_call([&] {
        kb.for_read(k_len);
        kis.clear();
        rk2.load(kis);
    }, kreps);
    std::ostringstream kagain;
    rk2.save(kagain);
    if (kagain.str() != kblob) {
        std::fprintf(stderr, "relinearization key save -> load -> save changed the bytes\n");
        return 1;
    }
    std::printf("{\"ct_bytes\": %zu, \"ct_save_ms\": %.4f, \"ct_load_ms\": %.4f, \"relin_key_bytes\": %zu, "
                "\"relin_key_save_ms\": %.3f, \"relin_key_load_ms\": %.3f, \"relin_keygen_ms\": %.3f, "
                "\"galois_keygen_ms\": %.3f, \"reps\": %d, \"key_reps\": %d}\n",
                blob.size(), save_ms, load_ms, kblob.size(), ksave_ms, kload_ms, rk_ms, gk_ms, reps, kreps);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc > 1 && std::strcmp(argv[1], "io") == 0) return io_mode(argc > 2 ? std::atoi(argv[2]) : 8);
    const int ops = argc > 1 ? std::atoi(argv[1]) : 128;
    const int warmup = argc > 2 ? std::atoi(argv[2]) : 16;
    const size_t n = 65536;
    troyn::KernelProvider::initialize();
    troyn::EncryptionParameters parms(troyn::SchemeType::ckks);
    parms.setPolyModulusDegree(n);
    parms.setCoeffModulus(troyn::CoeffModulus::Create(n, std::vector<int>(25, 60)));
    // 1500 bits exceed the 128-bit security table at N=2^16 (the reference's own
    // benchmarks use SecurityLevel::none the same way)
    troyn::SEALContext context(parms, true, troyn::SecurityLevel::none);
    troyn::KeyGenerator keygen(context);
    troyn::RelinKeys rk = keygen.createRelinKeys();
    troyn::Encryptor encryptor(context, keygen.createPublicKey());
    troyn::CKKSEncoder encoder(context);
    troyn::Evaluator evaluator(context);
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    const int pairs = 8;
    std::vector<troyn::Ciphertext> a(pairs), b(pairs), out(pairs);
    for (int i = 0; i < pairs; i++) {
        std::vector<std::complex<double>> v(encoder.slotCount());
        for (auto &x : v) x = {u(rng), u(rng)};
        troyn::Plaintext p;
        encoder.encode(v, std::pow(2.0, 40), p);
        encryptor.encrypt(p, a[i]);
        encryptor.encrypt(p, b[i]);
    }
    auto op = [&](int i) {
        evaluator.multiply(a[i % pairs], b[i % pairs], out[i % pairs]);
        evaluator.relinearizeInplace(out[i % pairs], rk);
    };
    for (int i = 0; i < warmup; i++) op(i);
    (void)hipDeviceSynchronize();
    const size_t allocs = troyn::MemoryPool::deviceAllocations();
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < ops; i++) op(i);
    (void)hipDeviceSynchronize();
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("{\"ops_s\": %.3f, \"ms_per_op\": %.5f, \"ops\": %d, \"device_allocations_in_timed_loop\": %zu}\n",
                ops / s, s * 1e3 / ops, ops, troyn::MemoryPool::deviceAllocations() - allocs);
    return 0;
}
