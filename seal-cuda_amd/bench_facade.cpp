// bench_facade.cpp -- times the plain drop-in path: the troyn:: class API (include/troyhip.hpp)
// used one ciphertext at a time, as a reference program calls EvaluatorCuda
// (evaluator.multiply(a, b, out); evaluator.relinearizeInplace(out, rk)), on the headline
// shape: CKKS, N=2^16, CoeffModulus::Create(65536, {60 x 25}), device-generated keys,
// encryptions of encoded random slots.  Prints one JSON line.
//   bench_facade [ops] [warmup]
#include "troyhip.hpp"
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>

namespace troyn = troyhip;

int main(int argc, char **argv) {
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
