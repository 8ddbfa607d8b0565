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

}  // extern "C"
