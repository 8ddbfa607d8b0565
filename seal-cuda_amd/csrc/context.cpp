// context.cpp -- builds and uploads the per-parameter-set tables.
#include "context.h"
#include "ckks_encoder.h"
#include <cstring>

namespace shc {

DevPrime upload_prime(Context &c, uint64_t q) {
    // the NTT's lazy ranges reach 8q, which must fit 64 bits
    if (q >> 61) fail(ERR_INVALID_ARGUMENT, "modulus above 61 bits");
    const hm::NttTables t = hm::make_ntt_tables(c.logn, q);
    DevPrime p{};
    p.q = t.q;
    p.two_q = t.q << 1;
    p.neg_four_q = 0 - (t.q << 2);
    p.neg_eight_q = 0 - (t.q << 3);
    hm::barrett_ratio(t.q, p.ratio_lo, p.ratio_hi);
    p.inv_n = t.inv_n;
    p.inv_n_shoup = t.inv_n_shoup;
    const size_t bytes = t.fwd.size() * sizeof(uint64_t);
    DevBuf buf(2 * bytes);
    SHC_HIP(hipMemcpy(buf.ptr, t.fwd.data(), bytes, hipMemcpyHostToDevice));
    SHC_HIP(hipMemcpy(static_cast<char *>(buf.ptr) + bytes, t.inv.data(), bytes, hipMemcpyHostToDevice));
    p.fwd = buf.as<const ulonglong2>();
    p.inv = reinterpret_cast<const ulonglong2 *>(static_cast<char *>(buf.ptr) + bytes);
    c.table_bufs.push_back(std::move(buf));
    return p;
}

DevPrime Context::aux_prime(uint64_t q) {
    auto it = aux.find(q);
    if (it != aux.end()) return it->second;
    for (size_t i = 0; i < moduli.size(); i++)
        if (moduli[i] == q) return aux[q] = host_primes[i];
    return aux[q] = upload_prime(*this, q);
}

const BehzLevel &Context::behz(size_t level) const {
    if (level >= behz_levels.size() || !behz_levels[level])
        fail(ERR_LOGIC, "BEHZ tables are only built for BFV data levels");
    return *behz_levels[level];
}

template <class T>
static DevBuf upload(const std::vector<T> &v) {
    DevBuf b(v.size() * sizeof(T));
    if (!v.empty()) SHC_HIP(hipMemcpy(b.ptr, v.data(), b.bytes, hipMemcpyHostToDevice));
    return b;
}

std::unique_ptr<Context> make_context(Scheme scheme, int logn, const std::vector<uint64_t> &moduli, uint64_t plain) {
    if (logn < 1 || logn > 17) fail(ERR_INVALID_ARGUMENT, "poly_modulus_degree must be a power of two in [2, 131072]");
    // At most 64 primes below 2^60 (the reference's bounds, src/utils/defines.h:19-27):
    // the key switch relies on them (16q fits a word in its ModUp transforms; 63 products below 4q^2 fit 128 bits).
    if (moduli.empty() || moduli.size() > 64) fail(ERR_INVALID_ARGUMENT, "coeff_modulus size is invalid");
    for (size_t i = 0; i < moduli.size(); i++) {
        if (moduli[i] >> 60 || moduli[i] < 2) fail(ERR_INVALID_ARGUMENT, "coeff_modulus bit count is invalid");
        for (size_t j = 0; j < i; j++)
            if (moduli[i] == moduli[j]) fail(ERR_INVALID_ARGUMENT, "coeff_modulus must be pairwise coprime");
    }
    if (scheme == SCHEME_BFV && (plain < 2 || plain >> 60)) fail(ERR_INVALID_ARGUMENT, "plain_modulus is invalid");
    auto c = std::make_unique<Context>();
    SHC_HIP(hipGetDevice(&c->device));
    c->scheme = scheme;
    c->logn = logn;
    c->n = uint64_t(1) << logn;
    c->moduli = moduli;
    c->plain = plain;
    for (uint64_t q : moduli) c->host_primes.push_back(upload_prime(*c, q));

    const size_t K = moduli.size();
    c->d_primes = upload(c->host_primes);
    if (K >= 2) {
        std::vector<DevPrime> ks((K - 1) * K);
        for (size_t L = 1; L < K; L++) {
            DevPrime *row = ks.data() + (L - 1) * K;
            for (size_t i = 0; i < L; i++) row[i] = c->host_primes[i];
            row[L] = c->host_primes[K - 1];
        }
        c->d_keyswitch_primes = upload(ks);
    }
    // divide-and-round constants: RNSTool::inv_q_last_mod_q_ (src/utils/rns.cpp:778-790)
    // for every level, and the special prime's for key switching
    auto divconst = [](uint64_t p, uint64_t qi) {
        DivRoundConst d{};
        d.inv = hm::invmod(p % qi, qi);
        if (!d.inv) fail(ERR_LOGIC, "invalid rns bases");
        d.inv_shoup = hm::shoup(d.inv, qi);
        d.half_mod = (p >> 1) % qi;
        d.p_mod = p % qi;
        return d;
    };
    std::vector<DivRoundConst> ksc(K), rsc((K + 1) * K);
    for (size_t i = 0; i + 1 < K; i++) ksc[i] = divconst(moduli[K - 1], moduli[i]);
    for (size_t L = 2; L <= K; L++)
        for (size_t i = 0; i + 1 < L; i++) rsc[L * K + i] = divconst(moduli[L - 1], moduli[i]);
    c->d_ks_consts = upload(ksc);
    c->d_rescale_consts = upload(rsc);
    if (K >= 2) {
        std::vector<uint64_t> f;
        for (size_t j = 0; j + 1 < K; j++) {
            const uint64_t w = moduli[K - 1] % moduli[j];
            f.push_back(w);
            f.push_back(hm::shoup(w, moduli[j]));
        }
        c->d_kswitch_factors = upload(f);
    }

    if (scheme == SCHEME_BFV) {
        for (uint64_t q : moduli)
            if (hm::invmod(plain % q, q) == 0) fail(ERR_INVALID_ARGUMENT, "plain_modulus must be coprime to coeff_modulus");
        // BEHZ tables for every data level (the first data level drops the special prime)
        const size_t top = K > 1 ? K - 1 : 1;
        c->behz_levels.resize(top + 1);
        for (size_t L = 1; L <= top; L++) {
            c->behz_levels[L] = std::make_unique<BehzLevel>();
            build_behz_level(*c, (uint32_t)L, *c->behz_levels[L]);
        }
        // batching needs t = 1 mod 2N (plain NTT tables, src/context.cpp:246-254)
        uint64_t root;
        if (hm::is_prime(plain) && hm::minimal_primitive_root(2 * c->n, plain, root)) {
            std::vector<DevPrime> tp{c->aux_prime(plain)};
            c->d_plain_prime = upload(tp);
        }
    }
    if (scheme == SCHEME_CKKS) build_ckks_tables(*c);
    return c;
}

}  // namespace shc
