// Host check of the C++ API's BLAKE2b and parms_id (include/troyhip.hpp, no GPU).
//   parms_hash_check blake2b <text>              -> hex BLAKE2b-512 and -256 of the text
//   parms_hash_check parms <scheme> <n> <t> <q0> <q1> ...  -> the 4 parms_id words
//   parms_hash_check hexpoly <string>             -> hexadecimal plaintext parse and print
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "troyhip.hpp"

int main(int argc, char **argv) {
    if (argc >= 3 && !std::strcmp(argv[1], "blake2b")) {
        const char *text = argv[2];
        for (size_t len : {64, 32}) {
            uint8_t out[64];
            troyhip::detail::blake2b(reinterpret_cast<const uint8_t *>(text), std::strlen(text), out, len);
            for (size_t i = 0; i < len; i++) std::printf("%02x", out[i]);
            std::printf("\n");
        }
        return 0;
    }
    if (argc >= 5 && !std::strcmp(argv[1], "parms")) {
        const uint64_t scheme = std::strtoull(argv[2], nullptr, 10), n = std::strtoull(argv[3], nullptr, 10),
                       t = std::strtoull(argv[4], nullptr, 10);
        std::vector<uint64_t> q;
        for (int i = 5; i < argc; i++) q.push_back(std::strtoull(argv[i], nullptr, 10));
        const auto h = troyhip::detail::parms_hash(scheme, n, q, t);
        std::printf("%llu %llu %llu %llu\n", (unsigned long long)h[0], (unsigned long long)h[1],
                    (unsigned long long)h[2], (unsigned long long)h[3]);
        return 0;
    }
    if (argc >= 3 && !std::strcmp(argv[1], "hexpoly")) {
        // hexpoly <string>: "invalid", or the coefficient count, the coefficients and
        // the to_string() form, one per line
        try {
            const troyhip::host::Plaintext p(argv[2]);
            std::printf("%zu\n", p.coeffCount());
            for (size_t i = 0; i < p.coeffCount(); i++) std::printf("%llu\n", (unsigned long long)p[i]);
            std::printf("%s\n", p.to_string().c_str());
        } catch (const std::invalid_argument &) {
            std::printf("invalid\n");
        }
        return 0;
    }
    return 2;
}
