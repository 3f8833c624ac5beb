// ref_digits.cpp -- TEST INFRASTRUCTURE ONLY.  Exposes the reference library's
// own gadget decomposition (compiled from its sources by oracle/Makefile into
// oracle/_ref/) through a C interface, so the oracle's restatement can be
// checked against the real thing.  Built twice: -DTORUS32 and -DTORUS56.
#include <cstdint>
#include <vector>
#include <thread>
#include <atomic>
#include "yatfhe/yatfhe_parameters.h"
#include "yatfhe/gadget_decomposition.h"

extern "C" {

int32_t ref_torus_bits(void) { return YatfheParameters{}.torusBits; }

// decomposeRow (gadget_decomposition.h:189-236): out is [l][N].
int32_t ref_decompose_row(const int64_t* in, int32_t N, int32_t l, int32_t radixBits, int64_t* out) {
    YatfheParameters param{};
    param.setRadixBits(radixBits);
    std::vector<Torus> row(in, in + N);
    std::vector<std::vector<Torus>> o(l, std::vector<Torus>(N));
    std::vector<Torus*> ptr(l);
    for (int j = 0; j < l; j++) ptr[j] = o[j].data();
    try {
        decomposeRow(ptr.data(), row.data(), N, l, param);
    } catch (...) {
        return -1;
    }
    for (int j = 0; j < l; j++)
        for (int i = 0; i < N; i++) out[static_cast<size_t>(j) * N + i] = o[j][i];
    return 0;
}

// signedGadgetDecompositionKs (gadget_decomposition.cpp:53-56).
void ref_decompose_ks(int64_t x, int32_t l, int32_t ksRadixBits, int32_t ksWidthBits, int64_t* out) {
    YatfheParameters param{};
    param.ksRadixBits = ksRadixBits;
    param.ksWidthBits = ksWidthBits;
    DecomposedData d{l};
    signedGadgetDecompositionKs(d, static_cast<Torus>(x), param);
    for (int j = 0; j < l; j++) out[j] = static_cast<int64_t>(d.value[j]) * d.sign;
}

// Order-dependent hash of the digits of x = lo + i*stride (i < count, low
// torusBits bits, sign-extended), via decomposeRow in rows of 4096, threaded.
// The oracle exports the same hash of its own digits.
uint64_t ref_digits_hash(int32_t l, int32_t radixBits, uint64_t lo, uint64_t stride, uint64_t count, int32_t threads) {
    YatfheParameters base{};
    const int w = base.torusBits;
    const uint64_t chunk = 4096;
    const uint64_t nChunks = (count + chunk - 1) / chunk;
    std::vector<uint64_t> part(nChunks, 0);
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        YatfheParameters param{};
        param.setRadixBits(radixBits);
        std::vector<Torus> row(chunk);
        std::vector<std::vector<Torus>> o(l, std::vector<Torus>(chunk));
        std::vector<Torus*> ptr(l);
        for (int j = 0; j < l; j++) ptr[j] = o[j].data();
        for (;;) {
            const uint64_t c = next.fetch_add(1);
            if (c >= nChunks) break;
            const uint64_t s = c * chunk;
            const int n = static_cast<int>(std::min(chunk, count - s));
            for (int i = 0; i < n; i++) {
                const uint64_t u = lo + (s + i) * stride;
                row[i] = static_cast<Torus>(static_cast<int64_t>(u << (64 - w)) >> (64 - w));
            }
            decomposeRow(ptr.data(), row.data(), n, l, param);
            uint64_t h = 0x9E3779B97F4A7C15ULL ^ c;
            for (int i = 0; i < n; i++)
                for (int j = 0; j < l; j++) {
                    h ^= static_cast<uint64_t>(static_cast<int64_t>(o[j][i])) + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
                }
            part[c] = h;
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < (threads > 0 ? threads : 1); t++) ts.emplace_back(work);
    for (auto& t : ts) t.join();
    uint64_t h = 0;
    for (auto x : part) h = h * 0x100000001B3ULL ^ x;
    return h;
}

}  // extern "C"
