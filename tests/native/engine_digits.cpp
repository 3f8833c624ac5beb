// Host build of the engine's device digit decompositions (digits_packed_add32,
// digits_packed_wide from ygpu_device.h), for the CPU tests that compare them
// with the oracle's zeroMeanDigits restatement.  Test infrastructure only.
//
// eng_digits_hash(l, b, w, lo, stride, count, threads) hashes the digits of the
// inputs sext_w(lo + i * stride), i < count, exactly as the oracle's
// or_digits_hash does (same chunking, same mixing), so equal hashes mean equal
// digit sequences.  Shapes: w = 32 with l * b <= 32 (the single-addition carry
// chain), w = 56 with l * b <= 32 (the top-32-bit form).
#include <hip/hip_runtime.h>
#include "ygpu_device.h"

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <thread>
#include <vector>

using namespace ygpu;

namespace {

template <int W, int L, int RB>
void digits_one(int64_t x, int64_t* d) {
    uint32_t w0, w1;
    if constexpr (W == 32) digits_packed_add32<L, RB>(static_cast<uint32_t>(x), w0, w1);
    else digits_packed_wide<W, L, RB>(x, w0, w1);
    for (int lvl = 0; lvl < L; lvl++) {
        if constexpr (RB <= 7) d[lvl] = static_cast<int8_t>((lvl < 4 ? w0 : w1) >> (8 * (lvl & 3)));
        else d[lvl] = static_cast<int16_t>((lvl < 2 ? w0 : w1) >> (16 * (lvl & 1)));
    }
}

template <int W, int L, int RB>
uint64_t hash_range(uint64_t lo, uint64_t stride, uint64_t count, int threads) {
    const uint64_t chunk = 4096;
    const uint64_t nChunks = (count + chunk - 1) / chunk;
    std::vector<uint64_t> part(nChunks, 0);
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        int64_t d[8];
        for (;;) {
            const uint64_t c = next.fetch_add(1);
            if (c >= nChunks) break;
            const uint64_t s = c * chunk;
            const int n = static_cast<int>(std::min(chunk, count - s));
            uint64_t h = 0x9E3779B97F4A7C15ULL ^ c;
            for (int i = 0; i < n; i++) {
                const int64_t x = sext_t<W>(static_cast<int64_t>(lo + (s + i) * stride));
                digits_one<W, L, RB>(x, d);
                for (int j = 0; j < L; j++) h ^= static_cast<uint64_t>(d[j]) + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
            }
            part[c] = h;
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < std::max(1, threads); t++) ts.emplace_back(work);
    for (auto& t : ts) t.join();
    uint64_t h = 0;
    for (auto x : part) h = h * 0x100000001B3ULL ^ x;
    return h;
}

}  // namespace

extern "C" {
// 0 for a shape this helper does not instantiate.
uint64_t eng_digits_hash(int32_t l, int32_t b, int32_t w, uint64_t lo, uint64_t stride, uint64_t count, int32_t threads) {
    if (w == 32 && l == 3 && b == 7) return hash_range<32, 3, 7>(lo, stride, count, threads);
    if (w == 32 && l == 3 && b == 10) return hash_range<32, 3, 10>(lo, stride, count, threads);
    if (w == 32 && l == 4 && b == 7) return hash_range<32, 4, 7>(lo, stride, count, threads);
    if (w == 32 && l == 2 && b == 8) return hash_range<32, 2, 8>(lo, stride, count, threads);
    if (w == 56 && l == 3 && b == 10) return hash_range<56, 3, 10>(lo, stride, count, threads);
    if (w == 56 && l == 2 && b == 14) return hash_range<56, 2, 14>(lo, stride, count, threads);
    return 0;
}
}
