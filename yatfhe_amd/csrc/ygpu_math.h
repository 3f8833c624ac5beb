// ygpu_math.h -- host-side exact modular arithmetic, transform tables, seeded RNG.
//
// Host code only (keys, tables, key generation).  The device arithmetic lives in
// ygpu_device.h.
#pragma once
#include <cstdint>
#include <cmath>
#include <vector>
#include <stdexcept>

namespace ygpu {

typedef unsigned __int128 u128;
typedef __int128 i128;

// NTT primes of the reference (torus.h:151, torus.h:157).
constexpr uint64_t Q_49P_T32 = 562941363486721ULL;      // 2^49 - 2^33 + 1
constexpr uint64_t Q_61P_T56 = 1945555039024054273ULL;  // 27 * 2^56 + 1

// Primes for the exact double-precision transform.  Every value the device
// transform holds is an integer of magnitude < 2^53, and the bound analysis
// (DESIGN) needs 2^53 / p >= 16, i.e. p < 2^49.  TORUS32 transforms mod the
// reference prime itself; TORUS56 transforms mod two such primes whose product
// (~2^98) exceeds twice the largest external-product coefficient (~2^83), and
// recovers the value mod Q_61P_T56 by CRT.
constexpr uint64_t P_AUX_0 = Q_49P_T32;
constexpr uint64_t P_AUX_1 = 562949952700417ULL;       // largest prime < 2^49 that is 1 mod 2^16 (other than P_AUX_0)

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return static_cast<uint64_t>((static_cast<u128>(a) * b) % q); }
inline uint64_t powmod(uint64_t a, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = mulmod(r, a, q);
        a = mulmod(a, a, q);
        e >>= 1;
    }
    return r;
}
inline uint64_t invmod(uint64_t a, uint64_t q) { return powmod(a, q - 2, q); }

inline bool is_prime_u64(uint64_t n) {
    if (n < 2) return false;
    for (uint64_t p : {2ULL, 3ULL, 5ULL, 7ULL, 11ULL, 13ULL, 17ULL, 19ULL, 23ULL, 29ULL, 31ULL, 37ULL}) {
        if (n % p == 0) return n == p;
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) { d >>= 1; s++; }
    for (uint64_t a : {2ULL, 3ULL, 5ULL, 7ULL, 11ULL, 13ULL, 17ULL, 19ULL, 23ULL, 29ULL, 31ULL, 37ULL}) {
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int r = 1; r < s; r++) {
            x = mulmod(x, x, n);
            if (x == n - 1) { comp = false; break; }
        }
        if (comp) return false;
    }
    return true;
}

inline uint64_t generator(uint64_t q) {
    std::vector<uint64_t> f;
    uint64_t m = q - 1;
    if ((m & 1) == 0) { f.push_back(2); while ((m & 1) == 0) m >>= 1; }
    for (uint64_t p = 3; p * p <= m; p += 2) {
        if (m % p == 0) { f.push_back(p); while (m % p == 0) m /= p; }
    }
    if (m > 1) f.push_back(m);
    for (uint64_t g = 2;; g++) {
        bool ok = true;
        for (auto p : f) if (powmod(g, (q - 1) / p, q) == 1) { ok = false; break; }
        if (ok) return g;
    }
}

inline int bit_reverse(int x, int bits) {
    int r = 0;
    for (int i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

inline double centred_double(uint64_t x, uint64_t p) {
    const int64_t c = x > (p - 1) / 2 ? static_cast<int64_t>(x) - static_cast<int64_t>(p) : static_cast<int64_t>(x);
    return static_cast<double>(c);
}

// Negacyclic transform mod a prime p < 2^62 with the psi powers merged
// (Cooley-Tukey forward, natural -> bit-reversed).  The device transform is
// the same sequence of butterflies, so this defines the engine's NTT order.
struct HostNtt {
    uint64_t p = 0;
    int N = 0, logN = 0;
    std::vector<uint64_t> psi, ipsi;   // bit-reversed powers of psi, psi^-1
    uint64_t ninv = 0;
    HostNtt() = default;
    HostNtt(uint64_t p_, int N_) : p(p_), N(N_) {
        while ((1 << logN) < N) logN++;
        const uint64_t g = generator(p);
        const uint64_t root = powmod(g, (p - 1) / (2 * static_cast<uint64_t>(N)), p);
        if (powmod(root, N, p) != p - 1) throw std::runtime_error("no primitive 2N-th root of unity");
        const uint64_t iroot = invmod(root, p);
        psi.resize(N);
        ipsi.resize(N);
        for (int i = 0; i < N; i++) {
            const int e = bit_reverse(i, logN);
            psi[i] = powmod(root, e, p);
            ipsi[i] = powmod(iroot, e, p);
        }
        ninv = invmod(N, p);
    }
    void forward(uint64_t* a) const {
        for (int m = 1, t = N / 2; m < N; m <<= 1, t >>= 1) {
            for (int i = 0; i < m; i++) {
                const uint64_t w = psi[m + i];
                for (int j = 2 * i * t; j < 2 * i * t + t; j++) {
                    const uint64_t U = a[j];
                    const uint64_t V = mulmod(a[j + t], w, p);
                    a[j] = (U + V) % p;
                    a[j + t] = (U + p - V) % p;
                }
            }
        }
    }
    // Gentleman-Sande inverse (bit-reversed -> natural), including 1/N.
    void inverse(uint64_t* a) const {
        for (int m = N / 2, t = 1; m >= 1; m >>= 1, t <<= 1) {
            for (int i = 0; i < m; i++) {
                const uint64_t w = ipsi[m + i];
                for (int j = 2 * i * t; j < 2 * i * t + t; j++) {
                    const uint64_t U = a[j], V = a[j + t];
                    a[j] = (U + V) % p;
                    a[j + t] = mulmod((U + p - V) % p, w, p);
                }
            }
        }
        for (int j = 0; j < N; j++) a[j] = mulmod(a[j], ninv, p);
    }
};

// ------------------------------------------------------------------ RNG
// SplitMix64 streams: stream(seed, id) is an independent sequence, so keys are
// identical for any thread count.
inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t stream) : s(seed * 0xD1B54A32D192ED03ULL ^ (stream + 0x632BE59BD9B4E019ULL) * 0x8CB92BA72F3D8DD7ULL) {
        splitmix64(s);
    }
    uint64_t next() { return splitmix64(s); }
    // uniform in [0, bound) by rejection on the smallest covering power of two
    uint64_t below(uint64_t bound) {
        if (bound <= 1) return 0;
        const int bits = 64 - __builtin_clzll(bound - 1);
        const uint64_t mask = bits >= 64 ? ~0ULL : ((1ULL << bits) - 1);
        for (;;) {
            const uint64_t x = next() & mask;
            if (x < bound) return x;
        }
    }
    // uniform in (0, 1]
    double unit() { return (static_cast<double>(next() >> 11) + 1.0) * 0x1.0p-53; }
    // Gaussian with standard deviation sigma (Box-Muller), rounded to an integer
    int64_t gauss(double sigma) {
        const double u1 = unit(), u2 = unit();
        const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925 * u2);
        return static_cast<int64_t>(std::llround(z * sigma));
    }
};

inline int64_t sext(int64_t x, int w) {
    const int shift = 64 - w;
    return static_cast<int64_t>(static_cast<uint64_t>(x) << shift) >> shift;
}

}  // namespace ygpu
