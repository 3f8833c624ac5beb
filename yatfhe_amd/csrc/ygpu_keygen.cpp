// ygpu_keygen.cpp -- seeded key generation and LWE encryption on the host.
//
// The reference draws keys from an OS-seeded ChaCha20 (csprng.cpp:63-72) with
// no seed API, so its key bits cannot be reproduced; this generator follows the
// same encryption equations with a seeded, thread-count-independent stream per
// ciphertext row.  Noise is either Gaussian, rounded to integers, or the
// reference's own TUniform law (genTUniformNoise, numeric.cpp:47-55): integers
// in [-2^b, 2^b], the end points with probability 2^-(b+2) and every other
// point with probability 2^-(b+1).
//   genTlweKey / genTrlweKey        binary keys (tlwe.cpp, trlwe.cpp:70-81)
//   encryptTrgswMPNtt               (trgsw.cpp:36-51): row r < k encrypts 0 and adds
//                                   mu_lvl to its mask polynomial a_r; row k (cPrime)
//                                   encrypts mu_lvl in the body; masks uniform mod q
//                                   (trlwe.cpp:25-48 samples them in the NTT domain,
//                                   which is the same distribution)
//   genTlweKeySwitchingKey          (keyswitching.cpp:19-29): KSK[i][j] = LWE_s(z_i * 2^(w - (j+1) b))
//   genBootstrappingKey (classic)   (bootstrapping.cpp:40-44, 98-104; encryptTrgswNtt,
//                                   trgsw.cpp:281-284): torus-uniform masks, mu added in the
//                                   torus, the cPrime body reduced to the torus before it
//   genBootstrappingKeyWWL24        (bootstrapping.cpp:106-126): the succinct key, body-digit
//                                   rows only (torus-uniform masks, symEncTrlweMultiSampleNtt)
//                                   plus the TRLEV of s^2 (symEncTrlevWithKeyNtt, trlev.cpp:75-94)
#include <cstdint>
#include <cstring>
#include <vector>
#include <thread>
#include <atomic>
#include <string>
#include <algorithm>
#include <functional>
#include "yatfhe_gpu.h"
#include "ygpu_math.h"

using namespace ygpu;

namespace {

void par_for(int count, int threads, const std::function<void(int)>& fn) {
    std::atomic<int> next{0};
    auto work = [&]() {
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= count) break;
            fn(i);
        }
    };
    std::vector<std::thread> ts;
    const int nt = std::max(1, std::min(threads, count));
    for (int t = 1; t < nt; t++) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
}

inline uint64_t lift(int64_t v, uint64_t q) {
    int64_t m = v % static_cast<int64_t>(q);
    if (m < 0) m += static_cast<int64_t>(q);
    return static_cast<uint64_t>(m);
}

// A noise law in the units of the torus it is added to.
struct Noise {
    bool tuniform = false;
    double sigma = 0;  // Gaussian standard deviation
    int b = 0;         // TUniform bound exponent
    static Noise gaussian_log2(double sigma_log2_fraction, int torus_bits) {
        Noise z;
        z.sigma = std::pow(2.0, sigma_log2_fraction + torus_bits);
        return z;
    }
    static Noise tuniform_bound(int b) {
        Noise z;
        z.tuniform = true;
        z.b = b;
        return z;
    }
    // genTUniformNoise (numeric.cpp:47-55): a uniform (b+2)-bit candidate c,
    // (c >> 1) + (c & 1) - 2^b.
    int64_t draw(Rng& r) const {
        if (!tuniform) return r.gauss(sigma);
        const int bits = b + 2;
        const uint64_t mask = bits >= 64 ? ~0ULL : ((uint64_t{1} << bits) - 1);
        uint64_t c = r.next() & mask;
        c = (c >> 1) + (c & 1);
        return static_cast<int64_t>(c) - (int64_t{1} << b);
    }
};

// stream identifiers
constexpr uint64_t S_LWE_KEY = 1, S_GLWE_KEY = 2;
constexpr uint64_t S_BSK = 1ULL << 40, S_KSK = 2ULL << 40, S_ENC = 3ULL << 40;
constexpr uint64_t S_HALF = 4ULL << 40, S_S2 = 5ULL << 40, S_CLASSIC = 6ULL << 40;

int worker_threads(int threads) {
    return threads > 0 ? threads : static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
}

// genTlweKey / genTrlweKey: binary secret keys.
void gen_secret_keys(const ygpu_params& p, uint64_t seed, int32_t* lwe_key, int32_t* glwe_key) {
    Rng rl(seed, S_LWE_KEY);
    for (int i = 0; i < p.n; i++) lwe_key[i] = static_cast<int32_t>(rl.next() & 1);
    Rng rg(seed, S_GLWE_KEY);
    for (int i = 0; i < p.k * p.N; i++) glwe_key[i] = static_cast<int32_t>(rg.next() & 1);
}

// genTlweKeySwitchingKey (keyswitching.cpp:19-29): rows [i][j][n+1], i < kN.
void gen_ksk(const ygpu_params& p, uint64_t seed, const Noise& nl, const int32_t* lwe_key, const int32_t* glwe_key,
             int32_t* ksk, int nt) {
    const int n = p.n, t = p.ks_level;
    par_for(p.k * p.N * t, nt, [&](int idx) {
        const int i = idx / t, j = idx % t;
        Rng r(seed, S_KSK + static_cast<uint64_t>(idx));
        int32_t* out = ksk + static_cast<size_t>(idx) * (n + 1);
        uint32_t dot = 0;
        for (int u = 0; u < n; u++) {
            out[u] = static_cast<int32_t>(static_cast<uint32_t>(r.next()));
            if (lwe_key[u]) dot += static_cast<uint32_t>(out[u]);
        }
        const uint32_t m = static_cast<uint32_t>(glwe_key[i]) << (p.ks_width_bits - (j + 1) * p.ks_radix_bits);
        const uint32_t e = static_cast<uint32_t>(nl.draw(r));
        out[n] = static_cast<int32_t>(dot + e + m);
    });
}

// Negacyclic product in the torus, a * z mod (X^N + 1, 2^w) for a binary z,
// accumulated into acc (multTorusPolynomialAcc, polynomial.h): wrapping 64-bit
// sums, whose low w bits are the torus result.
void torus_mul_binary_acc(int N, int w, const int64_t* a, const int32_t* z, int64_t* acc) {
    std::vector<uint64_t> sum(N, 0);
    for (int m = 0; m < N; m++) {
        if (!z[m]) continue;
        for (int j = 0; j < N; j++) {
            const int d = j + m;
            const uint64_t v = static_cast<uint64_t>(a[j]);
            if (d < N) sum[d] += v; else sum[d - N] -= v;
        }
    }
    for (int j = 0; j < N; j++) acc[j] = sext(static_cast<int64_t>(static_cast<uint64_t>(acc[j]) + sum[j]), w);
}

int keygen_mp(const ygpu_params* params, uint64_t seed, const Noise& nl, const Noise& ng, int32_t* lwe_key,
              int32_t* glwe_key, uint64_t* bsk_coeff, int32_t* ksk, int32_t threads) {
    if (!params || !lwe_key || !glwe_key || !bsk_coeff || !ksk) return YGPU_ERR_INVALID_ARG;
    const ygpu_params p = *params;
    if (p.k < 1 || p.N < 2 || p.n < 1 || p.level < 1 || p.ks_level < 1) return YGPU_ERR_INVALID_ARG;
    const int N = p.N, K = p.k, K1 = p.k + 1, L = p.level, n = p.n, w = p.torus_bits;
    const uint64_t q = p.q_ntt;
    try {
        gen_secret_keys(p, seed, lwe_key, glwe_key);
        HostNtt ntt(q, N);
        // transformed GLWE key
        std::vector<std::vector<uint64_t>> zt(K, std::vector<uint64_t>(N));
        for (int c = 0; c < K; c++) {
            for (int j = 0; j < N; j++) zt[c][j] = static_cast<uint64_t>(glwe_key[c * N + j]);
            ntt.forward(zt[c].data());
        }
        const int nt = worker_threads(threads);
        // BSK rows: [i][lvl][row][col][N]
        par_for(n * L * K1, nt, [&](int idx) {
            const int i = idx / (L * K1);
            const int lvl = (idx / K1) % L;
            const int row = idx % K1;
            Rng r(seed, S_BSK + static_cast<uint64_t>(idx));
            uint64_t* out = bsk_coeff + static_cast<size_t>(idx) * K1 * N;
            std::vector<uint64_t> body(N, 0), tmp(N);
            for (int c = 0; c < K; c++) {
                uint64_t* a = out + static_cast<size_t>(c) * N;
                for (int j = 0; j < N; j++) a[j] = r.below(q);
                std::memcpy(tmp.data(), a, sizeof(uint64_t) * N);
                ntt.forward(tmp.data());
                for (int j = 0; j < N; j++) tmp[j] = mulmod(tmp[j], zt[c][j], q);
                ntt.inverse(tmp.data());
                for (int j = 0; j < N; j++) body[j] = (body[j] + tmp[j]) % q;
            }
            for (int j = 0; j < N; j++) body[j] = (body[j] + lift(ng.draw(r), q)) % q;
            // mu_lvl = s_i * 2^(w - (lvl+1) b) as a torus value, lifted mod q
            const int64_t mu = sext(static_cast<int64_t>(lwe_key[i]) << (w - (lvl + 1) * p.radix_bits), w);
            if (row == K) {
                body[0] = (body[0] + lift(mu, q)) % q;
            } else {
                uint64_t* a = out + static_cast<size_t>(row) * N;
                a[0] = (a[0] + lift(mu, q)) % q;
            }
            std::memcpy(out + static_cast<size_t>(K) * N, body.data(), sizeof(uint64_t) * N);
        });
        gen_ksk(p, seed, nl, lwe_key, glwe_key, ksk, nt);
    } catch (...) {
        return YGPU_ERR_INVALID_ARG;
    }
    return YGPU_OK;
}

int keygen_succinct(const ygpu_params* params, uint64_t seed, const Noise& nl, const Noise& ng, int32_t s2_level,
                    int32_t* lwe_key, int32_t* glwe_key, uint64_t* bsk_half, uint64_t* s2, int32_t* ksk, int32_t threads) {
    if (!params || !lwe_key || !glwe_key || !bsk_half || !s2 || !ksk) return YGPU_ERR_INVALID_ARG;
    const ygpu_params p = *params;
    if (p.k < 1 || p.N < 2 || p.n < 1 || p.level < 1 || p.ks_level < 1) return YGPU_ERR_INVALID_ARG;
    if (s2_level < 1 || s2_level * p.radix_bits > p.torus_bits) return YGPU_ERR_INVALID_ARG;
    const int N = p.N, K = p.k, K1 = p.k + 1, L = p.level, n = p.n, w = p.torus_bits;
    const uint64_t q = p.q_ntt;
    try {
        gen_secret_keys(p, seed, lwe_key, glwe_key);
        HostNtt ntt(q, N);
        std::vector<std::vector<uint64_t>> zt(K, std::vector<uint64_t>(N));
        for (int c = 0; c < K; c++) {
            for (int j = 0; j < N; j++) zt[c][j] = static_cast<uint64_t>(glwe_key[c * N + j]);
            ntt.forward(zt[c].data());
        }
        const int nt = worker_threads(threads);
        // cPrime rows [i][lvl][col][N] (genBootstrappingKeyWWL24, bootstrapping.cpp:106-126):
        // symEncTrlweMultiSampleNtt of mu = s_i 2^(w - (lvl+1) b) at coefficient 0
        par_for(n * L, nt, [&](int idx) {
            const int i = idx / L, lvl = idx % L;
            Rng r(seed, S_HALF + static_cast<uint64_t>(idx));
            uint64_t* out = bsk_half + static_cast<size_t>(idx) * K1 * N;
            std::vector<uint64_t> body(N, 0), tmp(N);
            for (int c = 0; c < K; c++) {
                uint64_t* a = out + static_cast<size_t>(c) * N;
                for (int j = 0; j < N; j++) a[j] = lift(sext(static_cast<int64_t>(r.next()), w), q);
                std::memcpy(tmp.data(), a, sizeof(uint64_t) * N);
                ntt.forward(tmp.data());
                for (int j = 0; j < N; j++) tmp[j] = mulmod(tmp[j], zt[c][j], q);
                ntt.inverse(tmp.data());
                for (int j = 0; j < N; j++) body[j] = (body[j] + tmp[j]) % q;
            }
            const int64_t mu = sext(static_cast<int64_t>(lwe_key[i]) << (w - (lvl + 1) * p.radix_bits), w);
            for (int j = 0; j < N; j++) {
                const int64_t bt = sext(ng.draw(r) + (j == 0 ? mu : 0), w);
                body[j] = (body[j] + lift(bt, q)) % q;
            }
            std::memcpy(out + static_cast<size_t>(K) * N, body.data(), sizeof(uint64_t) * N);
        });
        // TRLEV of s^2 [l2][col][N] (symEncTrlevWithKeyNtt, trlev.cpp:75-94): a
        // TRLWE of 0 in the torus (encTrlevSingleSample), then a_c -= s_c 2^(w - (l2+1) b)
        par_for(s2_level, nt, [&](int l2) {
            Rng r(seed, S_S2 + static_cast<uint64_t>(l2));
            std::vector<int64_t> a(static_cast<size_t>(K) * N), b(N);
            for (int j = 0; j < N; j++) b[j] = sext(ng.draw(r), w);
            for (int c = 0; c < K; c++) {
                int64_t* ac = &a[static_cast<size_t>(c) * N];
                for (int j = 0; j < N; j++) ac[j] = sext(static_cast<int64_t>(r.next()), w);
                torus_mul_binary_acc(N, w, ac, glwe_key + static_cast<size_t>(c) * N, b.data());
            }
            const int shift = w - (l2 + 1) * p.radix_bits;
            uint64_t* out = s2 + static_cast<size_t>(l2) * K1 * N;
            for (int c = 0; c < K; c++) {
                for (int j = 0; j < N; j++) {
                    const int64_t v = sext(a[static_cast<size_t>(c) * N + j] - (static_cast<int64_t>(glwe_key[c * N + j]) << shift), w);
                    out[static_cast<size_t>(c) * N + j] = lift(v, q);
                }
            }
            for (int j = 0; j < N; j++) out[static_cast<size_t>(K) * N + j] = lift(b[j], q);
        });
        gen_ksk(p, seed, nl, lwe_key, glwe_key, ksk, nt);
    } catch (...) {
        return YGPU_ERR_INVALID_ARG;
    }
    return YGPU_OK;
}

int keygen_classic(const ygpu_params* params, uint64_t seed, const Noise& nl, const Noise& ng, int32_t* lwe_key,
                   int32_t* glwe_key, uint64_t* bsk_coeff, int32_t* ksk, int32_t threads) {
    if (!params || !lwe_key || !glwe_key || !bsk_coeff || !ksk) return YGPU_ERR_INVALID_ARG;
    const ygpu_params p = *params;
    if (p.k < 1 || p.N < 2 || p.n < 1 || p.level < 1 || p.ks_level < 1) return YGPU_ERR_INVALID_ARG;
    const int N = p.N, K = p.k, K1 = p.k + 1, L = p.level, n = p.n, w = p.torus_bits;
    const uint64_t q = p.q_ntt;
    try {
        gen_secret_keys(p, seed, lwe_key, glwe_key);
        HostNtt ntt(q, N);
        std::vector<std::vector<uint64_t>> zt(K, std::vector<uint64_t>(N));
        for (int c = 0; c < K; c++) {
            for (int j = 0; j < N; j++) zt[c][j] = static_cast<uint64_t>(glwe_key[c * N + j]);
            ntt.forward(zt[c].data());
        }
        const int nt = worker_threads(threads);
        const uint64_t qhalf = (q - 1) / 2;
        // rows [i][lvl][row][col][N]: encryptTrgswNtt (trgsw.cpp:281-284) =
        // encZeroTrgswNtt (185-191: symEncTrlweSingleSampleNtt, torus-uniform
        // mask, body completed mod q) + addIntegerToTrgswNtt (227-246: mu_lvl
        // added in the torus to a_row[0] for row < k, re-transformed; for row k
        // to the torus value of the body, i.e. after applyIntt's centring)
        par_for(n * L * K1, nt, [&](int idx) {
            const int i = idx / (L * K1);
            const int lvl = (idx / K1) % L;
            const int row = idx % K1;
            Rng r(seed, S_CLASSIC + static_cast<uint64_t>(idx));
            uint64_t* out = bsk_coeff + static_cast<size_t>(idx) * K1 * N;
            std::vector<int64_t> at(static_cast<size_t>(K) * N);
            std::vector<uint64_t> body(N, 0), tmp(N);
            for (int j = 0; j < N; j++) body[j] = lift(sext(ng.draw(r), w), q);
            for (int c = 0; c < K; c++) {
                uint64_t* a = out + static_cast<size_t>(c) * N;
                for (int j = 0; j < N; j++) {
                    at[static_cast<size_t>(c) * N + j] = sext(static_cast<int64_t>(r.next()), w);
                    a[j] = lift(at[static_cast<size_t>(c) * N + j], q);
                }
                std::memcpy(tmp.data(), a, sizeof(uint64_t) * N);
                ntt.forward(tmp.data());
                for (int j = 0; j < N; j++) tmp[j] = mulmod(tmp[j], zt[c][j], q);
                ntt.inverse(tmp.data());
                for (int j = 0; j < N; j++) body[j] = (body[j] + tmp[j]) % q;
            }
            const int64_t mu = sext(static_cast<int64_t>(lwe_key[i]) << (w - (lvl + 1) * p.radix_bits), w);
            if (row < K) {
                uint64_t* a = out + static_cast<size_t>(row) * N;
                a[0] = lift(sext(at[static_cast<size_t>(row) * N] + mu, w), q);
            } else {
                for (int j = 0; j < N; j++) {
                    const int64_t centred = body[j] > qhalf ? static_cast<int64_t>(body[j]) - static_cast<int64_t>(q)
                                                            : static_cast<int64_t>(body[j]);
                    const int64_t bt = sext(centred + (j == 0 ? mu : 0), w);
                    body[j] = lift(bt, q);
                }
            }
            std::memcpy(out + static_cast<size_t>(K) * N, body.data(), sizeof(uint64_t) * N);
        });
        gen_ksk(p, seed, nl, lwe_key, glwe_key, ksk, nt);
    } catch (...) {
        return YGPU_ERR_INVALID_ARG;
    }
    return YGPU_OK;
}

int lwe_encrypt(int32_t n, const int32_t* lwe_key, uint64_t seed, const Noise& nl, int32_t count, const int32_t* mu,
                int32_t* out) {
    if (n < 1 || !lwe_key || count < 0 || (!mu && count) || (!out && count)) return YGPU_ERR_INVALID_ARG;
    for (int c = 0; c < count; c++) {
        Rng r(seed, S_ENC + static_cast<uint64_t>(c));
        int32_t* o = out + static_cast<size_t>(c) * (n + 1);
        uint32_t dot = 0;
        for (int u = 0; u < n; u++) {
            o[u] = static_cast<int32_t>(static_cast<uint32_t>(r.next()));
            if (lwe_key[u]) dot += static_cast<uint32_t>(o[u]);
        }
        const uint32_t e = static_cast<uint32_t>(nl.draw(r));
        o[n] = static_cast<int32_t>(dot + e + static_cast<uint32_t>(mu[c]));
    }
    return YGPU_OK;
}

// zeroMeanDigits (gadget_decomposition.h:75-95): the l balanced digits of the
// low w bits of x in base 2^b, rounding bias for the dropped bits, ties at
// B/2 resolved by the bit above, sign applied last.  out[level], level 0 the
// most significant.
void zero_mean_digits(int64_t x, int l, int b, int w, int64_t* out) {
    const uint64_t mask = w >= 64 ? ~0ULL : ((uint64_t{1} << w) - 1);
    const uint64_t u = static_cast<uint64_t>(x) & mask;
    const bool negative = (u >> (w - 1)) & 1;
    const int dropped = w - l * b;
    const uint64_t m = (negative ? (0ULL - u) & mask : u) + (dropped > 0 ? (uint64_t{1} << (dropped - 1)) : 0);
    const int64_t B = int64_t{1} << b, halfB = B >> 1;
    int64_t carry = 0;
    for (int level = l - 1; level >= 0; --level) {
        const int shift = w - (level + 1) * b;
        const int64_t digit = static_cast<int64_t>((m >> shift) & static_cast<uint64_t>(B - 1)) + carry;
        const int64_t tie = level > 0 ? static_cast<int64_t>((m >> (shift + 2 * b - 1)) & 1) : 0;
        carry = (digit > halfB) | ((digit == halfB) & tie);
        out[level] = (negative ? -1 : 1) * (carry ? digit - B : digit);
    }
}

constexpr uint64_t S_FIRST = 7ULL << 40;

// genBootstrappingKeyMPLazyPipe (bootstrapping.cpp:269-320), binary keys: rows
// of R = bsk.level = lApprox = p.level gadget levels, each cut into D = param.l
// digits.  The half RGSW rows cPrime[lvl] of s_1 .. s_{n-1} are those of the
// succinct key for the same seed (encryptTrgswMP with isHalf, trgsw.cpp:18-34,
// whose torus rows are symEncTrlweMultiSampleNtt's applyIntt of the mod-q
// body: centred, low w bits); each torus coefficient is cut into D zero-mean
// digits (signedGadgetDecomposition, gadget_decomposition.h:103-106) stored as
// the reference's TORUS32 Decomp, int8.  bskFirst encrypts v when s_0 = 1
// (symEncTrlweMultiSample) and 0 otherwise (symEncTrlweSingleSample), in torus
// arithmetic; s2 is the TRLEV of s^2 with D levels (s2Dft, TrlevDft{p, p.l}).
int keygen_lazy(const ygpu_params* params, uint64_t seed, const Noise& nl, const Noise& ng, int32_t digit_levels, const int32_t* v,
                int32_t* lwe_key, int32_t* glwe_key, int32_t* first, int8_t* digits_a, int8_t* digits_b, uint64_t* s2, int32_t* ksk,
                int32_t threads) {
    if (!params || !v || !lwe_key || !glwe_key || !first || !digits_a || !digits_b || !s2 || !ksk) return YGPU_ERR_INVALID_ARG;
    const ygpu_params p = *params;
    if (p.torus_bits != 32) return YGPU_ERR_UNSUPPORTED;  // digits are int8 (TORUS32's Decomp)
    if (p.n < 2 || p.k < 1 || p.level < 1 || p.radix_bits > 7) return YGPU_ERR_INVALID_ARG;
    const int D = digit_levels;
    if (D < 1 || D > 8 || D * p.radix_bits > p.torus_bits) return YGPU_ERR_INVALID_ARG;
    const int N = p.N, K = p.k, K1 = p.k + 1, L = p.level, n = p.n, w = p.torus_bits;
    const uint64_t q = p.q_ntt, qhalf = (q - 1) / 2;
    try {
        std::vector<uint64_t> half(static_cast<size_t>(n) * L * K1 * N);
        int rc = keygen_succinct(params, seed, nl, ng, D, lwe_key, glwe_key, half.data(), s2, ksk, threads);
        if (rc) return rc;
        const int nt = worker_threads(threads);
        par_for((n - 1) * L, nt, [&](int idx) {
            const int i = idx / L + 1, lvl = idx % L;  // component idx / L holds s_i
            const uint64_t* row = half.data() + (static_cast<size_t>(i) * L + lvl) * K1 * N;
            std::vector<int64_t> d(D);
            for (int c = 0; c < K1; c++) {
                for (int j = 0; j < N; j++) {
                    const uint64_t r = row[static_cast<size_t>(c) * N + j];
                    const int64_t t = sext(r > qhalf ? static_cast<int64_t>(r) - static_cast<int64_t>(q) : static_cast<int64_t>(r), w);
                    zero_mean_digits(t, D, p.radix_bits, w, d.data());
                    for (int l = 0; l < D; l++) {
                        const size_t cell = static_cast<size_t>(idx) * D + l;
                        if (c < K) digits_a[(cell * K + c) * N + j] = static_cast<int8_t>(d[l]);
                        else digits_b[cell * N + j] = static_cast<int8_t>(d[l]);
                    }
                }
            }
        });
        // bskFirst[0]: mask uniform over the torus, body e + mu + a s in the torus
        Rng r(seed, S_FIRST);
        std::vector<int64_t> a(static_cast<size_t>(K) * N), b(N);
        for (int j = 0; j < N; j++) b[j] = ng.draw(r) + (lwe_key[0] ? static_cast<int64_t>(v[j]) : 0);
        for (int c = 0; c < K; c++) {
            int64_t* ac = &a[static_cast<size_t>(c) * N];
            for (int j = 0; j < N; j++) ac[j] = sext(static_cast<int64_t>(r.next()), w);
            torus_mul_binary_acc(N, w, ac, glwe_key + static_cast<size_t>(c) * N, b.data());
        }
        for (int c = 0; c < K; c++)
            for (int j = 0; j < N; j++) first[static_cast<size_t>(c) * N + j] = static_cast<int32_t>(a[static_cast<size_t>(c) * N + j]);
        for (int j = 0; j < N; j++) first[static_cast<size_t>(K) * N + j] = static_cast<int32_t>(sext(b[j], w));
    } catch (...) {
        return YGPU_ERR_INVALID_ARG;
    }
    return YGPU_OK;
}

// genBootstrappingKeyMPLazyPipeAlt (bootstrapping.cpp:322-339), binary keys:
// encryptTrgswMP(bskPrime[i], s_i) with isHalf for every i < n (the succinct
// key's half rows for the same seed, in torus form) and the TRLEV of s^2 with
// D = param.l levels.  rows: int32 [n][R][k+1][N].
int keygen_lazy_alt(const ygpu_params* params, uint64_t seed, const Noise& nl, const Noise& ng, int32_t digit_levels, int32_t* lwe_key,
                    int32_t* glwe_key, int32_t* rows, uint64_t* s2, int32_t* ksk, int32_t threads) {
    if (!params || !lwe_key || !glwe_key || !rows || !s2 || !ksk) return YGPU_ERR_INVALID_ARG;
    const ygpu_params p = *params;
    if (p.torus_bits != 32) return YGPU_ERR_UNSUPPORTED;
    if (digit_levels < 1 || digit_levels > 8 || digit_levels * p.radix_bits > p.torus_bits) return YGPU_ERR_INVALID_ARG;
    const uint64_t q = p.q_ntt, qhalf = (q - 1) / 2;
    try {
        const size_t total = static_cast<size_t>(p.n) * p.level * (p.k + 1) * p.N;
        std::vector<uint64_t> half(total);
        int rc = keygen_succinct(params, seed, nl, ng, digit_levels, lwe_key, glwe_key, half.data(), s2, ksk, threads);
        if (rc) return rc;
        for (size_t i = 0; i < total; i++) {
            const uint64_t r = half[i];
            rows[i] = static_cast<int32_t>(sext(r > qhalf ? static_cast<int64_t>(r) - static_cast<int64_t>(q) : static_cast<int64_t>(r), 32));
        }
    } catch (...) {
        return YGPU_ERR_INVALID_ARG;
    }
    return YGPU_OK;
}

bool tuniform_ok(const ygpu_params* params, int lwe_b, int glwe_b) {
    return params && lwe_b >= 0 && lwe_b < 31 && glwe_b >= 0 && glwe_b < params->torus_bits - 1;
}

}  // namespace

extern "C" {

int ygpu_keygen(const ygpu_params* params, uint64_t seed, double sigma_log2_lwe, double sigma_log2_glwe, int32_t* lwe_key,
                int32_t* glwe_key, uint64_t* bsk_coeff, int32_t* ksk, int32_t threads) {
    if (!params) return YGPU_ERR_INVALID_ARG;
    return keygen_mp(params, seed, Noise::gaussian_log2(sigma_log2_lwe, 32), Noise::gaussian_log2(sigma_log2_glwe, params->torus_bits),
                     lwe_key, glwe_key, bsk_coeff, ksk, threads);
}

int ygpu_keygen_succinct(const ygpu_params* params, uint64_t seed, double sigma_log2_lwe, double sigma_log2_glwe, int32_t s2_level,
                         int32_t* lwe_key, int32_t* glwe_key, uint64_t* bsk_half, uint64_t* s2, int32_t* ksk, int32_t threads) {
    if (!params) return YGPU_ERR_INVALID_ARG;
    return keygen_succinct(params, seed, Noise::gaussian_log2(sigma_log2_lwe, 32),
                           Noise::gaussian_log2(sigma_log2_glwe, params->torus_bits), s2_level, lwe_key, glwe_key, bsk_half, s2, ksk,
                           threads);
}

int ygpu_keygen_classic(const ygpu_params* params, uint64_t seed, double sigma_log2_lwe, double sigma_log2_glwe, int32_t* lwe_key,
                        int32_t* glwe_key, uint64_t* bsk_coeff, int32_t* ksk, int32_t threads) {
    if (!params) return YGPU_ERR_INVALID_ARG;
    return keygen_classic(params, seed, Noise::gaussian_log2(sigma_log2_lwe, 32),
                          Noise::gaussian_log2(sigma_log2_glwe, params->torus_bits), lwe_key, glwe_key, bsk_coeff, ksk, threads);
}

int ygpu_keygen_tuniform(const ygpu_params* params, uint64_t seed, int32_t lwe_bound_bits, int32_t glwe_bound_bits, int32_t form,
                         int32_t s2_level, int32_t* lwe_key, int32_t* glwe_key, uint64_t* bsk_out, uint64_t* s2_out, int32_t* ksk,
                         int32_t threads) {
    if (!tuniform_ok(params, lwe_bound_bits, glwe_bound_bits)) return YGPU_ERR_INVALID_ARG;
    const Noise nl = Noise::tuniform_bound(lwe_bound_bits), ng = Noise::tuniform_bound(glwe_bound_bits);
    switch (form) {
        case YGPU_KEY_MP: return keygen_mp(params, seed, nl, ng, lwe_key, glwe_key, bsk_out, ksk, threads);
        case YGPU_KEY_CLASSIC: return keygen_classic(params, seed, nl, ng, lwe_key, glwe_key, bsk_out, ksk, threads);
        case YGPU_KEY_SUCCINCT: return keygen_succinct(params, seed, nl, ng, s2_level, lwe_key, glwe_key, bsk_out, s2_out, ksk, threads);
        default: return YGPU_ERR_INVALID_ARG;
    }
}

int ygpu_keygen_lazy(const ygpu_params* params, uint64_t seed, double sigma_log2_lwe, double sigma_log2_glwe, int32_t digit_levels,
                     const int32_t* v, int32_t* lwe_key, int32_t* glwe_key, int32_t* first, int8_t* digits_a, int8_t* digits_b,
                     uint64_t* s2, int32_t* ksk, int32_t threads) {
    if (!params) return YGPU_ERR_INVALID_ARG;
    return keygen_lazy(params, seed, Noise::gaussian_log2(sigma_log2_lwe, 32), Noise::gaussian_log2(sigma_log2_glwe, params->torus_bits),
                       digit_levels, v, lwe_key, glwe_key, first, digits_a, digits_b, s2, ksk, threads);
}

int ygpu_keygen_lazy_alt(const ygpu_params* params, uint64_t seed, double sigma_log2_lwe, double sigma_log2_glwe, int32_t digit_levels,
                         int32_t* lwe_key, int32_t* glwe_key, int32_t* rows, uint64_t* s2, int32_t* ksk, int32_t threads) {
    if (!params) return YGPU_ERR_INVALID_ARG;
    return keygen_lazy_alt(params, seed, Noise::gaussian_log2(sigma_log2_lwe, 32), Noise::gaussian_log2(sigma_log2_glwe, params->torus_bits),
                           digit_levels, lwe_key, glwe_key, rows, s2, ksk, threads);
}

int ygpu_keygen_lazy_tuniform(const ygpu_params* params, uint64_t seed, int32_t lwe_bound_bits, int32_t glwe_bound_bits,
                              int32_t digit_levels, const int32_t* v, int32_t* lwe_key, int32_t* glwe_key, int32_t* first,
                              int8_t* digits_a, int8_t* digits_b, uint64_t* s2, int32_t* ksk, int32_t threads) {
    if (!tuniform_ok(params, lwe_bound_bits, glwe_bound_bits)) return YGPU_ERR_INVALID_ARG;
    return keygen_lazy(params, seed, Noise::tuniform_bound(lwe_bound_bits), Noise::tuniform_bound(glwe_bound_bits), digit_levels, v,
                       lwe_key, glwe_key, first, digits_a, digits_b, s2, ksk, threads);
}

int ygpu_keygen_lazy_alt_tuniform(const ygpu_params* params, uint64_t seed, int32_t lwe_bound_bits, int32_t glwe_bound_bits,
                                  int32_t digit_levels, int32_t* lwe_key, int32_t* glwe_key, int32_t* rows, uint64_t* s2, int32_t* ksk,
                                  int32_t threads) {
    if (!tuniform_ok(params, lwe_bound_bits, glwe_bound_bits)) return YGPU_ERR_INVALID_ARG;
    return keygen_lazy_alt(params, seed, Noise::tuniform_bound(lwe_bound_bits), Noise::tuniform_bound(glwe_bound_bits), digit_levels,
                           lwe_key, glwe_key, rows, s2, ksk, threads);
}

int ygpu_lwe_encrypt(int32_t n, const int32_t* lwe_key, uint64_t seed, double sigma_log2, int32_t count, const int32_t* mu,
                     int32_t* out) {
    return lwe_encrypt(n, lwe_key, seed, Noise::gaussian_log2(sigma_log2, 32), count, mu, out);
}

int ygpu_lwe_encrypt_tuniform(int32_t n, const int32_t* lwe_key, uint64_t seed, int32_t bound_bits, int32_t count, const int32_t* mu,
                              int32_t* out) {
    if (bound_bits < 0 || bound_bits >= 31) return YGPU_ERR_INVALID_ARG;
    return lwe_encrypt(n, lwe_key, seed, Noise::tuniform_bound(bound_bits), count, mu, out);
}

}  // extern "C"
