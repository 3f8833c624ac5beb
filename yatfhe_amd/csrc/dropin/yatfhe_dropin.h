// yatfhe_dropin.h -- C++ drop-in for the bootstrapping path of the reference
// library (xsong93/yatfhe), implemented on the MI355X engine's C ABI.
//
// Header-only.  Like the reference, the torus width is chosen at compile time
// with -DTORUS32 or -DTORUS56 (torus.h:11-21).  Types and free functions keep
// the reference's names, member names, argument order and meaning; the
// bootstrapping path (blind rotation, sample extract, keyswitch) runs on the GPU.
//
// Differences a caller must know:
//  * BootstrappingKeyMP / BootstrappingKey hold the key in coefficient form
//    (residues mod qNtt, the reference's NTT-form words after an inverse
//    transform without centring) in `bskCoeff`; the reference's HEXL-order
//    `bskDft` words are internal to HEXL and are not exchanged.
//  * Key generation is seeded (setYatfheSeed); the reference's ChaCha20 stream
//    is OS-seeded and cannot be reproduced.
//  * extractTlweFromTrlwe supports index 0 (the index the bootstrap uses) and
//    throws std::invalid_argument otherwise.
//  * Only k = 1, N in {1024, 2048}, l = lApprox <= 4, qLwe = 2^32 are supported
//    (ygpu_init validates and the shim throws std::invalid_argument).
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "yatfhe_gpu.h"

#if defined(TORUS32)
using Torus = int32_t;
using UnsignedInteger = uint32_t;
using Decomp = int8_t;
#elif defined(TORUS56)
using Torus = int64_t;
using UnsignedInteger = uint64_t;
using Decomp = int32_t;
#else
#error "yatfhe_dropin.h: define TORUS32 or TORUS56"
#endif
#if defined(LWE_TORUS32)
using LweTorus = int32_t;
#else
using LweTorus = Torus;
#endif
using Integer = Torus;
using Binary = Integer;

constexpr int64_t Q_32 = INT64_C(1) << 32;
constexpr int64_t Q_56 = INT64_C(1) << 56;
constexpr uint64_t Q_49P_T32 = 562941363486721ULL;      // torus.h:151
constexpr uint64_t Q_61P_T56 = 1945555039024054273ULL;  // torus.h:157

// YatfheParameters (yatfhe_parameters.h:9-110), binary-secret defaults.
struct YatfheParameters {
    int n{680};
    int lweNoiseB{18};
    int64_t qLwe{Q_32};
    int qLweBits{32};
    int torusBase{4};
#if defined(TORUS32)
    int k{1};
    int N{2048};
    int rlweNoiseB{3};
    int64_t q{Q_32};
    int torusBits{32};
    int radixBits{4};
    int l{8};
    int lApprox{4};
    uint64_t qNtt{Q_49P_T32};
#else
    int k{1};
    int N{2048};
    int rlweNoiseB{8};
    int64_t q{Q_56};
    int torusBits{56};
    int radixBits{14};
    int l{4};
    int lApprox{1};
    uint64_t qNtt{Q_61P_T56};
#endif
    int driftPhase{0};
    int radixBase{1 << radixBits};
    int baseOverTwo{radixBase / 2};
    int digitMask{radixBase - 1};
    int ksRadixBits{2};
    int ksWidthBits{qLweBits};
    int ksLevel{ksWidthBits / ksRadixBits};
    int group{1};

    void setRadixBits(const int b) {
        radixBits = b;
        radixBase = 1 << radixBits;
        baseOverTwo = radixBase / 2;
        digitMask = radixBase - 1;
    }
    void setKsRadixBits(const int b) {
        ksRadixBits = b;
        ksWidthBits = qLweBits;
        ksLevel = ksWidthBits / ksRadixBits;
    }
};

struct TorusPolynomial {
    std::vector<Torus> coeffs;
    int N{};
    TorusPolynomial() = default;
    explicit TorusPolynomial(int n) : coeffs(n, 0), N(n) {}
    TorusPolynomial(int n, Torus val) : coeffs(n, val), N(n) {}
};

struct Tlwe {
    std::vector<LweTorus> a{};
    LweTorus b{};
    int n{};
    Tlwe() = default;
    explicit Tlwe(int n_) : a(n_), b(0), n(n_) {}
};

struct ScaledTlwe {
    std::vector<int32_t> a{};
    int32_t b{};
    int n{};
    int mod{};
    ScaledTlwe() = default;
    ScaledTlwe(int mod_, int n_) : a(n_), b(0), n(n_), mod(mod_) {}
};

struct Trlwe {
    std::vector<TorusPolynomial> a;
    TorusPolynomial b;
    int N{};
    int k{};
    Trlwe() = default;
    explicit Trlwe(const YatfheParameters& p) : a(p.k, TorusPolynomial(p.N)), b(p.N), N(p.N), k(p.k) {}
    Trlwe(int k_, int N_) : a(k_, TorusPolynomial(N_)), b(N_), N(N_), k(k_) {}
};

struct TlweKey {
    int n{};
    int errorB{};
    std::vector<Binary> s{};
    TlweKey(int n_, int errorB_) : n(n_), errorB(errorB_), s(n_) {}
    explicit TlweKey(int n_) : n(n_), errorB(0), s(n_) {}
    explicit TlweKey(const YatfheParameters& p) : n(p.n), errorB(p.lweNoiseB), s(p.n) {}
};

struct TrlweKey {
    int k{}, N{}, errorB{};
    std::vector<TorusPolynomial> s;
    explicit TrlweKey(const YatfheParameters& p) : k(p.k), N(p.N), errorB(p.rlweNoiseB), s(p.k, TorusPolynomial(p.N)) {}
};

struct TrgswKey {
    TrlweKey trlweKey;
    explicit TrgswKey(const YatfheParameters& p) : trlweKey(p) {}
};

// Keys in coefficient form; see the header comment.
struct BootstrappingKeyMP {
    std::vector<uint64_t> bskCoeff;  // [n][level][k+1][k+1][N]
    int n{};
    int level{};
    int group{1};
    uint64_t generation{0};          // bumped by key generation, invalidates device copies
    BootstrappingKeyMP() = default;
    explicit BootstrappingKeyMP(const YatfheParameters& p) : BootstrappingKeyMP(p, p.l) {}
    BootstrappingKeyMP(const YatfheParameters& p, int lv)
        : bskCoeff(static_cast<size_t>(p.n) * lv * (p.k + 1) * (p.k + 1) * p.N), n(p.n), level(lv), group(p.group) {}
};
using BootstrappingKey = BootstrappingKeyMP;  // same arithmetic when l == lApprox (classic row order)

namespace yatfhe_dropin {
inline int s2_levels(const YatfheParameters& p);
}

// WWL+24 succinct key (bootstrapping.h:116-136) in coefficient form: the
// body-digit rows cPrime of every step and the TRLEV of s^2 with param.l levels
// (the l the caller passed to initYatfhe).
struct BootstrappingKeyWWL24 {
    std::vector<uint64_t> halfCoeff;  // [n][level][k+1][N]
    std::vector<uint64_t> s2Coeff;    // [s2Level][k+1][N]
    int n{};
    int level{};
    int s2Level{};
    int group{1};
    uint64_t generation{0};
    BootstrappingKeyWWL24() = default;
    BootstrappingKeyWWL24(const YatfheParameters& p, int lv)
        : halfCoeff(static_cast<size_t>(p.n) * lv * (p.k + 1) * p.N), n(p.n), level(lv), s2Level(yatfhe_dropin::s2_levels(p)),
          group(p.group) {
        s2Coeff.assign(static_cast<size_t>(s2Level) * (p.k + 1) * p.N, 0);
    }
};

struct TlweKeySwitchingKey {
    std::vector<std::vector<Tlwe>> decomposedKsk;
    int nCurrKey{};
    int nTargetKey{};
    int level{};
    uint64_t generation{0};
    explicit TlweKeySwitchingKey(const YatfheParameters& p)
        : decomposedKsk(p.N * p.k, std::vector<Tlwe>(p.ksLevel, Tlwe(p.n))), nCurrKey(p.N * p.k), nTargetKey(p.n), level(p.ksLevel) {}
    TlweKeySwitchingKey(int nCurr, int nTarget, int l)
        : decomposedKsk(nCurr, std::vector<Tlwe>(l, Tlwe(nTarget))), nCurrKey(nCurr), nTargetKey(nTarget), level(l) {}
};

namespace yatfhe_dropin {

inline void check(int rc) {
    if (rc != YGPU_OK) {
        const std::string msg = ygpu_last_error();
        if (rc == YGPU_ERR_INVALID_ARG || rc == YGPU_ERR_UNSUPPORTED) throw std::invalid_argument(msg);
        throw std::runtime_error(msg);
    }
}

// Process-global engine state, the counterpart of the reference's globals set
// by initYatfhe (initializer.cpp:17-29): one context for the current parameters.
struct State {
    std::mutex mu;
    ygpu_ctx* ctx = nullptr;
    YatfheParameters param{};
    uint64_t seed = 0x5EED;
    uint64_t draws = 0;
    int requestedL = 0;  // param.l as passed to initYatfhe (the s^2 levels of WWL+24 keys)
    uint64_t nextGeneration = 1;
    // device key cache: (bsk generation, ksk generation) -> key id
    struct Entry { const void* bsk; uint64_t bg; const void* ksk; uint64_t kg; int32_t id; };
    std::vector<Entry> keys;
    ~State() {
        if (ctx) ygpu_destroy(ctx);
    }
};
inline State& state() {
    static State s;
    return s;
}

inline ygpu_params to_params(const YatfheParameters& p, int level) {
    ygpu_params q{};
    q.torus_bits = p.torusBits;
    q.n = p.n;
    q.k = p.k;
    q.N = p.N;
    q.radix_bits = p.radixBits;
    q.level = level;
    q.ks_radix_bits = p.ksRadixBits;
    q.ks_level = p.ksLevel;
    q.ks_width_bits = p.ksWidthBits;
    q.lwe_q_bits = p.qLweBits;
    q.q_ntt = p.qNtt;
    return q;
}

inline uint64_t next_seed() {
    State& s = state();
    return s.seed * 0x9E3779B97F4A7C15ULL + (++s.draws);
}

inline int32_t device_key(const BootstrappingKeyMP& bsk, const TlweKeySwitchingKey& ksk) {
    State& s = state();
    if (!s.ctx) throw std::logic_error("initYatfhe must be called first");
    for (const auto& e : s.keys)
        if (e.bsk == &bsk && e.bg == bsk.generation && e.ksk == &ksk && e.kg == ksk.generation) return e.id;
    const YatfheParameters& p = s.param;
    if (bsk.n != p.n || static_cast<int>(ksk.decomposedKsk.size()) != p.k * p.N || ksk.level != p.ksLevel)
        throw std::invalid_argument("key dimensions do not match the initialised parameters");
    std::vector<int32_t> flat(static_cast<size_t>(p.k) * p.N * p.ksLevel * (p.n + 1));
    size_t o = 0;
    for (const auto& row : ksk.decomposedKsk)
        for (const auto& t : row) {
            for (int i = 0; i < p.n; i++) flat[o++] = static_cast<int32_t>(t.a[i]);
            flat[o++] = static_cast<int32_t>(t.b);
        }
    int32_t id = -1;
    check(ygpu_key_upload(s.ctx, bsk.bskCoeff.data(), flat.data(), &id));
    s.keys.push_back({&bsk, bsk.generation, &ksk, ksk.generation, id});
    return id;
}

inline int s2_levels(const YatfheParameters& p) {
    const State& s = state();
    return s.ctx && s.requestedL > 0 ? s.requestedL : p.l;
}

// Device copy of a WWL+24 key (blind rotation only; an all-zero KSK rides along).
inline int32_t device_key_wwl24(const BootstrappingKeyWWL24& bsk) {
    State& s = state();
    if (!s.ctx) throw std::logic_error("initYatfhe must be called first");
    for (const auto& e : s.keys)
        if (e.bsk == &bsk && e.bg == bsk.generation && e.ksk == nullptr) return e.id;
    const YatfheParameters& p = s.param;
    if (bsk.n != p.n || bsk.level != p.lApprox) throw std::invalid_argument("WWL+24 key dimensions do not match the initialised parameters");
    std::vector<int32_t> zero(static_cast<size_t>(p.k) * p.N * p.ksLevel * (p.n + 1), 0);
    int32_t id = -1;
    check(ygpu_key_upload_succinct(s.ctx, bsk.halfCoeff.data(), bsk.s2Coeff.data(), bsk.s2Level, zero.data(), &id));
    s.keys.push_back({&bsk, bsk.generation, nullptr, 0, id});
    return id;
}

// Any resident key whose BSK matches (blind rotation needs no KSK).
inline int32_t device_key_bsk(const BootstrappingKeyMP& bsk) {
    State& s = state();
    for (const auto& e : s.keys)
        if (e.bsk == &bsk && e.bg == bsk.generation) return e.id;
    // upload with an all-zero KSK (unused by blind rotation)
    static TlweKeySwitchingKey* zero = nullptr;
    if (!zero || zero->nTargetKey != s.param.n || static_cast<int>(zero->decomposedKsk.size()) != s.param.k * s.param.N) {
        delete zero;
        zero = new TlweKeySwitchingKey(s.param);
    }
    return device_key(bsk, *zero);
}

}  // namespace yatfhe_dropin

// ------------------------------------------------------------------ init
// initYatfhe (initializer.cpp:50-85): binds the process to one parameter set.
inline void initYatfhe(YatfheParameters& param) {
    auto& s = yatfhe_dropin::state();
    std::lock_guard<std::mutex> lk(s.mu);
    s.requestedL = param.l;
    if (param.l != param.lApprox)
        param.l = param.lApprox;  // the engine decomposes into exactly the levels it multiplies
    if (s.ctx) {
        ygpu_destroy(s.ctx);
        s.ctx = nullptr;
        s.keys.clear();
    }
    const ygpu_params q = yatfhe_dropin::to_params(param, param.lApprox);
    yatfhe_dropin::check(ygpu_init(&q, 0, &s.ctx));
    s.param = param;
}

inline void setYatfheSeed(uint64_t seed) {
    auto& s = yatfhe_dropin::state();
    std::lock_guard<std::mutex> lk(s.mu);
    s.seed = seed;
    s.draws = 0;
}

// ------------------------------------------------------------------ scalar rules
// longModP (numeric.h:53-66)
inline int64_t longModP(const int64_t a, const int64_t p) {
    if (p > 1 && (p & (p - 1)) == 0) {
        const int shift = 64 - __builtin_ctzll(static_cast<uint64_t>(p));
        return static_cast<int64_t>(static_cast<uint64_t>(a) << shift) >> shift;
    }
    int64_t b = a % p;
    if (b > (p - 1) / 2) b -= p;
    else if (b < -p / 2) b += p;
    return b;
}

inline int intModP(const int a, const int p) {
    int b = a % p;
    if (b > (p - 1) / 2) b -= p;
    else if (b < -p / 2) b += p;
    return b;
}

// modSwitchToTorusGeneral (numeric.cpp:183-187)
inline Torus modSwitchToTorusGeneral(const int32_t mu, const uint32_t mSize, const int64_t torusQ) {
    const int64_t scale = torusQ / mSize;
    return static_cast<Torus>(intModP(mu, static_cast<int32_t>(mSize)) * scale);
}

// modSwitchFromTorusGeneral (numeric.cpp:189-208), power-of-two ratios
inline int64_t modSwitchFromTorusGeneral(const Torus in, const int64_t newMod, const int64_t torusQ) {
    if (torusQ >= newMod) {
        const int64_t scale = torusQ / newMod;
        const int64_t num = static_cast<int64_t>(in);
        const int64_t real = (num >= 0) ? (num + scale / 2) / scale : (num - scale / 2) / scale;
        return longModP(real, newMod);
    }
    const int64_t up = newMod / torusQ;
    if (up * torusQ != newMod) throw std::invalid_argument("modSwitchFromTorusGeneral: ratio is not an integer");
    return longModP(static_cast<int64_t>(in) * up, newMod);
}

// rescaleTlweToNewMod (tlwe.cpp:75-81)
inline void rescaleTlweToNewMod(ScaledTlwe& output, const Tlwe& input) {
    const int64_t lweQ = yatfhe_dropin::state().param.qLwe;
    output.b = static_cast<int32_t>(modSwitchFromTorusGeneral(static_cast<Torus>(input.b), output.mod, lweQ));
    for (int i = 0; i < input.n; i++) output.a[i] = static_cast<int32_t>(modSwitchFromTorusGeneral(static_cast<Torus>(input.a[i]), output.mod, lweQ));
}

// generateTestPolynomialFR (polynomial.cpp:92-107)
inline void generateTestPolynomialFR(TorusPolynomial& v, const int modP, const int modQ) {
    const int N = v.N;
    const int box = modQ / modP, half = box / 2, hh = half / 2;
    const int64_t torusQ = yatfhe_dropin::state().param.q;
    for (int i = 0; i < N; i++) v.coeffs[i] = modSwitchToTorusGeneral(intModP(i / half, modP), modP, torusQ);
    for (int i = 0; i < hh; i++) v.coeffs[i] = static_cast<Torus>(-v.coeffs[i]);
    std::vector<Torus> r(N);
    for (int i = 0; i < N; i++) r[i] = v.coeffs[(i + hh) % N];
    v.coeffs = r;
}

// genNoiselessTrlweSample (trlwe.cpp:175-182): accum = (0, X^{-b} v)
inline void genNoiselessTrlweSample(Trlwe& accum, const TorusPolynomial& v, const ScaledTlwe& scaledInput) {
    const int N = v.N;
    for (auto& poly : accum.a) std::fill(poly.coeffs.begin(), poly.coeffs.end(), 0);
    int a = -scaledInput.b;
    int aTrue = a % (2 * N);
    if (aTrue < 0) aTrue += 2 * N;
    const int isWrap = aTrue < N ? 1 : -1;
    aTrue = aTrue < N ? aTrue : aTrue - N;
    const int w = yatfhe_dropin::state().param.torusBits;
    for (int i = 0; i < N; i++) {
        const int64_t t = (i < aTrue) ? -static_cast<int64_t>(v.coeffs[i - aTrue + N]) * isWrap : static_cast<int64_t>(v.coeffs[i - aTrue]) * isWrap;
        accum.b.coeffs[i] = static_cast<Torus>(static_cast<int64_t>(static_cast<uint64_t>(t) << (64 - w)) >> (64 - w));
    }
}

// ------------------------------------------------------------------ keys
// genTlweKey / genTrlweKey / genBootstrappingKeyMP / genTlweKeySwitchingKey
// (tlwe.cpp, trlwe.cpp:70-81, bootstrapping.cpp:150-174, keyswitching.cpp:19-29):
// one seeded draw generates a consistent key set; the TLWE and TRLWE keys must
// be generated before the keys that depend on them.
namespace yatfhe_dropin {
struct KeyDraw {
    uint64_t seed = 0;
    std::vector<int32_t> lwe, glwe;
    std::vector<uint64_t> bsk;
    std::vector<int32_t> ksk;
};
inline KeyDraw& draw() {
    static KeyDraw d;
    return d;
}
inline void ensure_draw(bool fresh) {
    State& s = state();
    KeyDraw& d = draw();
    if (!fresh && !d.lwe.empty()) return;
    const YatfheParameters& p = s.param;
    d.seed = next_seed();
    const ygpu_params q = to_params(p, p.lApprox);
    d.lwe.assign(p.n, 0);
    d.glwe.assign(static_cast<size_t>(p.k) * p.N, 0);
    d.bsk.assign(static_cast<size_t>(p.n) * p.lApprox * (p.k + 1) * (p.k + 1) * p.N, 0);
    d.ksk.assign(static_cast<size_t>(p.k) * p.N * p.ksLevel * (p.n + 1), 0);
    const double sl = -15.0, sg = p.torusBits == 32 ? -25.0 : -48.0;
    check(ygpu_keygen(&q, d.seed, sl, sg, d.lwe.data(), d.glwe.data(), d.bsk.data(), d.ksk.data(), 0));
}
}  // namespace yatfhe_dropin

inline void genTlweKey(TlweKey& key) {
    yatfhe_dropin::ensure_draw(true);
    const auto& d = yatfhe_dropin::draw();
    if (key.n != static_cast<int>(d.lwe.size())) throw std::invalid_argument("genTlweKey: key dimension differs from param.n");
    for (int i = 0; i < key.n; i++) key.s[i] = d.lwe[i];
}

inline void genTrlweKey(TrlweKey& key) {
    yatfhe_dropin::ensure_draw(false);
    const auto& d = yatfhe_dropin::draw();
    for (int i = 0; i < key.k; i++)
        for (int j = 0; j < key.N; j++) key.s[i].coeffs[j] = d.glwe[static_cast<size_t>(i) * key.N + j];
}

inline void genBootstrappingKeyMP(BootstrappingKeyMP& bsk, TrgswKey&, const TlweKey&, const YatfheParameters&) {
    yatfhe_dropin::ensure_draw(false);
    const auto& d = yatfhe_dropin::draw();
    if (bsk.bskCoeff.size() != d.bsk.size()) throw std::invalid_argument("genBootstrappingKeyMP: level must equal param.lApprox");
    bsk.bskCoeff = d.bsk;
    bsk.generation = yatfhe_dropin::state().nextGeneration++;
}

inline void genTlweKeySwitchingKey(TlweKeySwitchingKey& ksk, const TrlweKey&, const TlweKey&, const YatfheParameters& p) {
    yatfhe_dropin::ensure_draw(false);
    const auto& d = yatfhe_dropin::draw();
    size_t o = 0;
    for (int i = 0; i < p.k * p.N; i++)
        for (int j = 0; j < p.ksLevel; j++) {
            Tlwe& t = ksk.decomposedKsk[i][j];
            for (int u = 0; u < p.n; u++) t.a[u] = static_cast<LweTorus>(d.ksk[o++]);
            t.b = static_cast<LweTorus>(d.ksk[o++]);
        }
    ksk.generation = yatfhe_dropin::state().nextGeneration++;
}

// genBootstrappingKey (bootstrapping.cpp:98-104), group 1: the classic key
// (encryptTrgswNtt per step), from the same draw as the other keys.
inline void genBootstrappingKey(BootstrappingKey& bsk, TrgswKey&, const TlweKey&, const YatfheParameters& p) {
    if (p.group != 1) throw std::invalid_argument("genBootstrappingKey: only group == 1 keys are supported");
    yatfhe_dropin::ensure_draw(false);
    auto& s = yatfhe_dropin::state();
    const auto& d = yatfhe_dropin::draw();
    if (bsk.bskCoeff.size() != d.bsk.size()) throw std::invalid_argument("genBootstrappingKey: level must equal param.lApprox");
    const ygpu_params q = yatfhe_dropin::to_params(p, p.lApprox);
    std::vector<int32_t> lwe(p.n), glwe(static_cast<size_t>(p.k) * p.N), ksk(d.ksk.size());
    const double sg = p.torusBits == 32 ? -25.0 : -48.0;
    yatfhe_dropin::check(ygpu_keygen_classic(&q, d.seed, -15.0, sg, lwe.data(), glwe.data(), bsk.bskCoeff.data(), ksk.data(), 0));
    bsk.generation = s.nextGeneration++;
}

// genBootstrappingKeyWWL24 (bootstrapping.cpp:106-126), from the same draw as
// the other keys (same secret keys), with the GLWE noise of param.rlweNoiseB
// (a TUniform bound 2^B in torus units: log2 sigma = B - 0.79).
inline void genBootstrappingKeyWWL24(BootstrappingKeyWWL24& bsk, TrgswKey&, const TlweKey&, const YatfheParameters& p) {
    yatfhe_dropin::ensure_draw(false);
    auto& s = yatfhe_dropin::state();
    const auto& d = yatfhe_dropin::draw();
    if (bsk.level != p.lApprox) throw std::invalid_argument("genBootstrappingKeyWWL24: level must equal param.lApprox");
    const ygpu_params q = yatfhe_dropin::to_params(p, p.lApprox);
    std::vector<int32_t> lwe(p.n), glwe(static_cast<size_t>(p.k) * p.N), ksk(d.ksk.size());
    const double sg = p.rlweNoiseB - 0.79 - p.torusBits;
    yatfhe_dropin::check(ygpu_keygen_succinct(&q, d.seed, -15.0, sg, bsk.s2Level, lwe.data(), glwe.data(), bsk.halfCoeff.data(),
                                              bsk.s2Coeff.data(), ksk.data(), 0));
    bsk.generation = s.nextGeneration++;
}

// symEncTlwe (tlwe.cpp:20-31), seeded
inline void symEncTlwe(Tlwe& tlweSample, const Torus message, const TlweKey& key) {
    std::vector<int32_t> s(key.s.begin(), key.s.end());
    const int32_t mu = static_cast<int32_t>(static_cast<uint32_t>(static_cast<uint64_t>(message)));
    std::vector<int32_t> out(key.n + 1);
    yatfhe_dropin::check(ygpu_lwe_encrypt(key.n, s.data(), yatfhe_dropin::next_seed(), -15.0, 1, &mu, out.data()));
    for (int i = 0; i < key.n; i++) tlweSample.a[i] = out[i];
    tlweSample.b = out[key.n];
}

// symDecTlweToInt (tlwe.cpp:44-53)
inline Integer symDecTlweToInt(Tlwe& in, const TlweKey& key, const int torusBase) {
    int64_t tmp = 0;
    for (int i = 0; i < key.n; i++) if (key.s[i]) tmp += static_cast<int64_t>(in.a[i]) * key.s[i];
    const int64_t lweQ = yatfhe_dropin::state().param.qLwe;
    const Torus phase = static_cast<Torus>(longModP(static_cast<int64_t>(in.b) - longModP(tmp, lweQ), lweQ));
    const int t = static_cast<int>(modSwitchFromTorusGeneral(phase, torusBase, lweQ));
    return static_cast<Integer>(modSwitchFromTorusGeneral(modSwitchToTorusGeneral(t, torusBase, lweQ), torusBase, lweQ));
}

// ------------------------------------------------------------------ the path
// blindRotateMP21Ntt (blind_rotate.cpp:1538-1565) / blindRotateJP22Ntt (1448-1478)
inline void blindRotateMP21Ntt(Trlwe& accum, const BootstrappingKeyMP& bsk, const ScaledTlwe& input, const YatfheParameters&) {
    auto& s = yatfhe_dropin::state();
    std::lock_guard<std::mutex> lk(s.mu);
    const int32_t id = yatfhe_dropin::device_key_bsk(bsk);
    const int N = accum.N;
    std::vector<Torus> acc(static_cast<size_t>(accum.k + 1) * N);
    for (int c = 0; c < accum.k; c++) std::copy(accum.a[c].coeffs.begin(), accum.a[c].coeffs.end(), acc.begin() + static_cast<size_t>(c) * N);
    std::copy(accum.b.coeffs.begin(), accum.b.coeffs.end(), acc.begin() + static_cast<size_t>(accum.k) * N);
    yatfhe_dropin::check(ygpu_blind_rotate_scaled(s.ctx, id, input.a.data(), acc.data()));
    for (int c = 0; c < accum.k; c++) std::copy(acc.begin() + static_cast<size_t>(c) * N, acc.begin() + static_cast<size_t>(c + 1) * N, accum.a[c].coeffs.begin());
    std::copy(acc.begin() + static_cast<size_t>(accum.k) * N, acc.end(), accum.b.coeffs.begin());
}
inline void blindRotateJP22Ntt(Trlwe& accum, const BootstrappingKeyMP& bsk, const ScaledTlwe& input, const YatfheParameters& p) {
    blindRotateMP21Ntt(accum, bsk, input, p);
}
// blindRotateNtt (blind_rotate.cpp:550-561), classic keys with group 1
inline void blindRotateNtt(Trlwe& accum, const BootstrappingKey& bsk, const ScaledTlwe& input, const YatfheParameters& p) {
    if (p.group != 1) throw std::invalid_argument("blindRotateNtt: only group == 1 keys are supported");
    blindRotateMP21Ntt(accum, bsk, input, p);
}

// blindRotateWWL24Ntt (blind_rotate.cpp:1383-1412): the mask-digit rows the
// reference derives at every step (switchTrlweToSecretEmbeddingNtt) are derived
// once, on the device, when the key is first used.
inline void blindRotateWWL24Ntt(Trlwe& accum, const BootstrappingKeyWWL24& bsk, const ScaledTlwe& input, const YatfheParameters&) {
    auto& s = yatfhe_dropin::state();
    std::lock_guard<std::mutex> lk(s.mu);
    const int32_t id = yatfhe_dropin::device_key_wwl24(bsk);
    const int N = accum.N;
    std::vector<Torus> acc(static_cast<size_t>(accum.k + 1) * N);
    for (int c = 0; c < accum.k; c++) std::copy(accum.a[c].coeffs.begin(), accum.a[c].coeffs.end(), acc.begin() + static_cast<size_t>(c) * N);
    std::copy(accum.b.coeffs.begin(), accum.b.coeffs.end(), acc.begin() + static_cast<size_t>(accum.k) * N);
    yatfhe_dropin::check(ygpu_blind_rotate_scaled(s.ctx, id, input.a.data(), acc.data()));
    for (int c = 0; c < accum.k; c++) std::copy(acc.begin() + static_cast<size_t>(c) * N, acc.begin() + static_cast<size_t>(c + 1) * N, accum.a[c].coeffs.begin());
    std::copy(acc.begin() + static_cast<size_t>(accum.k) * N, acc.end(), accum.b.coeffs.begin());
}

// extractTlweFromTrlwe (trlwe.cpp:241-250), index 0, on the GPU
inline void extractTlweFromTrlwe(Tlwe& out, const Trlwe& in, const int index) {
    if (index != 0) throw std::invalid_argument("extractTlweFromTrlwe: this engine extracts index 0 only");
    auto& s = yatfhe_dropin::state();
    std::lock_guard<std::mutex> lk(s.mu);
    const int N = in.N;
    std::vector<Torus> acc(static_cast<size_t>(in.k + 1) * N);
    for (int c = 0; c < in.k; c++) std::copy(in.a[c].coeffs.begin(), in.a[c].coeffs.end(), acc.begin() + static_cast<size_t>(c) * N);
    std::copy(in.b.coeffs.begin(), in.b.coeffs.end(), acc.begin() + static_cast<size_t>(in.k) * N);
    std::vector<int32_t> o(static_cast<size_t>(in.k) * N + 1);
    yatfhe_dropin::check(ygpu_sample_extract(s.ctx, acc.data(), o.data()));
    for (int i = 0; i < in.k * N; i++) out.a[i] = o[i];
    out.b = o[static_cast<size_t>(in.k) * N];
}

// switchKeyForTlwe (keyswitching.cpp:48-72)
inline void switchKeyForTlwe(Tlwe& output, const TlweKeySwitchingKey& ksk, const Tlwe& input, const YatfheParameters&) {
    auto& s = yatfhe_dropin::state();
    std::lock_guard<std::mutex> lk(s.mu);
    static BootstrappingKeyMP* zeroBsk = nullptr;
    // a KSK travels with a BSK on the device; pair it with a zero BSK when used alone
    if (!zeroBsk || zeroBsk->n != s.param.n) {
        delete zeroBsk;
        zeroBsk = new BootstrappingKeyMP(s.param, s.param.lApprox);
        zeroBsk->generation = s.nextGeneration++;
    }
    const int32_t id = yatfhe_dropin::device_key(*zeroBsk, ksk);
    std::vector<int32_t> in(input.n + 1), out(output.n + 1);
    for (int i = 0; i < input.n; i++) in[i] = static_cast<int32_t>(input.a[i]);
    in[input.n] = static_cast<int32_t>(input.b);
    yatfhe_dropin::check(ygpu_keyswitch(s.ctx, id, in.data(), out.data()));
    for (int i = 0; i < output.n; i++) output.a[i] = out[i];
    output.b = out[output.n];
}

// functionalBootstrappingMPNtt (bootstrapping.cpp:71-80): the whole path in one
// GPU call (rescale, accumulator init, blind rotation, extract, keyswitch).
inline void functionalBootstrappingMPNtt(Tlwe& out, const Tlwe& input, const BootstrappingKeyMP& bsk,
                                         const TlweKeySwitchingKey& ksk, const TorusPolynomial& v, const YatfheParameters&) {
    auto& s = yatfhe_dropin::state();
    std::lock_guard<std::mutex> lk(s.mu);
    const int32_t id = yatfhe_dropin::device_key(bsk, ksk);
    std::vector<int32_t> in(input.n + 1), o(out.n + 1);
    for (int i = 0; i < input.n; i++) in[i] = static_cast<int32_t>(input.a[i]);
    in[input.n] = static_cast<int32_t>(input.b);
    yatfhe_dropin::check(ygpu_pbs(s.ctx, id, in.data(), v.coeffs.data(), o.data(), nullptr, nullptr));
    for (int i = 0; i < out.n; i++) out.a[i] = o[i];
    out.b = o[out.n];
}
// functionalBootstrappingNtt (bootstrapping.cpp:60-69), classic keys, l == lApprox
inline void functionalBootstrappingNtt(Tlwe& out, const Tlwe& input, const BootstrappingKey& bsk,
                                       const TlweKeySwitchingKey& ksk, const TorusPolynomial& v, const YatfheParameters& p) {
    functionalBootstrappingMPNtt(out, input, bsk, ksk, v, p);
}
