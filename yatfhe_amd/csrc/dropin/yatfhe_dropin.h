// yatfhe_dropin.h -- C++ drop-in for the bootstrapping path of the reference
// library (xsong93/yatfhe), implemented on the MI355X engine's C ABI.
//
// Header-only.  Like the reference, the torus width is chosen at compile time
// with -DTORUS32 or -DTORUS56 (torus.h:11-21).  Types and free functions keep
// the reference's names, member names, argument order and meaning, so the
// reference's call sites compile unchanged, e.g.
//   blindRotateMP21Ntt(accum, bsKey.bskDft, inputModN2, param);   (compare_test.cpp:58)
//   blindRotateNtt(in2, bsk.bskDft, sTlwe, param);                 (blindrot_test.cpp:191)
//   extractTlweFromTrlwe(tmp, accum, param.driftPhase);
// The bootstrapping path (blind rotation, sample extract, keyswitch) runs on
// the GPU.
//
// What a caller must know:
//  * The key containers are the reference's (BootstrappingKeyMP::bskDft of
//    TrgswMPDft, BootstrappingKey::bskDft of TrgswDft, their TrlweDft rows of
//    NttPolynomial), but every NttPolynomial holds the COEFFICIENT form of the
//    reference's NTT-form polynomial: its residues mod qNtt, i.e. the HEXL
//    words after an inverse transform without centring.  HEXL's word order is
//    internal to HEXL and never crosses this boundary.  The key generation
//    functions below fill that form; a reference user converts HEXL-form keys
//    with HEXL's inverse transform.
//  * A key reaches the device on its first use and is found again by its
//    address and a fingerprint of sampled words; after changing a key's words
//    in place call yatfhe_dropin::invalidateDeviceKeys().
//  * Key generation is seeded (setYatfheSeed); the reference's ChaCha20 stream
//    is OS-seeded and cannot be reproduced.  Noise follows the reference's
//    TUniform law with param.lweNoiseB / param.rlweNoiseB (numeric.cpp:47-60).
//  * The MP path runs at the key's own level (bskDft[0][0].l, as
//    blindRotateMP21Ntt reads it, blind_rotate.cpp:1539): up to 8 levels for
//    radixBits <= 7, 4 above.  The classic path (blindRotateNtt,
//    functionalBootstrappingNtt) decomposes into param.l digits and multiplies
//    param.lApprox of them (trgsw.cpp:550-576, cmux.cpp:16-21); the engine
//    computes it only for l == lApprox and throws std::invalid_argument
//    otherwise rather than return a different result.
//  * Only k = 1, N in {1024, 2048}, qLwe = 2^32 and group = 1 are supported
//    (std::invalid_argument otherwise).
//  * The path functions may be called from several threads at once: each
//    call runs on a stream of its own (the reference's serving workers call
//    blindRotateJP22Ntt concurrently, blindrotate_serving.cpp:197-228).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
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
using NttType = uint64_t;
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

// ------------------------------------------------------------------ containers
// polynomial.h:84-110, tlwe.h:13-33, trlwe.h:29-117, trgsw.h:14-135, trlev.h:24-36
struct TorusPolynomial {
    std::vector<Torus> coeffs;
    int N{};
    TorusPolynomial() = default;
    explicit TorusPolynomial(int n) : coeffs(n, 0), N(n) {}
    TorusPolynomial(int n, Torus val) : coeffs(n, val), N(n) {}
};
using IntPolynomial = TorusPolynomial;

// Coefficient-form residues mod qNtt here (see the header comment).
struct NttPolynomial {
    std::vector<NttType> coeffs{};
    int N{};
    NttPolynomial() = default;
    explicit NttPolynomial(int n) : coeffs(n, 0), N(n) {}
    NttPolynomial(int n, NttType value) : coeffs(n, value), N(n) {}
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
    Trlwe(int k_, int N_, Torus val) : a(k_, TorusPolynomial(N_, val)), b(N_, val), N(N_), k(k_) {}
};

struct TrlweDft {
    std::vector<NttPolynomial> a;
    NttPolynomial b;
    int k{};
    bool a_initialized{};
    TrlweDft() = default;
    explicit TrlweDft(const YatfheParameters& p) : a(p.k, NttPolynomial(p.N)), b(p.N), k(p.k), a_initialized(true) {}
    TrlweDft(int k_, int N) : a(k_, NttPolynomial(N)), b(N), k(k_), a_initialized(true) {}
    TrlweDft(int k_, int N, bool onlyB) : b(N), k(k_), a_initialized(!onlyB) {
        if (a_initialized) a.resize(k_, NttPolynomial(N));
    }
};

// c[lvl][r]: the RGSW row of mask digit a_r; cPrime[lvl]: the row of the body
// digit (externalProductTrgswMPNttImpl, trgsw.cpp:725-748).
struct TrgswMPDft {
    std::vector<std::vector<TrlweDft>> c;
    std::vector<TrlweDft> cPrime;
    int l{};
    int k{};
    bool isHalf{false};
    TrgswMPDft() = default;
    explicit TrgswMPDft(const YatfheParameters& p) : TrgswMPDft(p, p.l) {}
    TrgswMPDft(const YatfheParameters& p, const int level)
        : c(level, std::vector<TrlweDft>(p.k, TrlweDft(p.k, p.N))), cPrime(level, TrlweDft(p.k, p.N)), l(level), k(p.k) {}
    TrgswMPDft(const YatfheParameters& p, const int level, bool half) : cPrime(level, TrlweDft(p.k, p.N)), l(level), k(p.k), isHalf(half) {
        if (!isHalf) c.resize(level, std::vector<TrlweDft>(p.k, TrlweDft(p.k, p.N)));
    }
};

// Coefficient-domain RGSW types: the reference's BootstrappingKey[MP] also
// carry them (member `bsk`), unused by the NTT path; here they are declared
// for source compatibility and left empty.
struct TrgswMP {
    std::vector<std::vector<Trlwe>> c;
    std::vector<Trlwe> cPrime;
    int l{};
    int k{};
    bool isHalf{false};
    TrgswMP() = default;
    // (trgsw.h:89-118): the lazy-pipe Alt key holds its torus half rows here
    TrgswMP(const YatfheParameters& p, const int level, const bool half = false)
        : cPrime(level, Trlwe(p.k, p.N)), l(level), k(p.k), isHalf(half) {
        if (!isHalf) c.assign(level, std::vector<Trlwe>(p.k, Trlwe(p.k, p.N)));
    }
};
struct Trgsw {
    std::vector<std::vector<Trlwe>> trlweSamples{};
    int l{};
    int k{};
};

// trlweDftSamples[lvl][row]: row r < k multiplies mask digit a_r, row k the
// body digit (externalProductTrgswNtt, trgsw.cpp:550-576).
struct TrgswDft {
    std::vector<std::vector<TrlweDft>> trlweDftSamples;
    int l{};
    int k{};
    TrgswDft() = default;
    explicit TrgswDft(const YatfheParameters& p) : TrgswDft(p, p.l) {}
    TrgswDft(const YatfheParameters& p, const int level)
        : trlweDftSamples(level, std::vector<TrlweDft>(p.k + 1, TrlweDft(p.k, p.N))), l(level), k(p.k) {}
};

struct TrlevDft {
    std::vector<TrlweDft> trlweDfts;
    int l{};
    TrlevDft() = default;
    explicit TrlevDft(const YatfheParameters& p) : TrlevDft(p, p.l) {}
    TrlevDft(const YatfheParameters& p, const int level) : trlweDfts(level, TrlweDft(p.k, p.N)), l(level) {}
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

// BootstrappingKey (bootstrapping.h:16-45): the classic key, group 1.
struct BootstrappingKey {
    std::vector<TrgswDft> bskDft{};
    std::vector<Trgsw> bsk{};
    int n{};
    int group{};
    explicit BootstrappingKey(const YatfheParameters& p) : BootstrappingKey(p, p.l) {}
    BootstrappingKey(const YatfheParameters& p, const int level) : group(p.group) {
        if (group != 1) throw std::invalid_argument("BootstrappingKey: only group == 1 keys are supported");
        n = p.n;
        bskDft.assign(n, TrgswDft(p, level));
    }
};

// BootstrappingKeyMP (bootstrapping.h:47-114), binary secrets: n x 1 TrgswMPDft.
struct BootstrappingKeyMP {
    std::vector<std::vector<TrgswMPDft>> bskDft{};
    std::vector<std::vector<TrgswMP>> bsk{};
    int n{};
    int group{};
    bool isHalf{false};
    BootstrappingKeyMP() = default;
    explicit BootstrappingKeyMP(const YatfheParameters& p) : BootstrappingKeyMP(p, p.l) {}
    BootstrappingKeyMP(const YatfheParameters& p, const int level) : group(p.group) {
        if (group != 1) throw std::invalid_argument("BootstrappingKeyMP: only group == 1 keys are supported");
        n = p.n;
        bskDft.assign(n, std::vector<TrgswMPDft>(1, TrgswMPDft(p, level)));
    }
};

// BootstrappingKeyWWL24 (bootstrapping.h:116-136): per step only the rows of
// the body digit (TrgswMPDft with isHalf), and one TRLEV of s^2 with param.l
// levels.
struct BootstrappingKeyWWL24 {
    std::vector<TrgswMPDft> bskDft{};
    TrlevDft s2Dft{};
    int n{};
    int group{};
    BootstrappingKeyWWL24() = default;
    BootstrappingKeyWWL24(const YatfheParameters& p, const int level) : group(p.group) {
        if (group != 1) throw std::invalid_argument("BootstrappingKeyWWL24: only group == 1 keys are supported");
        n = p.n;
        bskDft.assign(n, TrgswMPDft(p, level, true));
        s2Dft = TrlevDft(p, p.l);
    }
};

// DecompPolynomial (polynomial.h:79-92): one digit level of a polynomial, in
// the Decomp type (int8 at TORUS32).
struct DecompPolynomial {
    std::vector<Decomp> coeffs{};
    int N{};
    DecompPolynomial() = default;
    explicit DecompPolynomial(int n) : coeffs(n, 0), N(n) {}
    DecompPolynomial(int n, Decomp value) : coeffs(n, value), N(n) {}
};

// BootstrappingKeyMPLazyPipe (bootstrapping.h:250-282), group 1, binary keys:
// bskFirst[0] a TRLWE of v s_0; bskDecompA[decompIndex(i, lvl)][l][k] and
// bskDecompB[decompIndex(i, lvl)][l] the param.l digits of cPrime[lvl] of s_{i+1}.
// bskB (the reference's transform of cPrime.b) is not read by
// blindRotateLazyPipeNtt and is left zero here.
struct BootstrappingKeyMPLazyPipe {
    std::vector<Trlwe> bskFirst{};
    std::vector<std::vector<std::vector<DecompPolynomial>>> bskDecompA{};
    std::vector<NttPolynomial> bskB{};
    std::vector<std::vector<DecompPolynomial>> bskDecompB{};
    TrlevDft s2Dft;
    int n{};
    int level{};
    int group{};
    bool initialized{false};
    BootstrappingKeyMPLazyPipe() = default;
    BootstrappingKeyMPLazyPipe(const YatfheParameters& p, const int level_, const bool, const bool) : level(level_), group(p.group) {
        if (group != 1) throw std::invalid_argument("BootstrappingKeyMPLazyPipe: only group == 1 keys are supported");
        n = p.n;
        const size_t m = static_cast<size_t>(n - 1) * level;
        bskFirst.assign(1, Trlwe(p.k, p.N));
        bskDecompA.assign(m, std::vector<std::vector<DecompPolynomial>>(p.l, std::vector<DecompPolynomial>(p.k, DecompPolynomial(p.N))));
        bskB.assign(m, NttPolynomial(p.N));
        bskDecompB.assign(m, std::vector<DecompPolynomial>(p.l, DecompPolynomial(p.N)));
        s2Dft = TrlevDft(p, p.l);
    }
    size_t decompIndex(int i, int lvl) const { return static_cast<size_t>(i) * level + lvl; }
};

// BootstrappingKeyMPLazyPipeAlt (bootstrapping.h:284-306): the half rows of
// every s_i in torus form and the TRLEV of s^2 with param.l levels.
struct BootstrappingKeyMPLazyPipeAlt {
    std::vector<TrgswMP> bskPrime;
    TrlevDft s2Dft;
    int n{};
    int level{};
    int group{};
    BootstrappingKeyMPLazyPipeAlt() = default;
    BootstrappingKeyMPLazyPipeAlt(const YatfheParameters& p, const int level_, const bool isHalf) : level(level_), group(p.group) {
        if (group != 1) throw std::invalid_argument("BootstrappingKeyMPLazyPipeAlt: only group == 1 keys are supported");
        n = p.n;
        bskPrime.assign(p.n, TrgswMP(p, level_, isHalf));
        s2Dft = TrlevDft(p, p.l);
    }
};

// TlweKeySwitchingKey (keyswitching.h:14-31)
struct TlweKeySwitchingKey {
    std::vector<std::vector<Tlwe>> decomposedKsk;
    int nCurrKey{};
    int nTargetKey{};
    int level{};
    explicit TlweKeySwitchingKey(const YatfheParameters& p)
        : decomposedKsk(p.N * p.k, std::vector<Tlwe>(p.ksLevel, Tlwe(p.n))), nCurrKey(p.N * p.k), nTargetKey(p.n), level(p.ksLevel) {}
    TlweKeySwitchingKey(int nCurr, int nTarget, int l)
        : decomposedKsk(nCurr, std::vector<Tlwe>(l, Tlwe(nTarget))), nCurrKey(nCurr), nTargetKey(nTarget), level(l) {}
};

// ------------------------------------------------------------------ engine state
namespace yatfhe_dropin {

inline void check(int rc) {
    if (rc != YGPU_OK) {
        const std::string msg = ygpu_last_error();
        if (rc == YGPU_ERR_INVALID_ARG || rc == YGPU_ERR_UNSUPPORTED) throw std::invalid_argument(msg);
        throw std::runtime_error(msg);
    }
}

// Process-global engine state, the counterpart of the reference's globals set
// by initYatfhe (initializer.cpp:17-29): the parameters, and one engine
// context per (gadget level, extraction index) in use, each holding the device
// copies of the keys used with it.
struct State {
    std::mutex mu;
    bool ready = false;
    YatfheParameters param{};
    uint64_t seed = 0x5EED;
    uint64_t draws = 0;
    std::map<std::pair<int, int>, ygpu_ctx*> ctxs;
    // device key cache: (context, BSK address, BSK fingerprint, KSK address, KSK fingerprint) -> key id
    std::map<std::tuple<ygpu_ctx*, const void*, uint64_t, const void*, uint64_t>, int32_t> keys;
    void clear() {
        for (auto& c : ctxs) ygpu_destroy(c.second);
        ctxs.clear();
        keys.clear();
    }
    ~State() { clear(); }
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

inline void require_init() {
    if (!state().ready) throw std::logic_error("initYatfhe must be called first");
}

// The context of (level, extraction index), created on first use.  s.mu held.
inline ygpu_ctx* context_locked(int level, int extract_index) {
    State& s = state();
    auto it = s.ctxs.find({level, extract_index});
    if (it != s.ctxs.end()) return it->second;
    const ygpu_params q = to_params(s.param, level);
    ygpu_ctx* ctx = nullptr;
    check(ygpu_init(&q, 0, &ctx));
    const int rc = ygpu_set_extract_index(ctx, extract_index);
    if (rc) {
        ygpu_destroy(ctx);
        check(rc);
    }
    s.ctxs[{level, extract_index}] = ctx;
    return ctx;
}

inline uint64_t mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
    h ^= h >> 31;
    return h * 0xBF58476D1CE4E5B9ULL;
}

// Coefficient-form BSK [n][level][k+1][k+1][N] of an MP key (bskDft[i][0]).
inline void flatten_mp(const std::vector<std::vector<TrgswMPDft>>& bskDft, int level, int k, int N, std::vector<uint64_t>& out) {
    const size_t n = bskDft.size();
    out.assign(n * level * (k + 1) * (k + 1) * N, 0);
    size_t o = 0;
    for (size_t i = 0; i < n; i++) {
        const TrgswMPDft& g = bskDft[i].at(0);
        if (g.l != level || g.isHalf || static_cast<int>(g.c.size()) != level)
            throw std::invalid_argument("BootstrappingKeyMP: every step must hold a full TrgswMPDft of the same level");
        for (int lvl = 0; lvl < level; lvl++)
            for (int row = 0; row <= k; row++) {
                const TrlweDft& t = row < k ? g.c[lvl].at(row) : g.cPrime.at(lvl);
                for (int col = 0; col <= k; col++) {
                    const NttPolynomial& poly = col < k ? t.a.at(col) : t.b;
                    if (static_cast<int>(poly.coeffs.size()) != N) throw std::invalid_argument("key polynomial of the wrong degree");
                    std::memcpy(out.data() + o, poly.coeffs.data(), sizeof(uint64_t) * N);
                    o += N;
                }
            }
    }
}

// Coefficient-form BSK of a classic key (bskDft[i].trlweDftSamples[lvl][row]).
inline void flatten_classic(const std::vector<TrgswDft>& bskDft, int level, int k, int N, std::vector<uint64_t>& out) {
    const size_t n = bskDft.size();
    out.assign(n * level * (k + 1) * (k + 1) * N, 0);
    size_t o = 0;
    for (size_t i = 0; i < n; i++) {
        const TrgswDft& g = bskDft[i];
        if (static_cast<int>(g.trlweDftSamples.size()) < level) throw std::invalid_argument("BootstrappingKey: too few levels");
        for (int lvl = 0; lvl < level; lvl++)
            for (int row = 0; row <= k; row++) {
                const TrlweDft& t = g.trlweDftSamples[lvl].at(row);
                for (int col = 0; col <= k; col++) {
                    const NttPolynomial& poly = col < k ? t.a.at(col) : t.b;
                    if (static_cast<int>(poly.coeffs.size()) != N) throw std::invalid_argument("key polynomial of the wrong degree");
                    std::memcpy(out.data() + o, poly.coeffs.data(), sizeof(uint64_t) * N);
                    o += N;
                }
            }
    }
}

inline void unflatten_mp(const std::vector<uint64_t>& in, int level, int k, int N, std::vector<std::vector<TrgswMPDft>>& bskDft) {
    size_t o = 0;
    for (auto& step : bskDft) {
        TrgswMPDft& g = step.at(0);
        for (int lvl = 0; lvl < level; lvl++)
            for (int row = 0; row <= k; row++) {
                TrlweDft& t = row < k ? g.c[lvl][row] : g.cPrime[lvl];
                for (int col = 0; col <= k; col++) {
                    NttPolynomial& poly = col < k ? t.a[col] : t.b;
                    std::memcpy(poly.coeffs.data(), in.data() + o, sizeof(uint64_t) * N);
                    o += N;
                }
            }
    }
}

inline void unflatten_classic(const std::vector<uint64_t>& in, int level, int k, int N, std::vector<TrgswDft>& bskDft) {
    size_t o = 0;
    for (auto& g : bskDft)
        for (int lvl = 0; lvl < level; lvl++)
            for (int row = 0; row <= k; row++) {
                TrlweDft& t = g.trlweDftSamples[lvl][row];
                for (int col = 0; col <= k; col++) {
                    NttPolynomial& poly = col < k ? t.a[col] : t.b;
                    std::memcpy(poly.coeffs.data(), in.data() + o, sizeof(uint64_t) * N);
                    o += N;
                }
            }
}

// A fingerprint of 64 words spread over a flat-indexed key (get(idx) returns
// word idx of `words`), plus its size.
template <class Get>
inline uint64_t fingerprint(size_t words, Get&& get) {
    uint64_t h = mix(0x243F6A8885A308D3ULL, words);
    if (words == 0) return h;
    for (size_t s = 0; s < 64; s++) {
        const size_t idx = (s * 0x9E3779B97F4A7C15ULL) % words;
        h = mix(h, get(idx) ^ idx);
    }
    return h;
}

inline uint64_t fingerprint_mp(const std::vector<std::vector<TrgswMPDft>>& bskDft, int level, int k, int N) {
    const size_t per = static_cast<size_t>(level) * (k + 1) * (k + 1) * N;
    return fingerprint(bskDft.size() * per, [&](size_t idx) -> uint64_t {
        const size_t i = idx / per;
        size_t r = idx % per;
        const int lvl = static_cast<int>(r / ((k + 1) * (k + 1) * N));
        r %= (k + 1) * (k + 1) * N;
        const int row = static_cast<int>(r / ((k + 1) * N));
        r %= (k + 1) * N;
        const int col = static_cast<int>(r / N);
        const int j = static_cast<int>(r % N);
        const TrgswMPDft& g = bskDft[i].at(0);
        const TrlweDft& t = row < k ? g.c.at(lvl).at(row) : g.cPrime.at(lvl);
        const NttPolynomial& poly = col < k ? t.a.at(col) : t.b;
        return poly.coeffs.at(j);
    });
}

inline uint64_t fingerprint_classic(const std::vector<TrgswDft>& bskDft, int level, int k, int N) {
    const size_t per = static_cast<size_t>(level) * (k + 1) * (k + 1) * N;
    return fingerprint(bskDft.size() * per, [&](size_t idx) -> uint64_t {
        const size_t i = idx / per;
        size_t r = idx % per;
        const int lvl = static_cast<int>(r / ((k + 1) * (k + 1) * N));
        r %= (k + 1) * (k + 1) * N;
        const int row = static_cast<int>(r / ((k + 1) * N));
        r %= (k + 1) * N;
        const int col = static_cast<int>(r / N);
        const int j = static_cast<int>(r % N);
        const TrlweDft& t = bskDft[i].trlweDftSamples.at(lvl).at(row);
        const NttPolynomial& poly = col < k ? t.a.at(col) : t.b;
        return poly.coeffs.at(j);
    });
}

inline uint64_t fingerprint_ksk(const TlweKeySwitchingKey& ksk) {
    const size_t t = static_cast<size_t>(ksk.level), cols = static_cast<size_t>(ksk.nTargetKey) + 1;
    return fingerprint(ksk.decomposedKsk.size() * t * cols, [&](size_t idx) -> uint64_t {
        const size_t i = idx / (t * cols), j = (idx / cols) % t, c = idx % cols;
        const Tlwe& row = ksk.decomposedKsk.at(i).at(j);
        return static_cast<uint64_t>(static_cast<int64_t>(c < cols - 1 ? row.a.at(c) : row.b));
    });
}

inline void flatten_ksk(const TlweKeySwitchingKey& ksk, const YatfheParameters& p, std::vector<int32_t>& flat) {
    if (static_cast<int>(ksk.decomposedKsk.size()) != p.k * p.N || ksk.level != p.ksLevel || ksk.nTargetKey != p.n)
        throw std::invalid_argument("keyswitching key dimensions do not match the initialised parameters");
    flat.assign(static_cast<size_t>(p.k) * p.N * p.ksLevel * (p.n + 1), 0);
    size_t o = 0;
    for (const auto& row : ksk.decomposedKsk)
        for (const auto& t : row) {
            for (int i = 0; i < p.n; i++) flat[o++] = static_cast<int32_t>(t.a.at(i));
            flat[o++] = static_cast<int32_t>(t.b);
        }
}

// The device key (context + key id) for a coefficient-form BSK supplied by
// `flatten` (called only on a cache miss) and an optional KSK (null: an
// all-zero KSK rides along; blind rotation does not read it).  s.mu held.
template <class Flatten>
inline std::pair<ygpu_ctx*, int32_t> device_key_locked(int level, int extract_index, const void* bskAddr, uint64_t bskFp,
                                                        Flatten&& flatten, const TlweKeySwitchingKey* ksk) {
    State& s = state();
    ygpu_ctx* ctx = context_locked(level, extract_index);
    const uint64_t kfp = ksk ? fingerprint_ksk(*ksk) : 0;
    const auto key = std::make_tuple(ctx, bskAddr, bskFp, static_cast<const void*>(ksk), kfp);
    auto it = s.keys.find(key);
    if (it != s.keys.end()) return {ctx, it->second};
    std::vector<uint64_t> bsk;
    flatten(bsk);
    std::vector<int32_t> flat;
    if (ksk) flatten_ksk(*ksk, s.param, flat);
    else flat.assign(static_cast<size_t>(s.param.k) * s.param.N * s.param.ksLevel * (s.param.n + 1), 0);
    int32_t id = -1;
    check(ygpu_key_upload(ctx, bsk.data(), flat.data(), &id));
    s.keys[key] = id;
    return {ctx, id};
}

inline std::pair<ygpu_ctx*, int32_t> mp_key(const std::vector<std::vector<TrgswMPDft>>& bskDft, const TlweKeySwitchingKey* ksk,
                                            const YatfheParameters& p) {
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    require_init();
    if (bskDft.empty() || bskDft[0].empty()) throw std::invalid_argument("empty bootstrapping key");
    if (static_cast<int>(bskDft.size()) != s.param.n) throw std::invalid_argument("bootstrapping key has the wrong number of steps");
    const int level = bskDft[0][0].l;
    const int k = s.param.k, N = s.param.N;
    const uint64_t fp = fingerprint_mp(bskDft, level, k, N);
    return device_key_locked(level, p.driftPhase, &bskDft, fp, [&](std::vector<uint64_t>& out) { flatten_mp(bskDft, level, k, N, out); }, ksk);
}

// The classic path computes what the MP path computes only when the key is
// decomposed into exactly the levels it multiplies.
inline std::pair<ygpu_ctx*, int32_t> classic_key(const std::vector<TrgswDft>& bskDft, const TlweKeySwitchingKey* ksk,
                                                 const YatfheParameters& p) {
    if (p.group != 1) throw std::invalid_argument("blindRotateNtt: only group == 1 keys are supported");
    if (p.l != p.lApprox)
        throw std::invalid_argument("blindRotateNtt: the classic path decomposes into param.l digits and multiplies param.lApprox of "
                                    "them; this engine computes it only for l == lApprox (l = " + std::to_string(p.l) +
                                    ", lApprox = " + std::to_string(p.lApprox) + ")");
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    require_init();
    if (static_cast<int>(bskDft.size()) != s.param.n) throw std::invalid_argument("bootstrapping key has the wrong number of steps");
    const int level = p.lApprox;
    const int k = s.param.k, N = s.param.N;
    const uint64_t fp = fingerprint_classic(bskDft, level, k, N);
    return device_key_locked(level, p.driftPhase, &bskDft, fp, [&](std::vector<uint64_t>& out) { flatten_classic(bskDft, level, k, N, out); },
                             ksk);
}

inline uint64_t next_seed() {
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    return s.seed * 0x9E3779B97F4A7C15ULL + (++s.draws);
}

// Accumulator <-> flat Torus[(k+1)N].
inline void to_flat(const Trlwe& t, std::vector<Torus>& acc) {
    const int N = t.N;
    acc.resize(static_cast<size_t>(t.k + 1) * N);
    for (int c = 0; c < t.k; c++) std::copy(t.a[c].coeffs.begin(), t.a[c].coeffs.end(), acc.begin() + static_cast<size_t>(c) * N);
    std::copy(t.b.coeffs.begin(), t.b.coeffs.end(), acc.begin() + static_cast<size_t>(t.k) * N);
}
inline void from_flat(const std::vector<Torus>& acc, Trlwe& t) {
    const int N = t.N;
    for (int c = 0; c < t.k; c++) std::copy(acc.begin() + static_cast<size_t>(c) * N, acc.begin() + static_cast<size_t>(c + 1) * N, t.a[c].coeffs.begin());
    std::copy(acc.begin() + static_cast<size_t>(t.k) * N, acc.end(), t.b.coeffs.begin());
}

// Forgets every device copy of a key (call after changing a key's words in place).
inline void invalidateDeviceKeys() {
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    for (auto& e : s.keys) ygpu_key_release(std::get<0>(e.first), e.second);
    s.keys.clear();
}

}  // namespace yatfhe_dropin

// ------------------------------------------------------------------ init
// initYatfhe (initializer.cpp:50-85): binds the process to one parameter set.
// The parameters are validated (std::invalid_argument) and left unchanged.
inline void initYatfhe(YatfheParameters& param) {
    auto& s = yatfhe_dropin::state();
    std::lock_guard<std::mutex> lk(s.mu);
    s.clear();
    s.ready = false;
    s.param = param;
    // validate now: the contexts of the levels keys will use are made on first use
    yatfhe_dropin::context_locked(param.lApprox, param.driftPhase);
    s.ready = true;
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
// modSwitchToTorus32 (numeric.cpp:210-214): the same on the torus modulus
inline Torus modSwitchToTorus32(const int32_t mu, const uint32_t mSize) {
    return modSwitchToTorusGeneral(mu, mSize, yatfhe_dropin::state().param.q);
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

namespace yatfhe_dropin {
// The box layout shared by the test-polynomial generators (polynomial.cpp:75-127):
// coefficient i takes the value of message slot i / box, the first half box is
// negated and the polynomial is rotated left by half a box.
inline void box_polynomial(TorusPolynomial& v, const int box, const std::vector<Torus>& slotValue) {
    const int N = v.N, half = box / 2;
    for (int i = 0; i < N; i++) v.coeffs[i] = slotValue[i / box];
    for (int i = 0; i < half; i++) v.coeffs[i] = static_cast<Torus>(-v.coeffs[i]);
    std::rotate(v.coeffs.begin(), v.coeffs.begin() + half, v.coeffs.end());
}
}  // namespace yatfhe_dropin

// generateTestPolynomial (polynomial.cpp:75-89): slot value intModP(slot, modP)
inline void generateTestPolynomial(TorusPolynomial& v, const int modP, const int modQ) {
    const int box = modQ / modP;
    std::vector<Torus> vals(v.N / box + 1);
    for (size_t s = 0; s < vals.size(); s++) vals[s] = modSwitchToTorus32(intModP(static_cast<int>(s), modP), modP);
    yatfhe_dropin::box_polynomial(v, box, vals);
}

// generateTestPolynomialFR (polynomial.cpp:92-107): the full plaintext range,
// boxes of modQ / modP / 2 coefficients
inline void generateTestPolynomialFR(TorusPolynomial& v, const int modP, const int modQ) {
    const int box = modQ / modP, half = box / 2, hh = half / 2;
    const int N = v.N;
    const int64_t torusQ = yatfhe_dropin::state().param.q;
    for (int i = 0; i < N; i++) v.coeffs[i] = modSwitchToTorusGeneral(intModP(i / half, modP), modP, torusQ);
    for (int i = 0; i < hh; i++) v.coeffs[i] = static_cast<Torus>(-v.coeffs[i]);
    std::rotate(v.coeffs.begin(), v.coeffs.begin() + hh, v.coeffs.end());
}

// generateTestPolynomialLt1 (polynomial.cpp:110-127): slot value 1 when its
// message is below t (intModP over param.torusBase), else 0, scaled to the torus
inline void generateTestPolynomialLt1(TorusPolynomial& v, const int t) {
    const int modP = yatfhe_dropin::state().param.torusBase;
    const int box = 2 * v.N / modP;
    const int threshold = intModP(t, modP);
    std::vector<Torus> vals(v.N / box + 1);
    for (size_t s = 0; s < vals.size(); s++) {
        const int msg = intModP(static_cast<int>(s), modP);
        vals[s] = modSwitchToTorus32(msg < threshold ? 1 : 0, modP);
    }
    yatfhe_dropin::box_polynomial(v, box, vals);
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
// genTlweKey / genTrlweKey / genBootstrappingKey[MP|WWL24] / genTlweKeySwitchingKey
// (tlwe.cpp, trlwe.cpp:70-81, bootstrapping.cpp:98-174, keyswitching.cpp:19-29):
// one seeded draw makes a consistent key set; genTlweKey starts a new draw, the
// other generators reuse its secret keys.  Noise: TUniform with param.lweNoiseB
// (LWE and KSK rows) and param.rlweNoiseB (bootstrapping-key rows).
namespace yatfhe_dropin {
struct KeyDraw {
    uint64_t seed = 0;
    bool valid = false;
    std::vector<int32_t> lwe, glwe;
};
inline KeyDraw& draw() {
    static KeyDraw d;
    return d;
}
inline void ensure_draw(bool fresh) {
    require_init();
    KeyDraw& d = draw();
    if (!fresh && d.valid) return;
    d.seed = next_seed();
    d.valid = true;
    const YatfheParameters& p = state().param;
    d.lwe.assign(p.n, 0);
    d.glwe.assign(static_cast<size_t>(p.k) * p.N, 0);
    // the secret keys depend only on the seed: a one-level draw gives them
    const ygpu_params q = to_params(p, 1);
    std::vector<uint64_t> bsk(static_cast<size_t>(p.n) * (p.k + 1) * (p.k + 1) * p.N);
    std::vector<int32_t> ksk(static_cast<size_t>(p.k) * p.N * p.ksLevel * (p.n + 1));
    check(ygpu_keygen_tuniform(&q, d.seed, p.lweNoiseB, p.rlweNoiseB, YGPU_KEY_MP, 0, d.lwe.data(), d.glwe.data(), bsk.data(), nullptr,
                               ksk.data(), 0));
}
// Keys of one form at `level` from the current draw.
inline void draw_keys(int form, int level, int s2Level, std::vector<uint64_t>& bsk, std::vector<uint64_t>* s2, std::vector<int32_t>& ksk) {
    ensure_draw(false);
    const YatfheParameters& p = state().param;
    const ygpu_params q = to_params(p, level);
    std::vector<int32_t> lwe(p.n), glwe(static_cast<size_t>(p.k) * p.N);
    const size_t rows = form == YGPU_KEY_SUCCINCT ? static_cast<size_t>(p.n) * level * (p.k + 1)
                                                  : static_cast<size_t>(p.n) * level * (p.k + 1) * (p.k + 1);
    bsk.assign(rows * p.N, 0);
    if (s2) s2->assign(static_cast<size_t>(s2Level) * (p.k + 1) * p.N, 0);
    ksk.assign(static_cast<size_t>(p.k) * p.N * p.ksLevel * (p.n + 1), 0);
    check(ygpu_keygen_tuniform(&q, draw().seed, p.lweNoiseB, p.rlweNoiseB, form, s2Level, lwe.data(), glwe.data(), bsk.data(),
                               s2 ? s2->data() : nullptr, ksk.data(), 0));
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

// genBootstrappingKeyMP (bootstrapping.cpp:150-174), at the key's level.
inline void genBootstrappingKeyMP(BootstrappingKeyMP& bsk, TrgswKey&, const TlweKey&, const YatfheParameters& p) {
    if (bsk.bskDft.empty() || bsk.bskDft[0].empty()) throw std::invalid_argument("genBootstrappingKeyMP: empty key");
    const int level = bsk.bskDft[0][0].l;
    std::vector<uint64_t> flat;
    std::vector<int32_t> ksk;
    yatfhe_dropin::draw_keys(YGPU_KEY_MP, level, 0, flat, nullptr, ksk);
    yatfhe_dropin::unflatten_mp(flat, level, p.k, p.N, bsk.bskDft);
}

// genBootstrappingKey (bootstrapping.cpp:98-104), group 1: encryptTrgswNtt per step.
inline void genBootstrappingKey(BootstrappingKey& bsk, TrgswKey&, const TlweKey&, const YatfheParameters& p) {
    if (p.group != 1) throw std::invalid_argument("genBootstrappingKey: only group == 1 keys are supported");
    if (bsk.bskDft.empty()) throw std::invalid_argument("genBootstrappingKey: empty key");
    const int level = bsk.bskDft[0].l;
    std::vector<uint64_t> flat;
    std::vector<int32_t> ksk;
    yatfhe_dropin::draw_keys(YGPU_KEY_CLASSIC, level, 0, flat, nullptr, ksk);
    yatfhe_dropin::unflatten_classic(flat, level, p.k, p.N, bsk.bskDft);
}

// genTlweKeySwitchingKey (keyswitching.cpp:19-29)
inline void genTlweKeySwitchingKey(TlweKeySwitchingKey& ksk, const TrlweKey&, const TlweKey&, const YatfheParameters& p) {
    std::vector<uint64_t> flat;
    std::vector<int32_t> k;
    yatfhe_dropin::draw_keys(YGPU_KEY_MP, 1, 0, flat, nullptr, k);
    size_t o = 0;
    for (int i = 0; i < p.k * p.N; i++)
        for (int j = 0; j < p.ksLevel; j++) {
            Tlwe& t = ksk.decomposedKsk[i][j];
            for (int u = 0; u < p.n; u++) t.a[u] = static_cast<LweTorus>(k[o++]);
            t.b = static_cast<LweTorus>(k[o++]);
        }
}

// genBootstrappingKeyWWL24 (bootstrapping.cpp:106-126): the body-digit rows of
// every step and the TRLEV of s^2 with bsk.s2Dft.l levels.
inline void genBootstrappingKeyWWL24(BootstrappingKeyWWL24& bsk, TrgswKey&, const TlweKey&, const YatfheParameters& p) {
    if (bsk.bskDft.empty()) throw std::invalid_argument("genBootstrappingKeyWWL24: empty key");
    const int level = bsk.bskDft[0].l, s2Level = bsk.s2Dft.l;
    std::vector<uint64_t> half, s2;
    std::vector<int32_t> ksk;
    yatfhe_dropin::draw_keys(YGPU_KEY_SUCCINCT, level, s2Level, half, &s2, ksk);
    size_t o = 0;
    for (auto& g : bsk.bskDft)
        for (int lvl = 0; lvl < level; lvl++)
            for (int col = 0; col <= p.k; col++) {
                NttPolynomial& poly = col < p.k ? g.cPrime[lvl].a[col] : g.cPrime[lvl].b;
                std::memcpy(poly.coeffs.data(), half.data() + o, sizeof(uint64_t) * p.N);
                o += p.N;
            }
    o = 0;
    for (int l2 = 0; l2 < s2Level; l2++)
        for (int col = 0; col <= p.k; col++) {
            NttPolynomial& poly = col < p.k ? bsk.s2Dft.trlweDfts[l2].a[col] : bsk.s2Dft.trlweDfts[l2].b;
            std::memcpy(poly.coeffs.data(), s2.data() + o, sizeof(uint64_t) * p.N);
            o += p.N;
        }
}

// genBootstrappingKeyMPLazyPipe (bootstrapping.cpp:269-320) for test polynomial
// v, TORUS32: rows of bsk.level levels, param.l digits each.
inline void genBootstrappingKeyMPLazyPipe(BootstrappingKeyMPLazyPipe& bsk, const TrgswKey&, const TlweKey&, const TorusPolynomial& v,
                                          const YatfheParameters& p) {
#if defined(TORUS32)
    using namespace yatfhe_dropin;
    if (bsk.bskDecompB.empty()) throw std::invalid_argument("genBootstrappingKeyMPLazyPipe: empty key");
    ensure_draw(false);
    const ygpu_params q = to_params(p, bsk.level);
    const size_t N = p.N, K = p.k, R = bsk.level, D = p.l, m = static_cast<size_t>(p.n - 1);
    std::vector<int32_t> lwe(p.n), glwe(K * N), first((K + 1) * N), ksk(K * N * p.ksLevel * (p.n + 1));
    std::vector<int8_t> da(m * R * D * K * N), db(m * R * D * N);
    std::vector<uint64_t> s2(D * (K + 1) * N);
    std::vector<int32_t> vv(v.coeffs.begin(), v.coeffs.end());
    check(ygpu_keygen_lazy_tuniform(&q, draw().seed, p.lweNoiseB, p.rlweNoiseB, static_cast<int32_t>(D), vv.data(), lwe.data(), glwe.data(),
                                    first.data(), da.data(), db.data(), s2.data(), ksk.data(), 0));
    for (size_t c = 0; c <= K; c++) {
        TorusPolynomial& t = c < K ? bsk.bskFirst[0].a[c] : bsk.bskFirst[0].b;
        for (size_t j = 0; j < N; j++) t.coeffs[j] = first[c * N + j];
    }
    for (size_t cell = 0; cell < m * R; cell++)
        for (size_t l = 0; l < D; l++) {
            for (size_t c = 0; c < K; c++)
                std::memcpy(bsk.bskDecompA[cell][l][c].coeffs.data(), da.data() + ((cell * D + l) * K + c) * N, N);
            std::memcpy(bsk.bskDecompB[cell][l].coeffs.data(), db.data() + (cell * D + l) * N, N);
        }
    for (size_t l = 0; l < D; l++)
        for (size_t c = 0; c <= K; c++) {
            NttPolynomial& poly = c < K ? bsk.s2Dft.trlweDfts[l].a[c] : bsk.s2Dft.trlweDfts[l].b;
            std::memcpy(poly.coeffs.data(), s2.data() + (l * (K + 1) + c) * N, sizeof(uint64_t) * N);
        }
#else
    (void)bsk; (void)v; (void)p;
    throw std::invalid_argument("genBootstrappingKeyMPLazyPipe: the lazy-pipe forms are built for TORUS32");
#endif
}

// genBootstrappingKeyMPLazyPipeAlt (bootstrapping.cpp:322-339), TORUS32.
inline void genBootstrappingKeyMPLazyPipeAlt(BootstrappingKeyMPLazyPipeAlt& bsk, const TrgswKey&, const TlweKey&, const YatfheParameters& p) {
#if defined(TORUS32)
    using namespace yatfhe_dropin;
    if (bsk.bskPrime.empty()) throw std::invalid_argument("genBootstrappingKeyMPLazyPipeAlt: empty key");
    ensure_draw(false);
    const ygpu_params q = to_params(p, bsk.level);
    const size_t N = p.N, K = p.k, R = bsk.level, D = p.l;
    std::vector<int32_t> lwe(p.n), glwe(K * N), rows(static_cast<size_t>(p.n) * R * (K + 1) * N), ksk(K * N * p.ksLevel * (p.n + 1));
    std::vector<uint64_t> s2(D * (K + 1) * N);
    check(ygpu_keygen_lazy_alt_tuniform(&q, draw().seed, p.lweNoiseB, p.rlweNoiseB, static_cast<int32_t>(D), lwe.data(), glwe.data(),
                                        rows.data(), s2.data(), ksk.data(), 0));
    size_t o = 0;
    for (auto& g : bsk.bskPrime)
        for (size_t lvl = 0; lvl < R; lvl++)
            for (size_t c = 0; c <= K; c++) {
                TorusPolynomial& t = c < K ? g.cPrime[lvl].a[c] : g.cPrime[lvl].b;
                for (size_t j = 0; j < N; j++) t.coeffs[j] = rows[o++];
            }
    for (size_t l = 0; l < D; l++)
        for (size_t c = 0; c <= K; c++) {
            NttPolynomial& poly = c < K ? bsk.s2Dft.trlweDfts[l].a[c] : bsk.s2Dft.trlweDfts[l].b;
            std::memcpy(poly.coeffs.data(), s2.data() + (l * (K + 1) + c) * N, sizeof(uint64_t) * N);
        }
#else
    (void)bsk; (void)p;
    throw std::invalid_argument("genBootstrappingKeyMPLazyPipeAlt: the lazy-pipe forms are built for TORUS32");
#endif
}

// symEncTlwe (tlwe.cpp:20-31): TUniform noise with the key's errorB, seeded
inline void symEncTlwe(Tlwe& tlweSample, const Torus message, const TlweKey& key) {
    std::vector<int32_t> s(key.s.begin(), key.s.end());
    const int32_t mu = static_cast<int32_t>(static_cast<uint32_t>(static_cast<uint64_t>(message)));
    std::vector<int32_t> out(key.n + 1);
    yatfhe_dropin::check(ygpu_lwe_encrypt_tuniform(key.n, s.data(), yatfhe_dropin::next_seed(), key.errorB, 1, &mu, out.data()));
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
namespace yatfhe_dropin {
inline void blind_rotate_on(std::pair<ygpu_ctx*, int32_t> key, Trlwe& accum, const ScaledTlwe& input) {
    std::vector<Torus> acc;
    to_flat(accum, acc);
    check(ygpu_blind_rotate_scaled(key.first, key.second, input.a.data(), acc.data()));
    from_flat(acc, accum);
}
inline void bootstrap_on(std::pair<ygpu_ctx*, int32_t> key, Tlwe& out, const Tlwe& input, const TorusPolynomial& v) {
    std::vector<int32_t> in(input.n + 1), o(out.n + 1);
    for (int i = 0; i < input.n; i++) in[i] = static_cast<int32_t>(input.a[i]);
    in[input.n] = static_cast<int32_t>(input.b);
    check(ygpu_pbs(key.first, key.second, in.data(), v.coeffs.data(), o.data(), nullptr, nullptr));
    for (int i = 0; i < out.n; i++) out.a[i] = o[i];
    out.b = o[out.n];
}
}  // namespace yatfhe_dropin

// blindRotateMP21Ntt (blind_rotate.cpp:1538-1565), at the key's level bskDft[0][0].l
inline void blindRotateMP21Ntt(Trlwe& accum, const std::vector<std::vector<TrgswMPDft>>& bskDft, const ScaledTlwe& input,
                               const YatfheParameters& param) {
    yatfhe_dropin::blind_rotate_on(yatfhe_dropin::mp_key(bskDft, nullptr, param), accum, input);
}
// blindRotateJP22Ntt (blind_rotate.cpp:1448-1478): the same arithmetic
inline void blindRotateJP22Ntt(Trlwe& accum, const BootstrappingKeyMP& bsk, const ScaledTlwe& input, const YatfheParameters& param) {
    blindRotateMP21Ntt(accum, bsk.bskDft, input, param);
}
// blindRotateNtt (blind_rotate.cpp:550-561 -> blindRotateNormalNtt, 202-212), group 1
inline void blindRotateNtt(Trlwe& accum, const std::vector<TrgswDft>& bskDft, const ScaledTlwe& input, const YatfheParameters& param) {
    yatfhe_dropin::blind_rotate_on(yatfhe_dropin::classic_key(bskDft, nullptr, param), accum, input);
}

// blindRotateWWL24Ntt (blind_rotate.cpp:1383-1412): the mask-digit rows the
// reference derives at every step (switchTrlweToSecretEmbeddingNtt) are derived
// once, on the device, when the key is first used.
inline void blindRotateWWL24Ntt(Trlwe& accum, const BootstrappingKeyWWL24& bsk, const ScaledTlwe& input, const YatfheParameters& param) {
    using namespace yatfhe_dropin;
    State& s = state();
    std::pair<ygpu_ctx*, int32_t> key;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        require_init();
        if (bsk.bskDft.empty() || static_cast<int>(bsk.bskDft.size()) != s.param.n)
            throw std::invalid_argument("WWL+24 key has the wrong number of steps");
        const int level = bsk.bskDft[0].l, s2Level = bsk.s2Dft.l;
        const YatfheParameters& p = s.param;
        const size_t per = static_cast<size_t>(level) * (p.k + 1) * p.N;
        const uint64_t fp = fingerprint(bsk.bskDft.size() * per + 1, [&](size_t idx) -> uint64_t {
            if (idx == bsk.bskDft.size() * per) return static_cast<uint64_t>(s2Level);
            const size_t i = idx / per;
            size_t r = idx % per;
            const int lvl = static_cast<int>(r / ((p.k + 1) * p.N));
            r %= (p.k + 1) * p.N;
            const int col = static_cast<int>(r / p.N), j = static_cast<int>(r % p.N);
            const TrlweDft& t = bsk.bskDft[i].cPrime.at(lvl);
            return (col < p.k ? t.a.at(col) : t.b).coeffs.at(j);
        });
        ygpu_ctx* ctx = context_locked(level, param.driftPhase);
        const auto ck = std::make_tuple(ctx, static_cast<const void*>(&bsk), fp, static_cast<const void*>(nullptr), uint64_t{1});
        auto it = s.keys.find(ck);
        if (it != s.keys.end()) {
            key = {ctx, it->second};
        } else {
            std::vector<uint64_t> half(static_cast<size_t>(p.n) * per), s2(static_cast<size_t>(s2Level) * (p.k + 1) * p.N);
            size_t o = 0;
            for (const auto& g : bsk.bskDft)
                for (int lvl = 0; lvl < level; lvl++)
                    for (int col = 0; col <= p.k; col++) {
                        const NttPolynomial& poly = col < p.k ? g.cPrime.at(lvl).a.at(col) : g.cPrime.at(lvl).b;
                        std::memcpy(half.data() + o, poly.coeffs.data(), sizeof(uint64_t) * p.N);
                        o += p.N;
                    }
            o = 0;
            for (int l2 = 0; l2 < s2Level; l2++)
                for (int col = 0; col <= p.k; col++) {
                    const TrlweDft& t = bsk.s2Dft.trlweDfts.at(l2);
                    const NttPolynomial& poly = col < p.k ? t.a.at(col) : t.b;
                    std::memcpy(s2.data() + o, poly.coeffs.data(), sizeof(uint64_t) * p.N);
                    o += p.N;
                }
            std::vector<int32_t> zero(static_cast<size_t>(p.k) * p.N * p.ksLevel * (p.n + 1), 0);
            int32_t id = -1;
            check(ygpu_key_upload_succinct(ctx, half.data(), s2.data(), s2Level, zero.data(), &id));
            s.keys[ck] = id;
            key = {ctx, id};
        }
    }
    blind_rotate_on(key, accum, input);
}

// prepareGdV (trgsw.cpp:911-927): the lApprox gadget digits of v, as
// coefficient-form residues (the NTT-form containers hold coefficient form
// here).  blindRotateLazyPipeAltNtt takes it for source compatibility; the
// engine derives the first component from v itself.
inline void prepareGdV(std::vector<NttPolynomial>& gdVntt, const TorusPolynomial& v, const YatfheParameters& param) {
    const int L = param.lApprox, N = param.N, b = param.radixBits, w = param.torusBits;
    gdVntt.assign(L, NttPolynomial(N));
    const int64_t B = int64_t{1} << b, halfB = B >> 1;
    const uint64_t mask = w >= 64 ? ~0ULL : ((uint64_t{1} << w) - 1);
    const int dropped = w - L * b;
    for (int j = 0; j < N; j++) {
        // zeroMeanDigits (gadget_decomposition.h:75-95)
        const uint64_t u = static_cast<uint64_t>(static_cast<int64_t>(v.coeffs[j])) & mask;
        const bool neg = (u >> (w - 1)) & 1;
        const uint64_t mm = (neg ? (0ULL - u) & mask : u) + (dropped > 0 ? (uint64_t{1} << (dropped - 1)) : 0);
        int64_t carry = 0;
        for (int lvl = L - 1; lvl >= 0; --lvl) {
            const int shift = w - (lvl + 1) * b;
            const int64_t digit = static_cast<int64_t>((mm >> shift) & static_cast<uint64_t>(B - 1)) + carry;
            const int64_t tie = lvl > 0 ? static_cast<int64_t>((mm >> (shift + 2 * b - 1)) & 1) : 0;
            carry = (digit > halfB) | ((digit == halfB) & tie);
            const int64_t d = (neg ? -1 : 1) * (carry ? digit - B : digit);
            gdVntt[lvl].coeffs[j] = d < 0 ? static_cast<NttType>(static_cast<int64_t>(param.qNtt) + d) : static_cast<NttType>(d);
        }
    }
}

namespace yatfhe_dropin {
inline void lazy_rotate_on(std::pair<ygpu_ctx*, int32_t> key, Trlwe& accum, const ScaledTlwe& input, const TorusPolynomial& v) {
    std::vector<int32_t> abar(input.a.begin(), input.a.end());
    abar.push_back(input.b);
    std::vector<Torus> acc(static_cast<size_t>(accum.k + 1) * accum.N);
    check(ygpu_blind_rotate_lazy(key.first, key.second, abar.data(), v.coeffs.data(), acc.data()));
    from_flat(acc, accum);
}
}  // namespace yatfhe_dropin

// blindRotateLazyPipeNtt (blind_rotate.cpp:822-902): accum := the lazy-pipe
// blind rotation of input under bsk (whose bskFirst encrypts v s_0).
inline void blindRotateLazyPipeNtt(Trlwe& accum, const BootstrappingKeyMPLazyPipe& bsk, const ScaledTlwe& input, const TorusPolynomial& v,
                                   const YatfheParameters& param) {
    using namespace yatfhe_dropin;
    std::pair<ygpu_ctx*, int32_t> key;
    {
        State& s = state();
        std::lock_guard<std::mutex> lk(s.mu);
        require_init();
        const YatfheParameters& p = s.param;
        if (bsk.n != p.n || bsk.bskDecompB.size() != static_cast<size_t>(p.n - 1) * bsk.level)
            throw std::invalid_argument("lazy-pipe key has the wrong number of components");
        const size_t N = p.N, K = p.k, R = bsk.level, D = bsk.bskDecompB.at(0).size(), m = static_cast<size_t>(p.n - 1);
        const size_t per = R * D * N;
        const uint64_t fp = fingerprint(m * per + 1, [&](size_t idx) -> uint64_t {
            if (idx == m * per) return static_cast<uint64_t>(static_cast<uint32_t>(bsk.bskFirst.at(0).b.coeffs.at(0)));
            const size_t cell = idx / (D * N), l = (idx / N) % D, j = idx % N;
            return static_cast<uint64_t>(static_cast<uint8_t>(bsk.bskDecompB.at(cell).at(l).coeffs.at(j))) |
                   (static_cast<uint64_t>(static_cast<uint8_t>(bsk.bskDecompA.at(cell).at(l).at(0).coeffs.at(j))) << 8);
        });
        ygpu_ctx* ctx = context_locked(bsk.level, param.driftPhase);
        const auto ck = std::make_tuple(ctx, static_cast<const void*>(&bsk), fp, static_cast<const void*>(nullptr), uint64_t{2});
        auto it = s.keys.find(ck);
        if (it != s.keys.end()) {
            key = {ctx, it->second};
        } else {
            std::vector<int32_t> first((K + 1) * N);
            for (size_t c = 0; c <= K; c++) {
                const TorusPolynomial& t = c < K ? bsk.bskFirst.at(0).a.at(c) : bsk.bskFirst.at(0).b;
                for (size_t j = 0; j < N; j++) first[c * N + j] = static_cast<int32_t>(t.coeffs.at(j));
            }
            std::vector<int8_t> da(m * R * D * K * N), db(m * R * D * N);
            for (size_t cell = 0; cell < m * R; cell++)
                for (size_t l = 0; l < D; l++) {
                    for (size_t c = 0; c < K; c++)
                        for (size_t j = 0; j < N; j++)
                            da[((cell * D + l) * K + c) * N + j] = static_cast<int8_t>(bsk.bskDecompA.at(cell).at(l).at(c).coeffs.at(j));
                    for (size_t j = 0; j < N; j++) db[(cell * D + l) * N + j] = static_cast<int8_t>(bsk.bskDecompB.at(cell).at(l).coeffs.at(j));
                }
            std::vector<uint64_t> s2(D * (K + 1) * N);
            for (size_t l = 0; l < D; l++)
                for (size_t c = 0; c <= K; c++) {
                    const TrlweDft& t = bsk.s2Dft.trlweDfts.at(l);
                    std::memcpy(s2.data() + (l * (K + 1) + c) * N, (c < K ? t.a.at(c) : t.b).coeffs.data(), sizeof(uint64_t) * N);
                }
            std::vector<int32_t> zero(K * N * p.ksLevel * (p.n + 1), 0);
            int32_t id = -1;
            check(ygpu_key_upload_lazy(ctx, static_cast<int32_t>(D), first.data(), da.data(), db.data(), s2.data(), zero.data(), &id));
            s.keys[ck] = id;
            key = {ctx, id};
        }
    }
    lazy_rotate_on(key, accum, input, v);
}

// blindRotateLazyPipeAltNtt (blind_rotate.cpp:996-1000, the sweep at 1002-1136).
inline void blindRotateLazyPipeAltNtt(Trlwe& accum, const BootstrappingKeyMPLazyPipeAlt& bsk, const ScaledTlwe& input,
                                      const TorusPolynomial& v, const std::vector<NttPolynomial>& gdVntt, const YatfheParameters& param) {
    using namespace yatfhe_dropin;
    if (static_cast<int>(gdVntt.size()) != param.lApprox) throw std::invalid_argument("blindRotateLazyPipeAltNtt: gdVntt needs lApprox levels");
    std::pair<ygpu_ctx*, int32_t> key;
    {
        State& s = state();
        std::lock_guard<std::mutex> lk(s.mu);
        require_init();
        const YatfheParameters& p = s.param;
        if (bsk.n != p.n || bsk.bskPrime.size() != static_cast<size_t>(p.n)) throw std::invalid_argument("Alt key has the wrong number of components");
        const size_t N = p.N, K = p.k, R = bsk.level, D = bsk.s2Dft.l;
        const size_t per = R * (K + 1) * N;
        const uint64_t fp = fingerprint(bsk.bskPrime.size() * per, [&](size_t idx) -> uint64_t {
            const size_t i = idx / per, lvl = (idx / ((K + 1) * N)) % R, c = (idx / N) % (K + 1), j = idx % N;
            const Trlwe& t = bsk.bskPrime.at(i).cPrime.at(lvl);
            return static_cast<uint64_t>(static_cast<uint32_t>((c < K ? t.a.at(c) : t.b).coeffs.at(j)));
        });
        ygpu_ctx* ctx = context_locked(bsk.level, param.driftPhase);
        const auto ck = std::make_tuple(ctx, static_cast<const void*>(&bsk), fp, static_cast<const void*>(nullptr), uint64_t{3});
        auto it = s.keys.find(ck);
        if (it != s.keys.end()) {
            key = {ctx, it->second};
        } else {
            std::vector<int32_t> rows(static_cast<size_t>(p.n) * per);
            size_t o = 0;
            for (const auto& g : bsk.bskPrime)
                for (size_t lvl = 0; lvl < R; lvl++)
                    for (size_t c = 0; c <= K; c++) {
                        const TorusPolynomial& t = c < K ? g.cPrime.at(lvl).a.at(c) : g.cPrime.at(lvl).b;
                        for (size_t j = 0; j < N; j++) rows[o++] = static_cast<int32_t>(t.coeffs.at(j));
                    }
            std::vector<uint64_t> s2(D * (K + 1) * N);
            for (size_t l = 0; l < D; l++)
                for (size_t c = 0; c <= K; c++) {
                    const TrlweDft& t = bsk.s2Dft.trlweDfts.at(l);
                    std::memcpy(s2.data() + (l * (K + 1) + c) * N, (c < K ? t.a.at(c) : t.b).coeffs.data(), sizeof(uint64_t) * N);
                }
            std::vector<int32_t> zero(K * N * p.ksLevel * (p.n + 1), 0);
            int32_t id = -1;
            check(ygpu_key_upload_lazy_alt(ctx, static_cast<int32_t>(D), rows.data(), s2.data(), zero.data(), &id));
            s.keys[ck] = id;
            key = {ctx, id};
        }
    }
    lazy_rotate_on(key, accum, input, v);
}

// extractTlweFromTrlwe (trlwe.cpp:241-250) at `index`, on the GPU
inline void extractTlweFromTrlwe(Tlwe& out, const Trlwe& in, const int index) {
    using namespace yatfhe_dropin;
    ygpu_ctx* ctx;
    {
        State& s = state();
        std::lock_guard<std::mutex> lk(s.mu);
        require_init();
        ctx = context_locked(s.param.lApprox, s.param.driftPhase);
    }
    std::vector<Torus> acc;
    to_flat(in, acc);
    std::vector<int32_t> o(static_cast<size_t>(in.k) * in.N + 1);
    check(ygpu_sample_extract_index(ctx, acc.data(), index, o.data()));
    for (int i = 0; i < in.k * in.N; i++) out.a[i] = o[i];
    out.b = o[static_cast<size_t>(in.k) * in.N];
}

// switchKeyForTlwe (keyswitching.cpp:48-72)
inline void switchKeyForTlwe(Tlwe& output, const TlweKeySwitchingKey& ksk, const Tlwe& input, const YatfheParameters& param) {
    using namespace yatfhe_dropin;
    std::pair<ygpu_ctx*, int32_t> key;
    {
        State& s = state();
        std::lock_guard<std::mutex> lk(s.mu);
        require_init();
        // a KSK travels with a BSK on the device: pair it with an all-zero BSK
        static const int zeroTag = 0;
        const YatfheParameters& p = s.param;
        const int level = p.lApprox;
        key = device_key_locked(level, param.driftPhase, &zeroTag, 0,
                                [&](std::vector<uint64_t>& out) { out.assign(static_cast<size_t>(p.n) * level * (p.k + 1) * (p.k + 1) * p.N, 0); },
                                &ksk);
    }
    std::vector<int32_t> in(input.n + 1), out(output.n + 1);
    for (int i = 0; i < input.n; i++) in[i] = static_cast<int32_t>(input.a[i]);
    in[input.n] = static_cast<int32_t>(input.b);
    check(ygpu_keyswitch(key.first, key.second, in.data(), out.data()));
    for (int i = 0; i < output.n; i++) output.a[i] = out[i];
    output.b = out[output.n];
}

// functionalBootstrappingMPNtt (bootstrapping.cpp:71-80): rescale, accumulator
// init, blind rotation at the key's level, extraction at param.driftPhase and
// keyswitch, in one GPU call.
inline void functionalBootstrappingMPNtt(Tlwe& out, const Tlwe& input, const BootstrappingKeyMP& bsk, const TlweKeySwitchingKey& ksk,
                                         const TorusPolynomial& v, const YatfheParameters& param) {
    yatfhe_dropin::bootstrap_on(yatfhe_dropin::mp_key(bsk.bskDft, &ksk, param), out, input, v);
}
// functionalBootstrappingNtt (bootstrapping.cpp:60-69), classic keys, l == lApprox
inline void functionalBootstrappingNtt(Tlwe& out, const Tlwe& input, const BootstrappingKey& bsk, const TlweKeySwitchingKey& ksk,
                                       const TorusPolynomial& v, const YatfheParameters& param) {
    yatfhe_dropin::bootstrap_on(yatfhe_dropin::classic_key(bsk.bskDft, &ksk, param), out, input, v);
}
