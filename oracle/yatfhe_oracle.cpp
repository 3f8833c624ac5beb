// yatfhe_oracle.cpp -- CPU oracle for the TFHE programmable-bootstrapping path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU, the arithmetic of
// the reference library's bootstrapping path (xsong93/yatfhe) so that the HIP
// engine in yatfhe_amd/ can be checked against it bit for bit.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// as the checker or as the timed CPU baseline; the product never calls it.
//
// Every function names the reference file:line it follows (paths relative to
// the reference repository root).  Torus values are carried as int64_t holding
// the sign-extended w-bit value (w = 32 or 56): the reference's TORUS32 int32
// wrap-around arithmetic equals int64 arithmetic followed by sign extension from
// bit 31, which is how every result here is formed.
//
// Pinning (what checks each part independently of this text):
//   modswitch (R9)          exhaustive closed-form check over all 2^32 inputs
//   rotation                the reference's POLY.ROT known-answer table
//   digits                  two restatements (scalar zeroMeanDigits form and the
//                           level-major row form) compared exhaustively, plus the
//                           reference's own header compiled under oracle/_ref
//   NTT / MAC / INTT        O(N^2) schoolbook negacyclic product mod q
//   external product        schoolbook route vs NTT route, bit for bit
//   blind rotate / extract  decryption identities (phase = X^{-b+sum a s} v)
//   keyswitch               int64 and uint32-wrap accumulations cross-checked,
//                           decryption of every message slot
//   succinct-key derivation schoolbook route vs NTT route, bit for bit; every
//                           derived row decrypts to -s_i 2^(w-(lvl+1)b) z (the RGSW
//                           mask-digit row) within the noise bound; blind rotation
//                           with the derived key decrypts to X^{sum a s} v
//   level-parallel PBS      equal to the single-threaded PBS, bit for bit
//   lazy-pipe rows          schoolbook route vs NTT route, bit for bit; the rows
//                           of component i rotated by a decrypt to the RGSW of
//                           1 + (X^a - 1) s_{i+1} (gadget times the message, within
//                           the digits' approximation); narrowing of rotated
//                           digits pinned on planted +-128 values
//   lazy-pipe rotation      schoolbook vs NTT route; decryption identity
//                           (phase = X^{-b + sum a s} v) over whole PBSs
//   lazy-pipe Alt           rows and first component: schoolbook vs NTT route;
//                           rows decrypt to the RGSW of 1 + (X^a - 1) s_i, the
//                           first component to v s_0; whole PBSs decrypt
#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <thread>
#include <atomic>
#include <algorithm>
#include <stdexcept>
#include <functional>
#include <chrono>

typedef unsigned __int128 u128;
typedef __int128 i128;

extern "C" {
typedef struct or_params {
    int32_t torus_bits;     // w: 32 (TORUS32) or 56 (TORUS56)
    int32_t n;              // LWE dimension
    int32_t k;              // GLWE dimension
    int32_t N;              // ring degree
    int32_t radix_bits;     // gadget base log2 (param.radixBits)
    int32_t level;          // gadget levels used (lApprox == l)
    int32_t ks_radix_bits;  // param.ksRadixBits
    int32_t ks_level;       // param.ksLevel
    int32_t ks_width_bits;  // param.ksWidthBits
    int32_t lwe_q_bits;     // log2 qLwe (32)
    uint64_t q_ntt;         // param.qNtt
    int32_t extract_index;  // param.driftPhase: the coefficient extractTlweFromTrlwe takes
    int32_t reserved;
} or_params;
}

namespace {

// ---------------------------------------------------------------- scalar rules
// longModPow2 (numeric.h:69-71): sign extension from the low (64-shift) bits.
inline int64_t sext(int64_t x, int w) {
    const int shift = 64 - w;
    return static_cast<int64_t>(static_cast<uint64_t>(x) << shift) >> shift;
}

// longModP (numeric.h:53-66), general modulus branch included for completeness.
inline int64_t long_mod_p(int64_t a, int64_t p) {
    if (p > 1 && (p & (p - 1)) == 0) {
        const int shift = 64 - __builtin_ctzll(static_cast<uint64_t>(p));
        return static_cast<int64_t>(static_cast<uint64_t>(a) << shift) >> shift;
    }
    int64_t b = a % p;
    if (b > (p - 1) / 2) b -= p;
    else if (b < -p / 2) b += p;
    return b;
}

// modSwitchFromTorusGeneral (numeric.cpp:189-208).
inline int64_t mod_switch_from_torus(int64_t in, int64_t newMod, int64_t torusQ) {
    if (torusQ >= newMod) {
        const int64_t scale = torusQ / newMod;
        const int64_t num = in;
        const int64_t real = (num >= 0) ? (num + scale / 2) / scale : (num - scale / 2) / scale;
        return long_mod_p(real, newMod);
    }
    const int64_t scaleUp = newMod / torusQ;
    if (scaleUp * torusQ == newMod) return long_mod_p(in * scaleUp, newMod);
    const double div = static_cast<double>(in) * static_cast<double>(newMod) / static_cast<double>(torusQ);
    const auto real = static_cast<int32_t>(std::round(div));
    return long_mod_p(static_cast<int64_t>(real), newMod);
}

// validateRotator (polynomial.cpp:22-29).
inline void validate_rotator(int& aTrue, int& isWrap, int a, int N) {
    aTrue = a % (2 * N);
    if (aTrue < 0) aTrue += 2 * N;
    isWrap = (aTrue < N) ? 1 : -1;
    aTrue = (aTrue < N) ? aTrue : aTrue - N;
}

// rotateTorusPolynomial (polynomial.cpp:184-193): out = X^a * in.
void rotate(int w, int N, int a, const int64_t* in, int64_t* out) {
    int aTrue, isWrap;
    validate_rotator(aTrue, isWrap, a, N);
    for (int i = 0; i < N; i++) {
        const int64_t tmp = (i < aTrue) ? (-in[i - aTrue + N] * isWrap) : (in[i - aTrue] * isWrap);
        out[i] = sext(tmp, w);
    }
}

// rotateTorusPolynomialMinusOne (polynomial.cpp:213-238): out = (X^a - 1) * in,
// power-of-two torus branch (subTorusPow2, numeric.h:74-76).
void rotate_minus_one(int w, int N, int a, const int64_t* in, int64_t* out) {
    int aTrue, isWrap;
    validate_rotator(aTrue, isWrap, a, N);
    for (int i = 0; i < aTrue; i++) {
        const int64_t tmp = sext(-in[i - aTrue + N] * isWrap, w);
        out[i] = sext(tmp - in[i], w);
    }
    for (int i = aTrue; i < N; i++) {
        const int64_t tmp = sext(in[i - aTrue] * isWrap, w);
        out[i] = sext(tmp - in[i], w);
    }
}

// ---------------------------------------------------------------- digits
// roundingBias / maskToWidth / windowShift (gadget_decomposition.h:52-68).
inline uint64_t width_mask(int w) { return w >= 64 ? ~0ULL : ((1ULL << w) - 1); }
inline uint64_t rounding_bias(int b, int w, int l) {
    const int dropped = w - l * b;
    return dropped > 0 ? (1ULL << (dropped - 1)) : 0;
}

// Form 1: zeroMeanDigits (gadget_decomposition.h:75-95), scalar, digit per level.
void digits_scalar(int64_t in, int l, int b, int w, int64_t* out) {
    const uint64_t mask = width_mask(w);
    const uint64_t u = static_cast<uint64_t>(in) & mask;
    const bool negative = (u >> (w - 1)) & 1;
    const uint64_t m = (negative ? (0ULL - u) & mask : u) + rounding_bias(b, w, l);
    const int64_t B = int64_t{1} << b;
    const int64_t halfB = B >> 1;
    const int64_t sign = negative ? -1 : 1;
    int64_t carry = 0;
    for (int level = l - 1; level >= 0; --level) {
        const int shift = w - (level + 1) * b;
        const int64_t window = static_cast<int64_t>((m >> shift) & static_cast<uint64_t>(B - 1));
        const int64_t digit = window + carry;
        const int64_t tieBit = level > 0 ? static_cast<int64_t>((m >> (shift + 2 * b - 1)) & 1) : 0;
        carry = (digit > halfB) | ((digit == halfB) & tieBit);
        out[level] = sign * (carry ? digit - B : digit);
    }
}

// Form 2: decomposeRowByLevel (gadget_decomposition.h:110-152), level-major over
// a whole row with flat per-value state.  out is [l][N].
void digits_row(const int64_t* in, int N, int l, int b, int w, int64_t* out) {
    const int64_t B = int64_t{1} << b;
    const int64_t halfB = B >> 1;
    const uint64_t mask = width_mask(w);
    const uint64_t round = rounding_bias(b, w, l);
    std::vector<uint64_t> mag(N);
    std::vector<int64_t> sgn(N), car(N);
    for (int j = 0; j < N; j++) {
        const uint64_t value = static_cast<uint64_t>(in[j]) & mask;
        const bool negative = (value >> (w - 1)) & 1;
        mag[j] = (negative ? (0ULL - value) & mask : value) + round;
        sgn[j] = negative ? -1 : 1;
        car[j] = 0;
    }
    for (int level = l - 1; level >= 0; --level) {
        const int shift = w - (level + 1) * b;
        const int tieShift = level > 0 ? shift + 2 * b - 1 : 0;
        const int64_t tieMask = level > 0 ? 1 : 0;
        int64_t* o = out + static_cast<size_t>(level) * N;
        for (int j = 0; j < N; j++) {
            const int64_t window = static_cast<int64_t>((mag[j] >> shift) & static_cast<uint64_t>(B - 1));
            const int64_t digit = window + car[j];
            const int64_t tieBit = static_cast<int64_t>((mag[j] >> tieShift) & 1) & tieMask;
            car[j] = (digit > halfB) | ((digit == halfB) & tieBit);
            o[j] = (digit - car[j] * B) * sgn[j];
        }
    }
}

// ---------------------------------------------------------------- mod q
inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) {
    return static_cast<uint64_t>((static_cast<u128>(a) * b) % q);
}
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
// Shoup's multiplication by a fixed operand w with wp = floor(w * 2^64 / q).
inline uint64_t shoup(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
    const uint64_t hi = static_cast<uint64_t>((static_cast<u128>(x) * wp) >> 64);
    uint64_t r = x * w - hi * q;
    return r >= q ? r - q : r;
}
inline uint64_t shoup_pre(uint64_t w, uint64_t q) {
    return static_cast<uint64_t>((static_cast<u128>(w) << 64) / q);
}

// Smallest generator of Z_q^*, q prime with q-1 = 2^e * (small odd part).
uint64_t find_generator(uint64_t q) {
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

inline int bitrev(int x, int bits) {
    int r = 0;
    for (int i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

// Textbook negacyclic NTT mod q (Cooley-Tukey forward, Gentleman-Sande inverse,
// psi powers merged).  The reference uses Intel HEXL here (ntt_hexl.h:35-44,
// ntt_hexl.cpp:153-175); the transform order and root are internal to each
// implementation and never cross the interface: only the negacyclic product mod
// q, which is unique, does.
struct Ntt {
    uint64_t q = 0;
    int N = 0, logN = 0;
    std::vector<uint64_t> psi, psip, ipsi, ipsip;  // bit-reversed powers, Shoup constants
    uint64_t ninv = 0, ninvp = 0;

    Ntt(uint64_t q_, int N_) : q(q_), N(N_) {
        while ((1 << logN) < N) logN++;
        const uint64_t g = find_generator(q);
        const uint64_t root = powmod(g, (q - 1) / (2 * static_cast<uint64_t>(N)), q);
        if (powmod(root, N, q) != q - 1) throw std::runtime_error("no primitive 2N-th root");
        const uint64_t iroot = powmod(root, q - 2, q);
        psi.resize(N); psip.resize(N); ipsi.resize(N); ipsip.resize(N);
        for (int i = 0; i < N; i++) {
            const int e = bitrev(i, logN);
            psi[i] = powmod(root, e, q);
            psip[i] = shoup_pre(psi[i], q);
            ipsi[i] = powmod(iroot, e, q);
            ipsip[i] = shoup_pre(ipsi[i], q);
        }
        ninv = powmod(N, q - 2, q);
        ninvp = shoup_pre(ninv, q);
    }
    void forward(uint64_t* a) const {
        for (int m = 1, t = N / 2; m < N; m <<= 1, t >>= 1) {
            for (int i = 0; i < m; i++) {
                const uint64_t w = psi[m + i], wp = psip[m + i];
                for (int j = 2 * i * t; j < 2 * i * t + t; j++) {
                    const uint64_t U = a[j];
                    const uint64_t V = shoup(a[j + t], w, wp, q);
                    uint64_t s = U + V; if (s >= q) s -= q;
                    a[j] = s;
                    a[j + t] = U >= V ? U - V : U + q - V;
                }
            }
        }
    }
    void inverse(uint64_t* a) const {
        for (int m = N / 2, t = 1; m >= 1; m >>= 1, t <<= 1) {
            for (int i = 0; i < m; i++) {
                const uint64_t w = ipsi[m + i], wp = ipsip[m + i];
                for (int j = 2 * i * t; j < 2 * i * t + t; j++) {
                    const uint64_t U = a[j], V = a[j + t];
                    uint64_t s = U + V; if (s >= q) s -= q;
                    a[j] = s;
                    a[j + t] = shoup(U >= V ? U - V : U + q - V, w, wp, q);
                }
            }
        }
        for (int j = 0; j < N; j++) a[j] = shoup(a[j], ninv, ninvp, q);
    }
};

// applyNtt's lift (ntt_hexl.h:38-42): signed v -> v + q if v < 0.
inline uint64_t lift(int64_t v, uint64_t q) {
    return static_cast<uint64_t>(v) + (v < 0 ? q : 0);
}
// applyIntt's exit (ntt_hexl.cpp:156-167): x in [0,q); x >= (q+1)/2 -> x - q;
// then sign-extend the low w bits.
inline int64_t intt_exit(uint64_t x, uint64_t q, int w) {
    const uint64_t halfQ = (q + 1) >> 1;
    const uint64_t centred = x - (x >= halfQ ? q : 0);
    return sext(static_cast<int64_t>(centred), w);
}

// Brute-force negacyclic product mod q: out = a * b in Z_q[X]/(X^N+1), a signed
// (small or torus), b residues in [0,q).  O(N^2), exact, no transform.
void schoolbook_modq(uint64_t q, int N, const int64_t* a, const uint64_t* b, uint64_t* out) {
    std::vector<i128> acc(N, 0);
    for (int i = 0; i < N; i++) {
        if (a[i] == 0) continue;
        const i128 ai = a[i];
        for (int j = 0; j < N; j++) {
            const i128 p = ai * static_cast<i128>(b[j]);
            const int d = i + j;
            if (d < N) acc[d] += p; else acc[d - N] -= p;
        }
    }
    for (int i = 0; i < N; i++) {
        i128 r = acc[i] % static_cast<i128>(q);
        if (r < 0) r += q;
        out[i] = static_cast<uint64_t>(r);
    }
}

// ---------------------------------------------------------------- keys
struct Key {
    or_params p;
    Ntt ntt;
    // BSK in this oracle's NTT order, [n][l][k+1 rows][k+1 cols][N], with Shoup constants.
    std::vector<uint64_t> bsk, bskp;
    // Coefficient-domain residues, kept for the schoolbook route.
    std::vector<uint64_t> coeff;
    Key(const or_params& p_, const uint64_t* bsk_coeff) : p(p_), ntt(p_.q_ntt, p_.N) {
        const size_t per = static_cast<size_t>(p.level) * (p.k + 1) * (p.k + 1) * p.N;
        const size_t total = per * p.n;
        coeff.assign(bsk_coeff, bsk_coeff + total);
        bsk = coeff;
        bskp.resize(total);
        for (size_t o = 0; o < total; o += p.N) ntt.forward(&bsk[o]);
        for (size_t i = 0; i < total; i++) bskp[i] = shoup_pre(bsk[i], p.q_ntt);
    }
};

// externalProductTrgswMPNttImpl (trgsw.cpp:712-767) for one BSK row, k+1 torus
// polynomials in, k+1 out.  Decomposition into exactly `level` digits
// (gadgetDecomposeTrlwe, trlwe.cpp:185-198, with DecomposedTrlwe{param, level},
// trgsw.cpp:700); row r < k is the digit of a_r, row k the digit of b; MAC order
// is irrelevant mod q (HEXL Eltwise*Mod are exact); one INTT per output column.
void external_product(const Key& key, size_t i, const int64_t* in, int64_t* out, bool schoolbook) {
    const or_params& p = key.p;
    const int N = p.N, K1 = p.k + 1, L = p.level, w = p.torus_bits;
    const uint64_t q = p.q_ntt;
    std::vector<int64_t> dig(static_cast<size_t>(K1) * L * N);   // [row][lvl][N]
    for (int r = 0; r < K1; r++) digits_row(in + static_cast<size_t>(r) * N, N, L, p.radix_bits, w, &dig[static_cast<size_t>(r) * L * N]);
    std::vector<uint64_t> acc(static_cast<size_t>(K1) * N, 0), tmp(N);
    const size_t rowBase = i * static_cast<size_t>(L) * K1 * K1 * N;
    for (int lvl = 0; lvl < L; lvl++) {
        for (int r = 0; r < K1; r++) {
            const int64_t* d = &dig[(static_cast<size_t>(r) * L + lvl) * N];
            const size_t keyOff = rowBase + (static_cast<size_t>(lvl) * K1 + r) * K1 * N;
            if (schoolbook) {
                for (int c = 0; c < K1; c++) {
                    schoolbook_modq(q, N, d, &key.coeff[keyOff + static_cast<size_t>(c) * N], tmp.data());
                    for (int j = 0; j < N; j++) {
                        uint64_t s = acc[static_cast<size_t>(c) * N + j] + tmp[j];
                        acc[static_cast<size_t>(c) * N + j] = s >= q ? s - q : s;
                    }
                }
            } else {
                for (int j = 0; j < N; j++) tmp[j] = lift(d[j], q);
                key.ntt.forward(tmp.data());
                for (int c = 0; c < K1; c++) {
                    const uint64_t* kb = &key.bsk[keyOff + static_cast<size_t>(c) * N];
                    const uint64_t* kp = &key.bskp[keyOff + static_cast<size_t>(c) * N];
                    uint64_t* a = &acc[static_cast<size_t>(c) * N];
                    for (int j = 0; j < N; j++) {
                        uint64_t s = a[j] + shoup(tmp[j], kb[j], kp[j], q);
                        a[j] = s >= q ? s - q : s;
                    }
                }
            }
        }
    }
    for (int c = 0; c < K1; c++) {
        uint64_t* a = &acc[static_cast<size_t>(c) * N];
        if (!schoolbook) key.ntt.inverse(a);
        for (int j = 0; j < N; j++) out[static_cast<size_t>(c) * N + j] = intt_exit(a[j], q, w);
    }
}

// blindRotateMP21Ntt (blind_rotate.cpp:1538-1565), binary-key branch:
// skip a_i == 0; tmp = (X^{a_i} - 1) acc; tmp = ExtProd(BSK_i, tmp); acc += tmp.
void blind_rotate(const Key& key, const int32_t* abar, int64_t* acc, bool schoolbook) {
    const or_params& p = key.p;
    const int N = p.N, K1 = p.k + 1, w = p.torus_bits;
    std::vector<int64_t> tmp(static_cast<size_t>(K1) * N), ext(static_cast<size_t>(K1) * N);
    for (int i = 0; i < p.n; i++) {
        if (abar[i] == 0) continue;
        for (int r = 0; r < K1; r++) rotate_minus_one(w, N, abar[i], acc + static_cast<size_t>(r) * N, &tmp[static_cast<size_t>(r) * N]);
        external_product(key, i, tmp.data(), ext.data(), schoolbook);
        // accumulateTrlwe -> accumulateTorusPolynomial (trlwe.h:268-274, polynomial.h:203-217)
        for (size_t j = 0; j < tmp.size(); j++) acc[j] = sext(acc[j] + ext[j], w);
    }
}

// switchTrlweToSecretEmbeddingNtt (trgsw.cpp:861-909) for one (step, level) of
// a WWL+24 succinct key (BootstrappingKeyWWL24, bootstrapping.h:116-136), in
// coefficient form.  cp: the body-digit row cPrime[lvl] [k+1][N]; s2: the TRLEV
// of s^2 [s2_level][k+1][N]; out: the full RGSW rows of this level [k+1 rows][k+1
// cols][N].  Row k1 < k is derived: the torus values of cPrime.a[k1] (applyIntt:
// centre, low w bits), decomposed into s2_level digits (decomposeRow with
// param.l levels), and
//   c[k1].a[k2] = sum_l2 D_l2 * s2[l2].a[k2] + cPrime.b,   c[k1].b = sum_l2 D_l2 * s2[l2].b   (mod q);
// row k is cPrime itself.  The products are negacyclic mod q, by this oracle's
// NTT (the reference's route, one inverse per output column) or by schoolbook.
void derive_succinct_row(const or_params& p, const Ntt& ntt, const std::vector<uint64_t>& s2ntt, int s2_level,
                         const uint64_t* cp, const uint64_t* s2, uint64_t* out, bool schoolbook) {
    const int N = p.N, K = p.k, K1 = p.k + 1, w = p.torus_bits;
    const uint64_t q = p.q_ntt;
    std::vector<int64_t> ta(N), dig(static_cast<size_t>(s2_level) * N);
    std::vector<uint64_t> acc(static_cast<size_t>(K1) * N), tmp(N), bntt(cp + static_cast<size_t>(K) * N, cp + static_cast<size_t>(K1) * N);
    if (!schoolbook) ntt.forward(bntt.data());
    for (int k1 = 0; k1 < K; k1++) {
        for (int j = 0; j < N; j++) ta[j] = intt_exit(cp[static_cast<size_t>(k1) * N + j], q, w);
        digits_row(ta.data(), N, s2_level, p.radix_bits, w, dig.data());
        std::fill(acc.begin(), acc.end(), 0);
        for (int l2 = 0; l2 < s2_level; l2++) {
            const int64_t* d = &dig[static_cast<size_t>(l2) * N];
            if (schoolbook) {
                for (int c = 0; c < K1; c++) {
                    schoolbook_modq(q, N, d, s2 + (static_cast<size_t>(l2) * K1 + c) * N, tmp.data());
                    for (int j = 0; j < N; j++) acc[static_cast<size_t>(c) * N + j] = (acc[static_cast<size_t>(c) * N + j] + tmp[j]) % q;
                }
            } else {
                for (int j = 0; j < N; j++) tmp[j] = lift(d[j], q);
                ntt.forward(tmp.data());
                for (int c = 0; c < K1; c++) {
                    const uint64_t* sk = &s2ntt[(static_cast<size_t>(l2) * K1 + c) * N];
                    for (int j = 0; j < N; j++) {
                        const uint64_t prod = static_cast<uint64_t>(static_cast<u128>(tmp[j]) * sk[j] % q);
                        acc[static_cast<size_t>(c) * N + j] = (acc[static_cast<size_t>(c) * N + j] + prod) % q;
                    }
                }
            }
        }
        for (int c = 0; c < K; c++) {
            const uint64_t* b = schoolbook ? cp + static_cast<size_t>(K) * N : bntt.data();
            for (int j = 0; j < N; j++) acc[static_cast<size_t>(c) * N + j] = (acc[static_cast<size_t>(c) * N + j] + b[j]) % q;
        }
        for (int c = 0; c < K1; c++) {
            if (!schoolbook) ntt.inverse(&acc[static_cast<size_t>(c) * N]);
            std::memcpy(out + (static_cast<size_t>(k1) * K1 + c) * N, &acc[static_cast<size_t>(c) * N], sizeof(uint64_t) * N);
        }
    }
    std::memcpy(out + static_cast<size_t>(K) * K1 * N, cp, sizeof(uint64_t) * K1 * N);
}

std::vector<uint64_t> s2_transformed(const Ntt& ntt, int s2_level, int K1, int N, const uint64_t* s2) {
    std::vector<uint64_t> t(s2, s2 + static_cast<size_t>(s2_level) * K1 * N);
    for (size_t o = 0; o < t.size(); o += N) ntt.forward(&t[o]);
    return t;
}

int64_t torus_q(const or_params& p) { return p.torus_bits >= 63 ? 0 : (int64_t{1} << p.torus_bits); }
int64_t lwe_q(const or_params& p) { return int64_t{1} << p.lwe_q_bits; }

// rescaleTlweToNewMod (tlwe.cpp:75-81) to mod 2N.
void rescale(const or_params& p, const int32_t* lwe, int32_t* out) {
    const int64_t m = 2 * static_cast<int64_t>(p.N);
    for (int i = 0; i <= p.n; i++) out[i] = static_cast<int32_t>(mod_switch_from_torus(lwe[i], m, lwe_q(p)));
}

// genNoiselessTrlweSample (trlwe.cpp:175-182): acc = (0, X^{-bbar} v).
void init_acc(const or_params& p, const int64_t* v, int32_t bbar, int64_t* acc) {
    const int N = p.N;
    std::fill(acc, acc + static_cast<size_t>(p.k) * N, 0);
    rotate(p.torus_bits, N, -bbar, v, acc + static_cast<size_t>(p.k) * N);
}

// extractTlweFromTrlwe(out, in, index) (trlwe.cpp:241-250) at index =
// p.extract_index, with the torus -> LWE_Q modswitch: a'_{iN+j} = a_i[index - j]
// for j <= index, else -a_i[N + index - j] (negated in the Torus type, so
// INT32_MIN wraps at 32 bits); b' = b[index].
void extract(const or_params& p, const int64_t* acc, int32_t* out) {
    const int N = p.N, w = p.torus_bits, index = p.extract_index;
    const int64_t tq = torus_q(p), lq = lwe_q(p);
    for (int i = 0; i < p.k; i++) {
        const int64_t* a = acc + static_cast<size_t>(i) * N;
        for (int j = 0; j < N; j++) {
            const int64_t x = (j <= index) ? a[index - j] : sext(-a[N + index - j], w == 32 ? 32 : 64);
            out[i * N + j] = static_cast<int32_t>(mod_switch_from_torus(x, lq, tq));
        }
    }
    out[p.k * N] = static_cast<int32_t>(mod_switch_from_torus(acc[static_cast<size_t>(p.k) * N + index], lq, tq));
}

// switchKeyForTlwe (keyswitching.cpp:48-72) with signedGadgetDecompositionKs
// (gadget_decomposition.cpp:53-56).  KSK layout [kN][t][n+1] (a[0..n-1], b).
void keyswitch(const or_params& p, const int32_t* ksk, const int32_t* in, int32_t* out) {
    const int nOut = p.n, nIn = p.k * p.N, t = p.ks_level;
    std::vector<int64_t> delta(nOut, 0);
    int64_t delta_b = 0;
    std::vector<int64_t> d(t);
    for (int i = 0; i < nIn; i++) {
        digits_scalar(in[i], t, p.ks_radix_bits, p.ks_width_bits, d.data());
        for (int j = 0; j < t; j++) {
            const int64_t coeff = d[j];
            if (coeff == 0) continue;
            const int32_t* row = ksk + (static_cast<size_t>(i) * t + j) * (nOut + 1);
            for (int kk = 0; kk < nOut; kk++) delta[kk] += coeff * row[kk];
            delta_b += coeff * row[nOut];
        }
    }
    const int64_t lq = lwe_q(p);
    for (int kk = 0; kk < nOut; kk++) out[kk] = static_cast<int32_t>(long_mod_p(-delta[kk], lq));
    out[nOut] = static_cast<int32_t>(long_mod_p(static_cast<int64_t>(in[nIn]) - delta_b, lq));
}

// blindRotateMP21Ntt with the reference's threading of the external product
// (externalProductTrgswMPNttImpl, trgsw.cpp:725-760): the calling thread rotates
// and decomposes, then each gadget level's forward transforms and MAC into its
// own partial sum run on a worker of a fixed pool (ThreadPool::run + join per
// step), and the calling thread sums the partials and inverse-transforms.
// Bit-identical to blind_rotate (the MAC order is irrelevant mod q); used to
// time the reference's single-ciphertext latency form.
class LevelPool {
  public:
    explicit LevelPool(int workers) : n_(workers) {
        for (int w = 0; w < n_; w++) ts_.emplace_back([this, w] { loop(w); });
    }
    ~LevelPool() {
        stop_.store(true);
        gen_.fetch_add(1);
        for (auto& t : ts_) t.join();
    }
    // Runs fn(w) for w < n on the workers and waits for all of them.
    void run(const std::function<void(int)>& fn) {
        fn_ = &fn;
        done_.store(0);
        gen_.fetch_add(1, std::memory_order_acq_rel);
        while (done_.load(std::memory_order_acquire) < n_) __builtin_ia32_pause();
    }

  private:
    void loop(int w) {
        uint64_t seen = 0;
        for (;;) {
            uint64_t g;
            while ((g = gen_.load(std::memory_order_acquire)) == seen) __builtin_ia32_pause();
            seen = g;
            if (stop_.load()) return;
            (*fn_)(w);
            done_.fetch_add(1, std::memory_order_acq_rel);
        }
    }
    int n_;
    std::vector<std::thread> ts_;
    std::atomic<uint64_t> gen_{0};
    std::atomic<int> done_{0};
    std::atomic<bool> stop_{false};
    const std::function<void(int)>* fn_ = nullptr;
};

void blind_rotate_level_parallel(const Key& key, const int32_t* abar, int64_t* acc, LevelPool& pool) {
    const or_params& p = key.p;
    const int N = p.N, K1 = p.k + 1, L = p.level, w = p.torus_bits;
    const uint64_t q = p.q_ntt;
    std::vector<int64_t> tmp(static_cast<size_t>(K1) * N), dig(static_cast<size_t>(K1) * L * N);
    std::vector<std::vector<uint64_t>> partial(L, std::vector<uint64_t>(static_cast<size_t>(K1) * N));
    size_t step = 0;
    const std::function<void(int)> level = [&](int lvl) {
        std::vector<uint64_t>& part = partial[lvl];
        std::fill(part.begin(), part.end(), 0);
        std::vector<uint64_t> t(N);
        const size_t rowBase = step * static_cast<size_t>(L) * K1 * K1 * N;
        for (int r = 0; r < K1; r++) {
            const int64_t* d = &dig[(static_cast<size_t>(r) * L + lvl) * N];
            for (int j = 0; j < N; j++) t[j] = lift(d[j], q);
            key.ntt.forward(t.data());
            const size_t keyOff = rowBase + (static_cast<size_t>(lvl) * K1 + r) * K1 * N;
            for (int c = 0; c < K1; c++) {
                const uint64_t* kb = &key.bsk[keyOff + static_cast<size_t>(c) * N];
                const uint64_t* kp = &key.bskp[keyOff + static_cast<size_t>(c) * N];
                uint64_t* a = &part[static_cast<size_t>(c) * N];
                for (int j = 0; j < N; j++) {
                    uint64_t s = a[j] + shoup(t[j], kb[j], kp[j], q);
                    a[j] = s >= q ? s - q : s;
                }
            }
        }
    };
    std::vector<uint64_t> sum(N);
    for (int i = 0; i < p.n; i++) {
        if (abar[i] == 0) continue;
        for (int r = 0; r < K1; r++) {
            rotate_minus_one(w, N, abar[i], acc + static_cast<size_t>(r) * N, &tmp[static_cast<size_t>(r) * N]);
            digits_row(&tmp[static_cast<size_t>(r) * N], N, L, p.radix_bits, w, &dig[static_cast<size_t>(r) * L * N]);
        }
        step = static_cast<size_t>(i);
        pool.run(level);
        for (int c = 0; c < K1; c++) {
            for (int j = 0; j < N; j++) {
                uint64_t s = 0;
                for (int lvl = 0; lvl < L; lvl++) {
                    s += partial[lvl][static_cast<size_t>(c) * N + j];
                    if (s >= q) s -= q;
                }
                sum[j] = s;
            }
            key.ntt.inverse(sum.data());
            for (int j = 0; j < N; j++) {
                int64_t& a = acc[static_cast<size_t>(c) * N + j];
                a = sext(a + intt_exit(sum[j], q, w), w);
            }
        }
    }
}
// ---------------------------------------------------------------- lazy pipe
// BootstrappingKeyMPLazyPipe (bootstrapping.h:250-282) as genBootstrappingKeyMPLazyPipe
// (bootstrapping.cpp:269-320) fills it, binary keys.  Components i < n-1 stand
// for s_{i+1}; the rows have R = bsk.level = param.lApprox = p.level gadget
// levels (the accumulator's decomposition), the stored digits D = param.l.
//   first [k+1][N]        bskFirst[0], torus: TRLWE of v * s_0
//   da [n-1][R][D][k][N]  bskDecompA[decompIndex(i, lvl)][l][k]: digit l of cPrime[lvl].a[k]
//   db [n-1][R][D][N]     bskDecompB[decompIndex(i, lvl)][l]: digit l of cPrime[lvl].b
//   s2 [D][k+1][N]        s2Dft (TrlevDft{p, p.l}) as coefficient residues mod q
// Digits are the reference's Decomp type (torus.h:14-18): int8 at TORUS32, int32
// at TORUS56, so rotated digits narrow to it.
struct LazyKey {
    or_params p;
    int D;  // digit levels (param.l)
    Ntt ntt;
    std::vector<int64_t> first;
    std::vector<int32_t> da, db;
    std::vector<uint64_t> s2, s2ntt, g;
    LazyKey(const or_params& p_, int digit_levels, const int64_t* first_, const int32_t* da_, const int32_t* db_, const uint64_t* s2_)
        : p(p_), D(digit_levels), ntt(p_.q_ntt, p_.N) {
        const size_t N = p.N, K = p.k, K1 = p.k + 1, R = p.level, m = static_cast<size_t>(p.n - 1);
        first.assign(first_, first_ + K1 * N);
        da.assign(da_, da_ + m * R * D * K * N);
        db.assign(db_, db_ + m * R * D * N);
        s2.assign(s2_, s2_ + D * K1 * N);
        s2ntt = s2;
        for (size_t o = 0; o < s2ntt.size(); o += N) ntt.forward(&s2ntt[o]);
        // getNttGadgetRecomper(l) (ntt_hexl.cpp:122-149, over param.l levels): the
        // transform of the constant polynomial 2^(w - (l+1) b), which is that
        // constant at every point
        for (int l = 0; l < D; l++) g.push_back(lift(sext(int64_t{1} << (p.torus_bits - (l + 1) * p.radix_bits), p.torus_bits), p.q_ntt));
    }
};

// static_cast<Decomp> of an int (torus.h:14-18).
inline int64_t narrow_decomp(int64_t x, int w) { return w == 32 ? static_cast<int64_t>(static_cast<int8_t>(x)) : static_cast<int64_t>(static_cast<int32_t>(x)); }

// rotateDecompPolynomialMinusOne (polynomial.cpp:273-282): out = Decomp(X^a in - in).
void rotate_decomp_minus_one(int w, int N, int a, const int32_t* in, int64_t* out) {
    int aTrue, isWrap;
    validate_rotator(aTrue, isWrap, a, N);
    for (int i = 0; i < N; i++) {
        const int64_t tmp = ((i < aTrue) ? -static_cast<int64_t>(in[i - aTrue + N]) : static_cast<int64_t>(in[i - aTrue])) * isWrap;
        out[i] = narrow_decomp(tmp - in[i], w);
    }
}

// rotateBskComponent (blind_rotate.cpp:490-507) of component `comp` by a, then
// switchDecompTrlweToSecretEmbeddingNtt (trgsw.cpp:1069-1102) per level: the
// rows of the per-ciphertext RGSW [L][k+1 rows][k+1 cols][N].  Row r < k is
// c[lvl][r] (columns a[0..k-1], b), row k is cPrime[lvl]:
//   cPrime.b    = sum_l D(b)_l g_l          cPrime.a[k1] = sum_l D(a_k1)_l g_l
//   c[k1].a[k2] = sum_l D(a_k1)_l s2[l].a[k2] + cPrime.b     c[k1].b = sum_l D(a_k1)_l s2[l].b
// (mod q), D the rotated digits with the gadget's +1 on the diagonal digit
// (rotatedB[lvl][lvl] coefficient 0).  The NTT route leaves the rows in this
// oracle's transform domain; the schoolbook route forms them in coefficient
// form (negacyclic products by brute force).
void lazy_rows(const LazyKey& key, int comp, int a, std::vector<uint64_t>& rows, bool schoolbook) {
    const or_params& p = key.p;
    const int N = p.N, K = p.k, K1 = p.k + 1, L = p.level, D = key.D, w = p.torus_bits;
    const uint64_t q = p.q_ntt;
    rows.assign(static_cast<size_t>(L) * K1 * K1 * N, 0);
    std::vector<int64_t> d(N);
    std::vector<uint64_t> t(N), prod(N);
    auto addto = [&](uint64_t* dst, const uint64_t* src) {
        for (int j = 0; j < N; j++) dst[j] = (dst[j] + src[j]) % q;
    };
    for (int lvl = 0; lvl < L; lvl++) {
        uint64_t* R = &rows[static_cast<size_t>(lvl) * K1 * K1 * N];
        auto row = [&](int r, int c) { return R + (static_cast<size_t>(r) * K1 + c) * N; };
        const size_t cell = static_cast<size_t>(comp) * L + lvl;
        for (int l = 0; l < D; l++) {
            // the b digits: cPrime.b
            rotate_decomp_minus_one(w, N, a, &key.db[(cell * D + l) * N], d.data());
            if (l == lvl) d[0] = narrow_decomp(d[0] + 1, w);
            for (int j = 0; j < N; j++) t[j] = lift(d[j], q);
            if (!schoolbook) key.ntt.forward(t.data());
            for (int j = 0; j < N; j++) prod[j] = mulmod(t[j], key.g[l], q);
            addto(row(K, K), prod.data());
            // the a digits: cPrime.a[k1] and c[k1]
            for (int k1 = 0; k1 < K; k1++) {
                rotate_decomp_minus_one(w, N, a, &key.da[((cell * D + l) * K + k1) * N], d.data());
                for (int j = 0; j < N; j++) t[j] = lift(d[j], q);
                if (!schoolbook) key.ntt.forward(t.data());
                for (int j = 0; j < N; j++) prod[j] = mulmod(t[j], key.g[l], q);
                addto(row(K, k1), prod.data());
                for (int c = 0; c < K1; c++) {
                    const size_t so = (static_cast<size_t>(l) * K1 + c) * N;
                    if (schoolbook) {
                        schoolbook_modq(q, N, d.data(), &key.s2[so], prod.data());
                    } else {
                        for (int j = 0; j < N; j++) prod[j] = mulmod(t[j], key.s2ntt[so + j], q);
                    }
                    addto(row(k1, c), prod.data());
                }
            }
        }
        // adding b: c[k1].a[k2] += cPrime.b
        for (int k1 = 0; k1 < K; k1++)
            for (int k2 = 0; k2 < K; k2++) addto(row(k1, k2), row(K, K));
    }
}

// externalProductTrgswMPNttInPlace (trgsw.cpp:774-776 -> 712-767) with the
// derived rows: acc := rows (x) acc, level digits of each accumulator row.
void lazy_external_product(const or_params& p, const Ntt& ntt, const std::vector<uint64_t>& rows, int64_t* acc, bool schoolbook) {
    const int N = p.N, K1 = p.k + 1, L = p.level, w = p.torus_bits;
    const uint64_t q = p.q_ntt;
    std::vector<int64_t> dig(static_cast<size_t>(K1) * L * N);
    for (int r = 0; r < K1; r++) digits_row(acc + static_cast<size_t>(r) * N, N, L, p.radix_bits, w, &dig[static_cast<size_t>(r) * L * N]);
    std::vector<uint64_t> sum(static_cast<size_t>(K1) * N, 0), t(N), prod(N);
    for (int lvl = 0; lvl < L; lvl++) {
        for (int r = 0; r < K1; r++) {
            const int64_t* d = &dig[(static_cast<size_t>(r) * L + lvl) * N];
            if (!schoolbook) {
                for (int j = 0; j < N; j++) t[j] = lift(d[j], q);
                ntt.forward(t.data());
            }
            for (int c = 0; c < K1; c++) {
                const uint64_t* kr = &rows[((static_cast<size_t>(lvl) * K1 + r) * K1 + c) * N];
                if (schoolbook) schoolbook_modq(q, N, d, kr, prod.data());
                else for (int j = 0; j < N; j++) prod[j] = mulmod(t[j], kr[j], q);
                uint64_t* s = &sum[static_cast<size_t>(c) * N];
                for (int j = 0; j < N; j++) s[j] = (s[j] + prod[j]) % q;
            }
        }
    }
    for (int c = 0; c < K1; c++) {
        uint64_t* s = &sum[static_cast<size_t>(c) * N];
        if (!schoolbook) ntt.inverse(s);
        for (int j = 0; j < N; j++) acc[static_cast<size_t>(c) * N + j] = intt_exit(s[j], q, w);
    }
}

// blindRotateLazyPipeNtt (blind_rotate.cpp:822-902), binary keys, as one
// sequential loop (the reference's pipeline only overlaps the next
// component's rotation and scheme switch with this one's product):
//   acc = X^{-b} (R(v) + (X^{a_0} - 1) bskFirst)
//   for i = 1 .. n-1, a_i != 0:  acc := RGSW_i (x) acc, RGSW_i the rows of
//   component i-1 rotated by a_i.
// abar: the rescaled input [n+1] (a_0 .. a_{n-1}, b).
void blind_rotate_lazy(const LazyKey& key, const int32_t* abar, const int64_t* v, int64_t* acc, bool schoolbook) {
    const or_params& p = key.p;
    const int N = p.N, K = p.k, K1 = p.k + 1, w = p.torus_bits;
    std::vector<int64_t> tmp(static_cast<size_t>(K1) * N);
    for (int r = 0; r < K1; r++) rotate_minus_one(w, N, abar[0], &key.first[static_cast<size_t>(r) * N], &tmp[static_cast<size_t>(r) * N]);
    for (int j = 0; j < N; j++) tmp[static_cast<size_t>(K) * N + j] = sext(tmp[static_cast<size_t>(K) * N + j] + v[j], w);
    for (int r = 0; r < K1; r++) rotate(w, N, -abar[p.n], &tmp[static_cast<size_t>(r) * N], acc + static_cast<size_t>(r) * N);
    std::vector<uint64_t> rows;
    for (int i = 1; i < p.n; i++) {
        if (abar[i] == 0) continue;
        lazy_rows(key, i - 1, abar[i], rows, schoolbook);
        lazy_external_product(p, key.ntt, rows, acc, schoolbook);
    }
}

// ---------------------------------------------------------------- lazy pipe, Alt
// BootstrappingKeyMPLazyPipeAlt (bootstrapping.h:284-306) as
// genBootstrappingKeyMPLazyPipeAlt (bootstrapping.cpp:322-339) fills it: for
// every s_i (i < n) the half RGSW rows cPrime[lvl] in torus form (TrgswMP,
// lvl < R = lApprox), and the TRLEV of s^2 with D = param.l levels.
//   rows [n][R][k+1][N]   torus: bskPrime[i].cPrime[lvl] (a[0..k-1], b)
//   s2 [D][k+1][N]        s2Dft as coefficient residues mod q
struct AltKey {
    or_params p;
    int D;
    Ntt ntt;
    std::vector<int64_t> rows;
    std::vector<uint64_t> s2, s2ntt, g;
    AltKey(const or_params& p_, int digit_levels, const int64_t* rows_, const uint64_t* s2_)
        : p(p_), D(digit_levels), ntt(p_.q_ntt, p_.N) {
        const size_t N = p.N, K1 = p.k + 1, R = p.level;
        rows.assign(rows_, rows_ + static_cast<size_t>(p.n) * R * K1 * N);
        s2.assign(s2_, s2_ + D * K1 * N);
        s2ntt = s2;
        for (size_t o = 0; o < s2ntt.size(); o += N) ntt.forward(&s2ntt[o]);
        for (int l = 0; l < D; l++) g.push_back(lift(sext(int64_t{1} << (p.torus_bits - (l + 1) * p.radix_bits), p.torus_bits), p.q_ntt));
    }
};

// The per-ciphertext rows of component i rotated by a (blindRotateLazyPipeAltNttSweep's
// stageA and stageB, blind_rotate.cpp:1035-1087): per level, rotateTrlweMinusOneBPlusOne
// (trlwe.cpp:291-297) of cPrime[lvl] with one = 2^(w - (lvl+1) b); the rotated
// mask decomposed into D digits (gadgetDecomposeTrlweA, trlwe.cpp:200-210);
//   cPrime.a[k1] = sum_l D_l g_l,  c[k1].a[k2] = sum_l D_l s2[l].a[k2] + cPrime.b,
//   c[k1].b = sum_l D_l s2[l].b,  cPrime.b = the rotated body itself (mod q).
// Layout and routes as lazy_rows.
void alt_rows(const AltKey& key, int comp, int a, std::vector<uint64_t>& rows, bool schoolbook) {
    const or_params& p = key.p;
    const int N = p.N, K = p.k, K1 = p.k + 1, R = p.level, D = key.D, w = p.torus_bits;
    const uint64_t q = p.q_ntt;
    rows.assign(static_cast<size_t>(R) * K1 * K1 * N, 0);
    std::vector<int64_t> rot(N), dig(static_cast<size_t>(D) * N);
    std::vector<uint64_t> t(N), prod(N);
    auto addto = [&](uint64_t* dst, const uint64_t* src) {
        for (int j = 0; j < N; j++) dst[j] = (dst[j] + src[j]) % q;
    };
    for (int lvl = 0; lvl < R; lvl++) {
        uint64_t* Rw = &rows[static_cast<size_t>(lvl) * K1 * K1 * N];
        auto row = [&](int r, int c) { return Rw + (static_cast<size_t>(r) * K1 + c) * N; };
        const int64_t* cp = &key.rows[((static_cast<size_t>(comp) * R + lvl) * K1) * N];
        // the rotated body plus the gadget: cPrime.b
        rotate_minus_one(w, N, a, cp + static_cast<size_t>(K) * N, rot.data());
        rot[0] = sext(rot[0] + (int64_t{1} << (w - (lvl + 1) * p.radix_bits)), w);
        for (int j = 0; j < N; j++) t[j] = lift(rot[j], q);
        if (!schoolbook) key.ntt.forward(t.data());
        std::memcpy(row(K, K), t.data(), sizeof(uint64_t) * N);
        for (int k1 = 0; k1 < K; k1++) {
            rotate_minus_one(w, N, a, cp + static_cast<size_t>(k1) * N, rot.data());
            digits_row(rot.data(), N, D, p.radix_bits, w, dig.data());
            for (int l = 0; l < D; l++) {
                const int64_t* d = &dig[static_cast<size_t>(l) * N];
                for (int j = 0; j < N; j++) t[j] = lift(d[j], q);
                if (!schoolbook) key.ntt.forward(t.data());
                for (int j = 0; j < N; j++) prod[j] = mulmod(t[j], key.g[l], q);
                addto(row(K, k1), prod.data());
                for (int c = 0; c < K1; c++) {
                    const size_t so = (static_cast<size_t>(l) * K1 + c) * N;
                    if (schoolbook) schoolbook_modq(q, N, d, &key.s2[so], prod.data());
                    else for (int j = 0; j < N; j++) prod[j] = mulmod(t[j], key.s2ntt[so + j], q);
                    addto(row(k1, c), prod.data());
                }
            }
        }
        for (int k1 = 0; k1 < K; k1++)
            for (int k2 = 0; k2 < K; k2++) addto(row(k1, k2), row(K, K));
    }
}

// deriveFirstComponentNtt (trgsw.cpp:929-955) with prepareGdV (trgsw.cpp:911-927):
// out = sum_lvl GD(v)_lvl * cPrime_0[lvl] (mod q, R = lApprox digits of v),
// the TRLWE of v s_0.
void alt_first(const AltKey& key, const int64_t* v, int64_t* out, bool schoolbook) {
    const or_params& p = key.p;
    const int N = p.N, K1 = p.k + 1, R = p.level, w = p.torus_bits;
    const uint64_t q = p.q_ntt;
    std::vector<int64_t> gd(static_cast<size_t>(R) * N);
    digits_row(v, N, R, p.radix_bits, w, gd.data());
    std::vector<uint64_t> acc(static_cast<size_t>(K1) * N, 0), t(N), r(N), prod(N);
    for (int lvl = 0; lvl < R; lvl++) {
        const int64_t* d = &gd[static_cast<size_t>(lvl) * N];
        for (int j = 0; j < N; j++) t[j] = lift(d[j], q);
        if (!schoolbook) key.ntt.forward(t.data());
        for (int c = 0; c < K1; c++) {
            const int64_t* cp = &key.rows[(static_cast<size_t>(lvl) * K1 + c) * N];  // component 0
            for (int j = 0; j < N; j++) r[j] = lift(cp[j], q);
            if (schoolbook) {
                schoolbook_modq(q, N, d, r.data(), prod.data());
            } else {
                key.ntt.forward(r.data());
                for (int j = 0; j < N; j++) prod[j] = mulmod(t[j], r[j], q);
            }
            for (int j = 0; j < N; j++) acc[static_cast<size_t>(c) * N + j] = (acc[static_cast<size_t>(c) * N + j] + prod[j]) % q;
        }
    }
    for (int c = 0; c < K1; c++) {
        uint64_t* a = &acc[static_cast<size_t>(c) * N];
        if (!schoolbook) key.ntt.inverse(a);
        for (int j = 0; j < N; j++) out[static_cast<size_t>(c) * N + j] = intt_exit(a[j], q, w);
    }
}

// blindRotateLazyPipeAltNtt (blind_rotate.cpp:996-1000 -> the sweep, 1002-1136)
// as one sequential loop: acc = X^{-b} (R(v) + (X^{a_0} - 1) derived), then for
// i = 1 .. n-1 with a_i != 0, acc := rows_i (x) acc.
void blind_rotate_alt(const AltKey& key, const int32_t* abar, const int64_t* v, int64_t* acc, bool schoolbook) {
    const or_params& p = key.p;
    const int N = p.N, K = p.k, K1 = p.k + 1, w = p.torus_bits;
    std::vector<int64_t> first(static_cast<size_t>(K1) * N), tmp(static_cast<size_t>(K1) * N);
    alt_first(key, v, first.data(), schoolbook);
    for (int r = 0; r < K1; r++) rotate_minus_one(w, N, abar[0], &first[static_cast<size_t>(r) * N], &tmp[static_cast<size_t>(r) * N]);
    for (int j = 0; j < N; j++) tmp[static_cast<size_t>(K) * N + j] = sext(tmp[static_cast<size_t>(K) * N + j] + v[j], w);
    for (int r = 0; r < K1; r++) rotate(w, N, -abar[p.n], &tmp[static_cast<size_t>(r) * N], acc + static_cast<size_t>(r) * N);
    std::vector<uint64_t> rows;
    for (int i = 1; i < p.n; i++) {
        if (abar[i] == 0) continue;
        alt_rows(key, i, abar[i], rows, schoolbook);
        lazy_external_product(p, key.ntt, rows, acc, schoolbook);
    }
}

void pbs_one(const Key& key, const int32_t* ksk, const int32_t* lwe, const int64_t* v,
             int64_t* acc, int32_t* ext, int32_t* out, bool schoolbook) {
    const or_params& p = key.p;
    std::vector<int32_t> scaled(p.n + 1);
    rescale(p, lwe, scaled.data());
    init_acc(p, v, scaled[p.n], acc);
    blind_rotate(key, scaled.data(), acc, schoolbook);
    extract(p, acc, ext);
    keyswitch(p, ksk, ext, out);
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int64_t or_sext(int64_t x, int32_t w) { return sext(x, w); }

int64_t or_mod_switch_from_torus(int64_t in, int64_t new_mod, int64_t torus_q_) {
    return mod_switch_from_torus(in, new_mod, torus_q_);
}

// Exhaustive check of the 2N modswitch against its closed form over [lo, hi):
// returns the number of mismatches.  Closed form: round-half-away-from-zero of
// x / scale with C truncating division, then sign extension mod 2N.
int64_t or_modswitch_closed_form_check(int32_t log2_2N, int64_t lo, int64_t hi) {
    const int64_t m = int64_t{1} << log2_2N;
    const int64_t scale = (int64_t{1} << 32) / m;
    int64_t bad = 0;
    for (int64_t u = lo; u < hi; u++) {
        const int64_t x = static_cast<int32_t>(static_cast<uint32_t>(u));
        const int64_t a = x >= 0 ? x : -x;
        const int64_t r = (a + scale / 2) / scale;        // |x| rounded half up
        const int64_t s = x >= 0 ? r : -r;
        const int64_t want = sext(s, log2_2N);
        if (mod_switch_from_torus(x, m, int64_t{1} << 32) != want) bad++;
    }
    return bad;
}

void or_rotate(int32_t w, int32_t N, int32_t a, const int64_t* in, int64_t* out) { rotate(w, N, a, in, out); }
void or_rotate_minus_one(int32_t w, int32_t N, int32_t a, const int64_t* in, int64_t* out) { rotate_minus_one(w, N, a, in, out); }

void or_digits_scalar(int64_t in, int32_t l, int32_t b, int32_t w, int64_t* out) { digits_scalar(in, l, b, w, out); }
void or_digits_row(const int64_t* in, int32_t N, int32_t l, int32_t b, int32_t w, int64_t* out) { digits_row(in, N, l, b, w, out); }

// Compare the two digit forms on the values x = lo + i*stride (i < count, low w
// bits), threaded.  Returns the number of mismatching values.
int64_t or_digits_cross_check(int32_t l, int32_t b, int32_t w, uint64_t lo, uint64_t stride, uint64_t count, int32_t threads) {
    std::atomic<int64_t> bad{0};
    const uint64_t chunk = 4096;
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        std::vector<int64_t> in(chunk), rowOut(static_cast<size_t>(l) * chunk), one(l);
        for (;;) {
            const uint64_t s = next.fetch_add(chunk);
            if (s >= count) break;
            const uint64_t e = std::min(count, s + chunk);
            const int n = static_cast<int>(e - s);
            for (int i = 0; i < n; i++) in[i] = sext(static_cast<int64_t>(lo + (s + i) * stride), w);
            digits_row(in.data(), n, l, b, w, rowOut.data());
            int64_t local = 0;
            for (int i = 0; i < n; i++) {
                digits_scalar(in[i], l, b, w, one.data());
                for (int lv = 0; lv < l; lv++) if (one[lv] != rowOut[static_cast<size_t>(lv) * n + i]) { local++; break; }
            }
            bad += local;
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < std::max(1, threads); t++) ts.emplace_back(work);
    for (auto& t : ts) t.join();
    return bad.load();
}

// Hash of the oracle's scalar-form digits of x = lo + i*stride (i < count),
// chunked and combined exactly as oracle/ref_digits.cpp hashes the reference's
// decomposeRow output, so equal hashes mean equal digits on every input.
uint64_t or_digits_hash(int32_t l, int32_t b, int32_t w, uint64_t lo, uint64_t stride, uint64_t count, int32_t threads) {
    const uint64_t chunk = 4096;
    const uint64_t nChunks = (count + chunk - 1) / chunk;
    std::vector<uint64_t> part(nChunks, 0);
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        std::vector<int64_t> d(l);
        std::vector<int64_t> all(static_cast<size_t>(l) * chunk);
        for (;;) {
            const uint64_t c = next.fetch_add(1);
            if (c >= nChunks) break;
            const uint64_t s = c * chunk;
            const int n = static_cast<int>(std::min(chunk, count - s));
            uint64_t h = 0x9E3779B97F4A7C15ULL ^ c;
            for (int i = 0; i < n; i++) {
                digits_scalar(sext(static_cast<int64_t>(lo + (s + i) * stride), w), l, b, w, d.data());
                for (int j = 0; j < l; j++) h ^= static_cast<uint64_t>(d[j]) + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
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

// Negacyclic product mod q via the oracle NTT (method 0) or schoolbook (1).
void or_negacyclic_mul_modq(uint64_t q, int32_t N, const int64_t* a, const uint64_t* b, uint64_t* out, int32_t method) {
    if (method == 1) { schoolbook_modq(q, N, a, b, out); return; }
    Ntt ntt(q, N);
    std::vector<uint64_t> x(N), y(b, b + N);
    for (int i = 0; i < N; i++) x[i] = lift(a[i], q);
    ntt.forward(x.data());
    ntt.forward(y.data());
    for (int i = 0; i < N; i++) x[i] = mulmod(x[i], y[i], q);
    ntt.inverse(x.data());
    std::memcpy(out, x.data(), sizeof(uint64_t) * N);
}

// Forward/inverse round trip of the oracle NTT (returns 1 on identity).
int32_t or_ntt_roundtrip(uint64_t q, int32_t N, const uint64_t* a) {
    Ntt ntt(q, N);
    std::vector<uint64_t> x(a, a + N);
    ntt.forward(x.data());
    ntt.inverse(x.data());
    return std::equal(x.begin(), x.end(), a) ? 1 : 0;
}

int64_t or_intt_exit(uint64_t x, uint64_t q, int32_t w) { return intt_exit(x, q, w); }

void* or_key_create(const or_params* p, const uint64_t* bsk_coeff) {
    try { return new Key(*p, bsk_coeff); } catch (...) { return nullptr; }
}
void or_key_free(void* k) { delete static_cast<Key*>(k); }

void or_external_product(void* key, int32_t i, const int64_t* in, int64_t* out, int32_t schoolbook) {
    external_product(*static_cast<Key*>(key), static_cast<size_t>(i), in, out, schoolbook != 0);
}

void or_rescale(const or_params* p, const int32_t* lwe, int32_t* out) { rescale(*p, lwe, out); }
void or_init_acc(const or_params* p, const int64_t* v, int32_t bbar, int64_t* acc) { init_acc(*p, v, bbar, acc); }
void or_blind_rotate(void* key, const int32_t* abar, int64_t* acc, int32_t schoolbook) {
    blind_rotate(*static_cast<Key*>(key), abar, acc, schoolbook != 0);
}
void or_extract(const or_params* p, const int64_t* acc, int32_t* out) { extract(*p, acc, out); }
void or_keyswitch(const or_params* p, const int32_t* ksk, const int32_t* in, int32_t* out) { keyswitch(*p, ksk, in, out); }

// Keyswitch with wrapping uint32 accumulation (rule R11 cross-check): must equal
// or_keyswitch for qLwe = 2^32.
void or_keyswitch_u32(const or_params* p, const int32_t* ksk, const int32_t* in, int32_t* out) {
    const int nOut = p->n, nIn = p->k * p->N, t = p->ks_level;
    std::vector<uint32_t> acc(nOut + 1, 0);
    std::vector<int64_t> d(t);
    for (int i = 0; i < nIn; i++) {
        digits_scalar(in[i], t, p->ks_radix_bits, p->ks_width_bits, d.data());
        for (int j = 0; j < t; j++) {
            const uint32_t c = static_cast<uint32_t>(d[j]);
            const int32_t* row = ksk + (static_cast<size_t>(i) * t + j) * (nOut + 1);
            for (int kk = 0; kk <= nOut; kk++) acc[kk] += c * static_cast<uint32_t>(row[kk]);
        }
    }
    for (int kk = 0; kk < nOut; kk++) out[kk] = static_cast<int32_t>(0u - acc[kk]);
    out[nOut] = static_cast<int32_t>(static_cast<uint32_t>(in[nIn]) - acc[nOut]);
}

// functionalBootstrappingMPNtt (bootstrapping.cpp:71-80) for one ciphertext;
// acc (k+1)N torus, ext kN+1, out n+1 (any of the three may be null).
void or_pbs(void* key, const int32_t* ksk, const int32_t* lwe, const int64_t* v,
            int64_t* acc, int32_t* ext, int32_t* out, int32_t schoolbook) {
    const Key& k = *static_cast<Key*>(key);
    const or_params& p = k.p;
    std::vector<int64_t> a(static_cast<size_t>(p.k + 1) * p.N);
    std::vector<int32_t> e(static_cast<size_t>(p.k) * p.N + 1), o(p.n + 1);
    pbs_one(k, ksk, lwe, v, a.data(), e.data(), o.data(), schoolbook != 0);
    if (acc) std::memcpy(acc, a.data(), a.size() * sizeof(int64_t));
    if (ext) std::memcpy(ext, e.data(), e.size() * sizeof(int32_t));
    if (out) std::memcpy(out, o.data(), o.size() * sizeof(int32_t));
}

// Many independent ciphertexts, one per thread at a time (the CPU baseline's
// throughput form).  lwe/out are [count][n+1]; v is one test polynomial.
// Many ciphertexts, one per thread at a time; acc_out ([count][(k+1)N]) and
// ext_out ([count][kN+1]) are optional.
void or_pbs_batch_all(void* key, const int32_t* ksk, int32_t count, const int32_t* lwe, const int64_t* v,
                      int32_t* out, int64_t* acc_out, int32_t* ext_out, int32_t threads) {
    const Key& k = *static_cast<Key*>(key);
    const or_params& p = k.p;
    const size_t accLen = static_cast<size_t>(p.k + 1) * p.N, extLen = static_cast<size_t>(p.k) * p.N + 1;
    std::atomic<int> next{0};
    auto work = [&]() {
        std::vector<int64_t> a(accLen);
        std::vector<int32_t> e(extLen);
        for (;;) {
            const int c = next.fetch_add(1);
            if (c >= count) break;
            pbs_one(k, ksk, lwe + static_cast<size_t>(c) * (p.n + 1), v, a.data(), e.data(),
                    out + static_cast<size_t>(c) * (p.n + 1), false);
            if (acc_out) std::memcpy(acc_out + c * accLen, a.data(), accLen * sizeof(int64_t));
            if (ext_out) std::memcpy(ext_out + c * extLen, e.data(), extLen * sizeof(int32_t));
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < std::max(1, threads); t++) ts.emplace_back(work);
    for (auto& t : ts) t.join();
}

// functionalBootstrappingMPNtt for one ciphertext with the external product's
// gadget levels on a pool of param.level workers (the reference's latency form).
void or_pbs_level_parallel(void* key, const int32_t* ksk, const int32_t* lwe, const int64_t* v, int32_t* out, int32_t reps,
                           double* seconds) {
    const Key& k = *static_cast<Key*>(key);
    const or_params& p = k.p;
    LevelPool pool(p.level);
    std::vector<int64_t> acc(static_cast<size_t>(p.k + 1) * p.N);
    std::vector<int32_t> scaled(p.n + 1), ext(static_cast<size_t>(p.k) * p.N + 1);
    for (int r = 0; r < reps; r++) {
        const auto t0 = std::chrono::steady_clock::now();
        rescale(p, lwe, scaled.data());
        init_acc(p, v, scaled[p.n], acc.data());
        blind_rotate_level_parallel(k, scaled.data(), acc.data(), pool);
        extract(p, acc.data(), ext.data());
        keyswitch(p, ksk, ext.data(), out);
        if (seconds) seconds[r] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
}

// The lazy-pipe key (layouts at LazyKey) and its bootstrapping.
void* or_lazy_key_new(const or_params* p, int32_t digit_levels, const int64_t* first, const int32_t* da, const int32_t* db,
                      const uint64_t* s2) {
    try {
        return new LazyKey(*p, digit_levels, first, da, db, s2);
    } catch (...) {
        return nullptr;
    }
}
void or_lazy_key_free(void* k) { delete static_cast<LazyKey*>(k); }
// The per-ciphertext rows of component `comp` rotated by a, in coefficient
// form [L][k+1][k+1][N] (the NTT route's rows inverse-transformed).
void or_lazy_rows(void* key, int32_t comp, int32_t a, uint64_t* out, int32_t schoolbook) {
    const LazyKey& k = *static_cast<LazyKey*>(key);
    std::vector<uint64_t> rows;
    lazy_rows(k, comp, a, rows, schoolbook != 0);
    if (!schoolbook)
        for (size_t o = 0; o < rows.size(); o += k.p.N) k.ntt.inverse(&rows[o]);
    std::memcpy(out, rows.data(), sizeof(uint64_t) * rows.size());
}
void or_blind_rotate_lazy(void* key, const int32_t* abar, const int64_t* v, int64_t* acc, int32_t schoolbook) {
    blind_rotate_lazy(*static_cast<LazyKey*>(key), abar, v, acc, schoolbook != 0);
}
// blindRotateLazyPipeNtt, extract and keyswitch of each ciphertext ([count][n+1]),
// threaded over ciphertexts; acc_out, ext_out optional.
void or_pbs_lazy_batch(void* key, const int32_t* ksk, int32_t count, const int32_t* lwe, const int64_t* v, int64_t* acc_out,
                       int32_t* ext_out, int32_t* out, int32_t threads) {
    const LazyKey& k = *static_cast<LazyKey*>(key);
    const or_params& p = k.p;
    const size_t accN = static_cast<size_t>(p.k + 1) * p.N, extN = static_cast<size_t>(p.k) * p.N + 1;
    std::atomic<int32_t> next{0};
    auto work = [&] {
        std::vector<int32_t> scaled(p.n + 1), ext(extN);
        std::vector<int64_t> acc(accN);
        for (int32_t c; (c = next.fetch_add(1)) < count;) {
            rescale(p, lwe + static_cast<size_t>(c) * (p.n + 1), scaled.data());
            blind_rotate_lazy(k, scaled.data(), v, acc.data(), false);
            extract(p, acc.data(), ext.data());
            if (acc_out) std::memcpy(acc_out + c * accN, acc.data(), sizeof(int64_t) * accN);
            if (ext_out) std::memcpy(ext_out + c * extN, ext.data(), sizeof(int32_t) * extN);
            if (out) keyswitch(p, ksk, ext.data(), out + static_cast<size_t>(c) * (p.n + 1));
        }
    };
    std::vector<std::thread> ts;
    for (int t = 1; t < std::max(1, threads); t++) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
}

// The Alt lazy-pipe key (layouts at AltKey) and its bootstrapping.
void* or_alt_key_new(const or_params* p, int32_t digit_levels, const int64_t* rows, const uint64_t* s2) {
    try {
        return new AltKey(*p, digit_levels, rows, s2);
    } catch (...) {
        return nullptr;
    }
}
void or_alt_key_free(void* k) { delete static_cast<AltKey*>(k); }
void or_alt_rows(void* key, int32_t comp, int32_t a, uint64_t* out, int32_t schoolbook) {
    const AltKey& k = *static_cast<AltKey*>(key);
    std::vector<uint64_t> rows;
    alt_rows(k, comp, a, rows, schoolbook != 0);
    if (!schoolbook)
        for (size_t o = 0; o < rows.size(); o += k.p.N) k.ntt.inverse(&rows[o]);
    std::memcpy(out, rows.data(), sizeof(uint64_t) * rows.size());
}
void or_alt_first(void* key, const int64_t* v, int64_t* out, int32_t schoolbook) {
    alt_first(*static_cast<AltKey*>(key), v, out, schoolbook != 0);
}
void or_blind_rotate_alt(void* key, const int32_t* abar, const int64_t* v, int64_t* acc, int32_t schoolbook) {
    blind_rotate_alt(*static_cast<AltKey*>(key), abar, v, acc, schoolbook != 0);
}
void or_pbs_alt_batch(void* key, const int32_t* ksk, int32_t count, const int32_t* lwe, const int64_t* v, int64_t* acc_out,
                      int32_t* ext_out, int32_t* out, int32_t threads) {
    const AltKey& k = *static_cast<AltKey*>(key);
    const or_params& p = k.p;
    const size_t accN = static_cast<size_t>(p.k + 1) * p.N, extN = static_cast<size_t>(p.k) * p.N + 1;
    std::atomic<int32_t> next{0};
    auto work = [&] {
        std::vector<int32_t> scaled(p.n + 1), ext(extN);
        std::vector<int64_t> acc(accN);
        for (int32_t c; (c = next.fetch_add(1)) < count;) {
            rescale(p, lwe + static_cast<size_t>(c) * (p.n + 1), scaled.data());
            blind_rotate_alt(k, scaled.data(), v, acc.data(), false);
            extract(p, acc.data(), ext.data());
            if (acc_out) std::memcpy(acc_out + c * accN, acc.data(), sizeof(int64_t) * accN);
            if (ext_out) std::memcpy(ext_out + c * extN, ext.data(), sizeof(int32_t) * extN);
            if (out) keyswitch(p, ksk, ext.data(), out + static_cast<size_t>(c) * (p.n + 1));
        }
    };
    std::vector<std::thread> ts;
    for (int t = 1; t < std::max(1, threads); t++) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
}

void or_pbs_batch(void* key, const int32_t* ksk, int32_t count, const int32_t* lwe, const int64_t* v,
                  int32_t* out, int32_t threads) {
    or_pbs_batch_all(key, ksk, count, lwe, v, out, nullptr, nullptr, threads);
}

// The full key of a WWL+24 succinct key: every (step, level) through
// derive_succinct_row by the NTT route.  half [n][level][k+1][N], s2
// [s2_level][k+1][N] -> full [n][level][k+1][k+1][N], all residues mod q.
// blindRotateWWL24Ntt (blind_rotate.cpp:1383-1412) derives the same rows at every
// step and then runs the MP21 external product on them, so blind rotation with
// the full key reproduces it.
void or_derive_succinct(const or_params* p, int32_t s2_level, const uint64_t* half, const uint64_t* s2, uint64_t* full, int32_t threads) {
    const Ntt ntt(p->q_ntt, p->N);
    const int K1 = p->k + 1, N = p->N;
    const std::vector<uint64_t> s2ntt = s2_transformed(ntt, s2_level, K1, N, s2);
    const int rows = p->n * p->level;
    std::atomic<int> next{0};
    auto work = [&]() {
        for (;;) {
            const int r = next.fetch_add(1);
            if (r >= rows) break;
            derive_succinct_row(*p, ntt, s2ntt, s2_level, half + static_cast<size_t>(r) * K1 * N, s2,
                                full + static_cast<size_t>(r) * K1 * K1 * N, false);
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < std::max(1, threads); t++) ts.emplace_back(work);
    for (auto& t : ts) t.join();
}

// One (step, level) of the derivation, by the NTT route (0) or schoolbook (1).
void or_derive_succinct_row(const or_params* p, int32_t s2_level, const uint64_t* cp, const uint64_t* s2, uint64_t* out, int32_t schoolbook) {
    const Ntt ntt(p->q_ntt, p->N);
    const std::vector<uint64_t> s2ntt = schoolbook ? std::vector<uint64_t>{} : s2_transformed(ntt, s2_level, p->k + 1, p->N, s2);
    derive_succinct_row(*p, ntt, s2ntt, s2_level, cp, s2, out, schoolbook != 0);
}

// Decryption helpers for semantic gates.
// LWE phase b - <a, s> mod 2^32 (symDecTlwe*, tlwe.cpp:33-53, before rounding).
int32_t or_lwe_phase(const int32_t* ct, const int32_t* s, int32_t n) {
    int64_t tmp = 0;
    for (int i = 0; i < n; i++) if (s[i]) tmp += static_cast<int64_t>(ct[i]) * s[i];
    return static_cast<int32_t>(static_cast<uint32_t>(ct[n]) - static_cast<uint32_t>(tmp));
}

// GLWE phase b - sum_i a_i * z_i mod 2^w (symDecTrlweWoRounding), binary key z [k][N].
void or_glwe_phase(int32_t w, int32_t k, int32_t N, const int64_t* acc, const int32_t* z, int64_t* out) {
    std::vector<i128> t(N, 0);
    for (int i = 0; i < k; i++) {
        const int64_t* a = acc + static_cast<size_t>(i) * N;
        const int32_t* zi = z + static_cast<size_t>(i) * N;
        for (int u = 0; u < N; u++) {
            if (!zi[u]) continue;
            for (int j = 0; j < N; j++) {
                const int d = u + j;
                if (d < N) t[d] += a[j]; else t[d - N] -= a[j];
            }
        }
    }
    const int64_t* b = acc + static_cast<size_t>(k) * N;
    for (int j = 0; j < N; j++) out[j] = sext(static_cast<int64_t>(static_cast<uint64_t>(b[j]) - static_cast<uint64_t>(static_cast<int64_t>(t[j]))), w);
}

// GLWE phase mod q of a coefficient-domain key row: (b - sum a_i z_i) mod q,
// centred to (-q/2, q/2).  rows: a[k][N] then b[N], residues in [0,q).
void or_glwe_phase_modq(uint64_t q, int32_t k, int32_t N, const uint64_t* row, const int32_t* z, int64_t* out) {
    std::vector<i128> t(N, 0);
    for (int i = 0; i < k; i++) {
        const uint64_t* a = row + static_cast<size_t>(i) * N;
        const int32_t* zi = z + static_cast<size_t>(i) * N;
        for (int u = 0; u < N; u++) {
            if (!zi[u]) continue;
            for (int j = 0; j < N; j++) {
                const int d = u + j;
                if (d < N) t[d] += a[j]; else t[d - N] -= a[j];
            }
        }
    }
    const uint64_t* b = row + static_cast<size_t>(k) * N;
    for (int j = 0; j < N; j++) {
        i128 r = (static_cast<i128>(b[j]) - t[j]) % static_cast<i128>(q);
        if (r < 0) r += q;
        int64_t c = static_cast<int64_t>(r);
        if (static_cast<uint64_t>(c) > q / 2) c -= static_cast<int64_t>(q);
        out[j] = c;
    }
}

}  // extern "C"
