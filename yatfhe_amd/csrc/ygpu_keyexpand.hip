// ygpu_keyexpand.hip -- device-side preparation of bootstrapping keys.
//
// A separate translation unit from the blind-rotation kernels: it instantiates
// the same inline transform code, and keeping it apart leaves the register
// allocation of k_blind_rotate untouched.
#include "ygpu_kernels.h"

namespace ygpu {

// ============================================================ key expansion
// One team of T = N/16 lanes per (polynomial, transform prime): unpack the
// residues, lift each to its centred representative mod q, reduce it to the
// centred range mod p, transform forward (natural -> last-pass order), scale by
// N^-1 and centre.  Transform bound: inputs |x| <= p/2 grow by at most
// p(0.5 + 1.5|x| 2^-53) per stage, i.e. below 10.3p after 11 stages (< 16p).
template <int LOGN>
__global__ void __launch_bounds__((1 << LOGN) / 16) k_bsk_expand(BskExpandArgs a) {
    constexpr int N = 1 << LOGN, T = N / 16, TW = T / 64;
    __shared__ double scr[Scratch<LOGN>::DOUBLES];
    const int pi = static_cast<int>(blockIdx.x) / a.np;
    const int np = static_cast<int>(blockIdx.x) % a.np;
    const int tl = static_cast<int>(threadIdx.x);
    const PrimeF64 pr = np ? a.prime[1] : a.prime[0];
    const int64_t p = static_cast<int64_t>(np ? a.p[1] : a.p[0]);
    const double ninv = np ? a.ninv[1] : a.ninv[0];
    const uint64_t mask = (a.bits >= 64) ? ~0ULL : ((1ULL << a.bits) - 1);
    const uint64_t qhalf = (a.q - 1) / 2;
    double x[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint64_t g = static_cast<uint64_t>(pi) * N + tl + T * r;
        const uint64_t bit = g * static_cast<uint64_t>(a.bits);
        const uint64_t w = bit >> 6;
        const int off = static_cast<int>(bit & 63);
        uint64_t v = a.packed[w] >> off;
        if (off + a.bits > 64) v |= a.packed[w + 1] << (64 - off);
        v &= mask;
        int64_t c = v > qhalf ? static_cast<int64_t>(v) - static_cast<int64_t>(a.q) : static_cast<int64_t>(v);
        c %= p;
        if (c > p / 2) c -= p;
        if (c < -(p / 2)) c += p;
        x[r] = static_cast<double>(c);
    }
    ntt_forward<LOGN, TW>(x, scr, tl, a.tw + static_cast<size_t>(np) * 2 * N, pr);
    double* out = a.out + (static_cast<size_t>(pi) * a.np + np) * N;
#pragma unroll
    for (int r = 0; r < 16; r++) out[r * T + tl] = centre(mulmod(x[r], ninv, pr), pr);
}

hipError_t launch_bsk_expand(int logn, const BskExpandArgs& a, hipStream_t s) {
    const unsigned blocks = static_cast<unsigned>(a.polys) * static_cast<unsigned>(a.np);
    if (logn == 10) hipLaunchKernelGGL(k_bsk_expand<10>, dim3(blocks), dim3(64), 0, s, a);
    else if (logn == 11) hipLaunchKernelGGL(k_bsk_expand<11>, dim3(blocks), dim3(128), 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// ===================================================== succinct-key derivation
// Residue number g of a packed array, as its centred representative mod q.
__device__ __forceinline__ int64_t unpack_centred(const uint64_t* packed, uint64_t g, int bits, uint64_t q) {
    const uint64_t mask = (bits >= 64) ? ~0ULL : ((1ULL << bits) - 1);
    const uint64_t bit = g * static_cast<uint64_t>(bits);
    const uint64_t w = bit >> 6;
    const int off = static_cast<int>(bit & 63);
    uint64_t v = packed[w] >> off;
    if (off + bits > 64) v |= packed[w + 1] << (64 - off);
    v &= mask;
    return v > (q - 1) / 2 ? static_cast<int64_t>(v) - static_cast<int64_t>(q) : static_cast<int64_t>(v);
}

// Digit of level `lvl` among the l zero-mean balanced digits of a TBITS-bit
// torus value (gadget_decomposition.h:75-95): the carry chain runs from the
// least significant level (l - 1) up to lvl.
template <int TBITS>
__device__ __forceinline__ int digit_at(int64_t x, int l, int rb, int lvl) {
    const uint64_t mask = (TBITS >= 64) ? ~0ULL : ((1ULL << TBITS) - 1);
    const uint64_t u = static_cast<uint64_t>(x) & mask;
    const bool neg = (u >> (TBITS - 1)) & 1;
    const int dropped = TBITS - l * rb;
    const uint64_t bias = dropped > 0 ? (1ULL << (dropped - 1)) : 0;
    const uint64_t m = (neg ? ((0ULL - u) & mask) : u) + bias;
    const int B = 1 << rb, halfB = B >> 1;
    int carry = 0, bal = 0;
    for (int v = l - 1; v >= lvl; v--) {
        const int shift = TBITS - (v + 1) * rb;
        const int digit = static_cast<int>((m >> shift) & static_cast<uint64_t>(B - 1)) + carry;
        const int tie = v > 0 ? static_cast<int>((m >> (shift + 2 * rb - 1)) & 1) : 0;
        carry = (digit > halfB) | ((digit == halfB) & tie);
        bal = carry ? digit - B : digit;
    }
    return neg ? -bal : bal;
}

// Engine form of one polynomial of centred residues mod q: reduce to the
// centred range mod p, transform forward, scale by N^-1, centre (x on return).
template <int LOGN, int TW>
__device__ __forceinline__ void engine_form(const int64_t (&c)[16], double (&x)[16], double* scr, int tl, const double* tw,
                                            const PrimeF64& pr, int64_t p, double ninv) {
#pragma unroll
    for (int r = 0; r < 16; r++) {
        int64_t v = c[r] % p;
        if (v > p / 2) v -= p;
        if (v < -(p / 2)) v += p;
        x[r] = static_cast<double>(v);
    }
    ntt_forward<LOGN, TW>(x, scr, tl, tw, pr);
#pragma unroll
    for (int r = 0; r < 16; r++) x[r] = centre(mulmod(x[r], ninv, pr), pr);
}

template <int LOGN>
__device__ __forceinline__ void store_poly(const double (&x)[16], double* out, int tl) {
    constexpr int T = (1 << LOGN) / 16;
#pragma unroll
    for (int r = 0; r < 16; r++) out[r * T + tl] = x[r];
}

// One workgroup per (step i, level lvl) of a WWL+24 succinct key
// (BootstrappingKeyWWL24, bootstrapping.h:116-136), one team of T = N/16 lanes
// per transform prime.  Writes the body-digit row cPrime[lvl] (copied) and the
// mask-digit row c[lvl][0] derived by switchTrlweToSecretEmbeddingNtt
// (trgsw.cpp:861-909):
//   c.a = sum_l2 D_l2(cPrime.a) * s2[l2].a + cPrime.b,   c.b = sum_l2 D_l2(cPrime.a) * s2[l2].b
// where D_l2 is digit level l2 of the torus value of cPrime.a (applyIntt's
// centring and sign extension, then decomposeRow with s2_level digits), each
// digit polynomial transformed once for both columns.  The products are exact
// negacyclic products mod q.  With one transform prime (p = q) the derived row
// is formed directly in the transform domain: the s^2 rows and cPrime.b are
// already scaled by N^-1, so sum NTT(D) s2 + cPrime.b is the engine form.  With
// two primes each team forms Y = sum D * centred(s2) exactly mod its prime
// (|Y| < s2_level N B/2 q/2 < p0 p1 / 2), inverts the transform, and the teams
// swap residues through LDS one column at a time; Garner CRT gives Y mod q, and
// the result is brought to engine form like an uploaded residue.  s2 arrives
// in engine form ([l2][col][prime][N]).
template <int LOGN, int TBITS, int NP>
__global__ void __launch_bounds__(NP * (1 << LOGN) / 16) k_bsk_derive(BskDeriveArgs a) {
    constexpr int N = 1 << LOGN, T = N / 16, TW = T / 64, K1 = 2;
    __shared__ double scratch[NP][Scratch<LOGN>::DOUBLES];
    __shared__ double ybuf[NP][NP == 2 ? N : 1];
    const int il = static_cast<int>(blockIdx.x);  // i * level + lvl
    const int j = static_cast<int>(threadIdx.x) / T;
    const int tl = static_cast<int>(threadIdx.x) % T;
    double* scr = scratch[j];
    const PrimeF64 pr = a.prime[j];
    const int64_t p = static_cast<int64_t>(a.p[j]);
    const double ninv = a.ninv[j];
    const double* tw = a.tw + static_cast<size_t>(j) * 2 * N;
    const uint64_t polyA = static_cast<uint64_t>(il) * K1;  // cPrime.a; cPrime.b follows
    double* rowBase = a.out + static_cast<size_t>(il) * K1 * K1 * NP * N;  // [row][col][prime][N]
    auto slot = [&](int row, int col) { return rowBase + ((static_cast<size_t>(row) * K1 + col) * NP + j) * N; };

    // body-digit row (row 1): cPrime unchanged; keep the torus values of cPrime.a
    int64_t c[16], t[16];
    double x[16], bx[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        c[r] = unpack_centred(a.half, polyA * N + tl + T * r, a.bits, a.q);
        t[r] = sext_t<TBITS>(c[r]);
    }
    engine_form<LOGN, TW>(c, x, scr, tl, tw, pr, p, ninv);
    store_poly<LOGN>(x, slot(1, 0), tl);
#pragma unroll
    for (int r = 0; r < 16; r++) c[r] = unpack_centred(a.half, (polyA + 1) * N + tl + T * r, a.bits, a.q);
    engine_form<LOGN, TW>(c, bx, scr, tl, tw, pr, p, ninv);
    store_poly<LOGN>(bx, slot(1, 1), tl);

    // mask-digit row (row 0): both columns from one transform per digit level
    double acc0[16], acc1[16];
#pragma unroll
    for (int r = 0; r < 16; r++) { acc0[r] = 0.0; acc1[r] = 0.0; }
    for (int l2 = 0; l2 < a.s2_level; l2++) {
#pragma unroll
        for (int r = 0; r < 16; r++) x[r] = static_cast<double>(digit_at<TBITS>(t[r], a.s2_level, a.radix_bits, l2));
        ntt_forward<LOGN, TW>(x, scr, tl, tw, pr);
        const double* s2a = a.s2 + ((static_cast<size_t>(l2) * K1 + 0) * NP + j) * N + tl;
        const double* s2b = a.s2 + ((static_cast<size_t>(l2) * K1 + 1) * NP + j) * N + tl;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            acc0[r] = centre(__dadd_rn(acc0[r], mulmod(x[r], s2a[r * T], pr)), pr);
            acc1[r] = centre(__dadd_rn(acc1[r], mulmod(x[r], s2b[r * T], pr)), pr);
        }
    }
    if constexpr (NP == 1) {
        // p = q: the engine form of the derived residues, directly
#pragma unroll
        for (int r = 0; r < 16; r++) {
            acc0[r] = centre(__dadd_rn(acc0[r], bx[r]), pr);
            acc1[r] = centre(acc1[r], pr);
        }
        store_poly<LOGN>(acc0, slot(0, 0), tl);
        store_poly<LOGN>(acc1, slot(0, 1), tl);
    } else {
        const int64_t q = static_cast<int64_t>(a.q);
        const int64_t qhalf = (q - 1) / 2;
        ntt_inverse<LOGN, TW>(acc0, scr, tl, tw + N, pr);
        ntt_inverse<LOGN, TW>(acc1, scr, tl, tw + N, pr);
        for (int col = 0; col < K1; col++) {
            // this team's residues of column col -> LDS; Y mod q from both primes
#pragma unroll
            for (int r = 0; r < 16; r++) ybuf[j][tl + T * r] = centre(col == 0 ? acc0[r] : acc1[r], pr);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 16; r++) {
                int64_t v = crt_mod_q(ybuf[0][tl + T * r], ybuf[1][tl + T * r], a.prime[1], a.crt);
                if (col == 0) {
                    v += unpack_centred(a.half, (polyA + 1) * N + tl + T * r, a.bits, a.q);
                    if (v > qhalf) v -= q;
                    if (v < -qhalf) v += q;
                }
                c[r] = v;
            }
            __syncthreads();
            engine_form<LOGN, TW>(c, x, scr, tl, tw, pr, p, ninv);
            store_poly<LOGN>(x, slot(0, col), tl);
        }
    }
}

hipError_t launch_bsk_derive(int logn, int tbits, const BskDeriveArgs& a, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>(a.n) * static_cast<unsigned>(a.level));
    const unsigned t = (1u << logn) / 16 * static_cast<unsigned>(a.np);
    if (a.np == 1) {
        if (logn == 10 && tbits == 32) hipLaunchKernelGGL((k_bsk_derive<10, 32, 1>), grid, dim3(t), 0, s, a);
        else if (logn == 11 && tbits == 32) hipLaunchKernelGGL((k_bsk_derive<11, 32, 1>), grid, dim3(t), 0, s, a);
        else if (logn == 10 && tbits == 56) hipLaunchKernelGGL((k_bsk_derive<10, 56, 1>), grid, dim3(t), 0, s, a);
        else if (logn == 11 && tbits == 56) hipLaunchKernelGGL((k_bsk_derive<11, 56, 1>), grid, dim3(t), 0, s, a);
        else return hipErrorInvalidValue;
    } else if (a.np == 2) {
        if (logn == 10 && tbits == 32) hipLaunchKernelGGL((k_bsk_derive<10, 32, 2>), grid, dim3(t), 0, s, a);
        else if (logn == 11 && tbits == 32) hipLaunchKernelGGL((k_bsk_derive<11, 32, 2>), grid, dim3(t), 0, s, a);
        else if (logn == 10 && tbits == 56) hipLaunchKernelGGL((k_bsk_derive<10, 56, 2>), grid, dim3(t), 0, s, a);
        else if (logn == 11 && tbits == 56) hipLaunchKernelGGL((k_bsk_derive<11, 56, 2>), grid, dim3(t), 0, s, a);
        else return hipErrorInvalidValue;
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ygpu
