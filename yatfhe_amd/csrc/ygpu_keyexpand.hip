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

}  // namespace ygpu
