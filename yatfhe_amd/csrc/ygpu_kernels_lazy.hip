// ygpu_kernels_lazy.hip -- the lazy-pipe key format's per-ciphertext key work.
//
// The lazy-pipe key (BootstrappingKeyMPLazyPipe, bootstrapping.h:250-282)
// stores, for every LWE key bit s_{i+1} (i < n - 1) and gadget level lvl, the
// zero-mean digits of the half RGSW row cPrime[lvl] (int8, l digit levels),
// one TRLWE bskFirst of v s_0, and the TRLEV of s^2.  blindRotateLazyPipeNtt
// (blind_rotate.cpp:822-902) rebuilds a full RGSW row set for every
// ciphertext and step: it rotates the stored digits by (X^{a_{i+1}} - 1)
// (rotateBskComponent, blind_rotate.cpp:490-507), adds the gadget's 1 on the
// diagonal digit, and scheme-switches through s^2
// (switchDecompTrlweToSecretEmbeddingNtt, trgsw.cpp:1069-1102); the
// accumulator is then replaced by the external product with those rows.
//
// Here the rows are built by k_lazy_derive, one wavefront per (ciphertext,
// step, level), into a workspace in the engine's BSK layout (transform domain
// times N^-1, last-pass order, [step][level][input row][output column][N]),
// a window of steps at a time; the batched blind-rotation kernel in its lazy
// form then runs those steps with each ciphertext reading its own rows.
// k_lazy_prepare forms the initial accumulator X^{-b} (R(v) + (X^{a_0} - 1)
// bskFirst) and the mod-switched rotation amounts of steps 1 .. n-1.
#include "ygpu_kernels.h"

namespace ygpu {

namespace {

constexpr int LZ_LOGN = 10, LZ_N = 1 << LZ_LOGN, LZ_T = LZ_N / 16, LZ_K1 = 2;
static_assert(LZ_T == 64, "one wavefront per N = 1024 transform");
constexpr int LZ_WAVES = 4;
constexpr int LZ_SCR = Scratch<LZ_LOGN>::DOUBLES;

}  // namespace

// acc = X^{-bbar} (R(v) + (X^{abar_0} - 1) bskFirst)  (blind_rotate.cpp:856-861:
// rotateTrlweMinusOne, addTorusPolynomial, rotateTrlwe), torus arithmetic mod
// 2^32; and the rotation amounts abar_1 .. abar_{n-1} (rescaleTlweToNewMod,
// tlwe.cpp:75-81).  One workgroup per ciphertext.
__global__ void __launch_bounds__(256) k_lazy_prepare(LazyArgs a) {
    constexpr int N = LZ_N, K1 = LZ_K1;
    const int ct = static_cast<int>(blockIdx.x);
    if (ct >= a.count) return;
    const int n = a.n;
    const int32_t* __restrict__ lwe = a.lwe_in + static_cast<size_t>(ct) * (n + 1);
    auto abar = [&](int i) { return a.prescaled ? lwe[i] : mod_switch_2n(lwe[i], LZ_LOGN + 1); };
    for (int i = 1 + static_cast<int>(threadIdx.x); i < n; i += 256) a.avals[static_cast<size_t>(ct) * (n - 1) + (i - 1)] = abar(i);
    if (threadIdx.x == 0) {
        a.row_tab[ct] = a.ws + static_cast<size_t>(ct) * a.ws_steps * a.level * K1 * K1 * N;
        a.row_slot[ct] = ct;
    }
    int aT1, w1, aT2, w2;
    validate_rotator(abar(0), N, aT1, w1);
    validate_rotator(-abar(n), N, aT2, w2);
    const int32_t* __restrict__ v = a.v + static_cast<size_t>(a.v_stride) * ct;
    int32_t* __restrict__ acc = a.acc + static_cast<size_t>(ct) * K1 * N;
    const int32_t* __restrict__ first = a.first + static_cast<size_t>(a.first_stride) * ct;
    for (int e = static_cast<int>(threadIdx.x); e < K1 * N; e += 256) {
        const int r = e / N, j = e % N;
        const int32_t* __restrict__ f = first + static_cast<size_t>(r) * N;
        // t_r[m] = (X^{a0} f_r)[m] - f_r[m] (+ v[m] for the body row), at m = the source of j under X^{-b}
        const int m = (j - aT2) & (N - 1);
        const int src = (m - aT1) & (N - 1);
        const uint32_t fs = static_cast<uint32_t>(f[src]);
        const uint32_t rot = ((m < aT1) != (w1 < 0)) ? 0u - fs : fs;
        uint32_t t = rot - static_cast<uint32_t>(f[m]);
        if (r == K1 - 1) t += static_cast<uint32_t>(v[m]);
        acc[e] = static_cast<int32_t>(((j < aT2) != (w2 < 0)) ? 0u - t : t);
    }
}

// One wavefront per (ciphertext, step si of the window, level lvl < R): the
// rows of level lvl of step s0 + si.  For digit level l < D the rotated B digits
// (gadget's 1 added on the diagonal, l == lvl) and A digits are transformed;
//   cPrime.b = sum_l D(b)_l g_l,  cPrime.a = sum_l D(a)_l g_l,
//   c.a = sum_l D(a)_l s2[l].a + cPrime.b,  c.b = sum_l D(a)_l s2[l].b
// mod p, with g_l = 2^(32 - (l+1) b) N^-1 and s2 in the engine form (N^-1
// included), so the rows carry N^-1 as the engine BSK does.
__global__ void __launch_bounds__(64 * LZ_WAVES, 2) k_lazy_derive(LazyArgs a) {
    constexpr int N = LZ_N, T = LZ_T, K1 = LZ_K1, LOGN = LZ_LOGN;
    __shared__ double scratch[LZ_WAVES][LZ_SCR];
    __shared__ __align__(16) int8_t stage[LZ_WAVES][N];
    const int w = static_cast<int>(threadIdx.x) / 64;
    const int tl = static_cast<int>(threadIdx.x) % 64;
    const int L = a.level, D = a.dlevel;
    const int64_t item = static_cast<int64_t>(blockIdx.x) * LZ_WAVES + w;
    if (item >= static_cast<int64_t>(a.count) * a.steps * L) return;  // whole wavefront: no workgroup barrier follows
    const int lvl = static_cast<int>(item % L);
    const int si = static_cast<int>((item / L) % a.steps);
    const int ct = static_cast<int>(item / (static_cast<int64_t>(L) * a.steps));
    const int comp = a.s0 + si;  // key component: s_{comp + 1}
    int aT, wrap;
    validate_rotator(a.avals[static_cast<size_t>(ct) * (a.n - 1) + comp], N, aT, wrap);
    const bool flip = wrap < 0;
    const PrimeF64 pr = a.prime;
    double* scr = scratch[w];
    int8_t* st = stage[w];
    const size_t cell = static_cast<size_t>(comp) * L + lvl;

    double* __restrict__ out = a.ws + ((static_cast<size_t>(ct) * a.ws_steps + si) * L + lvl) * K1 * K1 * N + tl;
    double cpa[16], ca[16], cb[16];
    // rotateDecompPolynomialMinusOne (polynomial.cpp:273-282) of a stored digit
    // polynomial, narrowed to int8, at this lane's pass-0 positions
    auto rotated = [&](const int8_t* __restrict__ src, bool diag, double (&x)[16]) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        reinterpret_cast<uint4*>(st)[tl] = reinterpret_cast<const uint4*>(src)[tl];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int j = tl + T * r;
            const int s = (j - aT) & (N - 1);
            const int vs = st[s];
            const int t = ((j < aT) != flip ? -vs : vs) - static_cast<int>(st[j]);
            int8_t d = static_cast<int8_t>(t);
            if (diag && j == 0) d = static_cast<int8_t>(d + 1);  // rotatedB[lvl][lvl].coeffs[0] + 1
            x[r] = static_cast<double>(d);
        }
    };
    // the body row first (cPrime.b, stored at once), which also starts c.a;
    // each sum then holds at most 8 products of magnitude <= 0.51 p
    {
        double cpb[16];
#pragma unroll
        for (int r = 0; r < 16; r++) cpb[r] = 0.0;
        for (int l = 0; l < D; l++) {
            double x[16];
            rotated(a.db + (cell * D + l) * N, l == lvl, x);
            ntt_forward<LOGN, 1>(x, scr, tl, a.tw, pr);
            const double g = a.g[l];
#pragma unroll
            for (int r = 0; r < 16; r++) cpb[r] = __dadd_rn(cpb[r], mulmod(x[r], g, pr));
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const double b = centre(cpb[r], pr);
            out[3 * N + r * T] = b;
            ca[r] = b;
            cpa[r] = 0.0;
            cb[r] = 0.0;
        }
    }
    for (int l = 0; l < D; l++) {
        const double g = a.g[l];
        double x[16];
        rotated(a.da + (cell * D + l) * N, false, x);
        ntt_forward<LOGN, 1>(x, scr, tl, a.tw, pr);
        const double* __restrict__ s2a = a.s2e + (static_cast<size_t>(l) * K1 + 0) * N + tl;
        const double* __restrict__ s2b = a.s2e + (static_cast<size_t>(l) * K1 + 1) * N + tl;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            cpa[r] = __dadd_rn(cpa[r], mulmod(x[r], g, pr));
            ca[r] = __dadd_rn(ca[r], mulmod(x[r], s2a[r * T], pr));
            cb[r] = __dadd_rn(cb[r], mulmod(x[r], s2b[r * T], pr));
        }
    }
    // rows of this level: input row 0 (the accumulator's mask digits) is c[0],
    // input row 1 (its body digits) is cPrime; columns a, b
#pragma unroll
    for (int r = 0; r < 16; r++) {
        out[0 * N + r * T] = centre(reduce(ca[r], pr), pr);
        out[1 * N + r * T] = centre(cb[r], pr);
        out[2 * N + r * T] = centre(cpa[r], pr);
    }
}

// The Alt form's first component, deriveFirstComponentNtt (trgsw.cpp:929-955)
// with prepareGdV (911-927): sum_lvl GD(v)_lvl * cPrime_0[lvl] mod p (R =
// lApprox digits of v), inverse-transformed, centred, reduced to the torus.
// One wavefront per test polynomial.
__global__ void __launch_bounds__(64) k_lazy_alt_first(LazyArgs a, int count_v) {
    constexpr int N = LZ_N, T = LZ_T, K1 = LZ_K1, LOGN = LZ_LOGN;
    __shared__ double scr[LZ_SCR];
    const int vi = static_cast<int>(blockIdx.x);
    if (vi >= count_v) return;
    const int tl = static_cast<int>(threadIdx.x);
    const int R = a.level;
    const int rb = a.radix_bits;
    const int32_t* __restrict__ v = a.v + static_cast<size_t>(a.v_stride) * vi;
    const PrimeF64 pr = a.prime;
    uint32_t dg0[16], dg1[16];
#pragma unroll
    for (int r = 0; r < 16; r++) digits_runtime_packed<32>(v[tl + T * r], R, rb, dg0[r], dg1[r]);
    double ca[16], cb[16];
#pragma unroll
    for (int r = 0; r < 16; r++) { ca[r] = 0.0; cb[r] = 0.0; }
    for (int lvl = 0; lvl < R; lvl++) {
        double x[16];
#pragma unroll
        for (int r = 0; r < 16; r++) x[r] = static_cast<double>(digit_runtime(dg0[r], dg1[r], lvl, rb));
        ntt_forward<LOGN, 1>(x, scr, tl, a.tw, pr);
        const double* __restrict__ ra = a.row0e + (static_cast<size_t>(lvl) * K1 + 0) * N + tl;
        const double* __restrict__ rbp = a.row0e + (static_cast<size_t>(lvl) * K1 + 1) * N + tl;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            ca[r] = __dadd_rn(ca[r], mulmod(x[r], ra[r * T], pr));
            cb[r] = __dadd_rn(cb[r], mulmod(x[r], rbp[r * T], pr));
        }
    }
    int32_t* __restrict__ out = a.first_out + static_cast<size_t>(vi) * K1 * N;
    for (int c = 0; c < K1; c++) {
        double x[16];
#pragma unroll
        for (int r = 0; r < 16; r++) x[r] = c == 0 ? ca[r] : cb[r];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        ntt_inverse<LOGN, 1>(x, scr, tl, a.tw + N, pr);
#pragma unroll
        for (int r = 0; r < 16; r++) out[static_cast<size_t>(c) * N + tl + T * r] = static_cast<int32_t>(f64_low32(centre(x[r], pr)));
    }
}

// The Alt form's rows: one wavefront per (ciphertext, step, level lvl < R) as
// k_lazy_derive, from the torus rows (blindRotateLazyPipeAltNttSweep's stageA
// and stageB, blind_rotate.cpp:1035-1087): the mask of cPrime[lvl] of s_{comp}
// times (X^a - 1), cut into D digits, transformed and multiplied into c and
// cPrime.a; the body times (X^a - 1) plus 2^(32 - (lvl+1) b) at coefficient 0
// (rotateTrlweMinusOneBPlusOne, trlwe.cpp:291-297), transformed whole: cPrime.b.
__global__ void __launch_bounds__(64 * LZ_WAVES, 2) k_lazy_alt_derive(LazyArgs a) {
    constexpr int N = LZ_N, T = LZ_T, K1 = LZ_K1, LOGN = LZ_LOGN;
    __shared__ double scratch[LZ_WAVES][LZ_SCR];
    __shared__ int32_t stage[LZ_WAVES][N];
    const int w = static_cast<int>(threadIdx.x) / 64;
    const int tl = static_cast<int>(threadIdx.x) % 64;
    const int L = a.level, D = a.dlevel, rb = a.radix_bits;
    const int64_t item = static_cast<int64_t>(blockIdx.x) * LZ_WAVES + w;
    if (item >= static_cast<int64_t>(a.count) * a.steps * L) return;  // whole wavefront: no workgroup barrier follows
    const int lvl = static_cast<int>(item % L);
    const int si = static_cast<int>((item / L) % a.steps);
    const int ct = static_cast<int>(item / (static_cast<int64_t>(L) * a.steps));
    const int step = a.s0 + si;  // steps 1 .. n-1 are 0 .. n-2 here
    int aT, wrap;
    validate_rotator(a.avals[static_cast<size_t>(ct) * (a.n - 1) + step], N, aT, wrap);
    const bool flip = wrap < 0;
    const PrimeF64 pr = a.prime;
    double* scr = scratch[w];
    int32_t* st = stage[w];
    const int32_t* __restrict__ row = a.rows_t + ((static_cast<size_t>(step + 1) * L + lvl) * K1) * N;
    // (X^a - 1) p at this lane's pass-0 positions, mod 2^32
    auto rotated = [&](const int32_t* __restrict__ src, uint32_t (&out)[16]) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int q = 0; q < 4; q++) reinterpret_cast<uint4*>(st)[tl + 64 * q] = reinterpret_cast<const uint4*>(src)[tl + 64 * q];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int j = tl + T * r;
            const uint32_t vs = static_cast<uint32_t>(st[(j - aT) & (N - 1)]);
            out[r] = (((j < aT) != flip) ? 0u - vs : vs) - static_cast<uint32_t>(st[j]);
        }
    };
    double* __restrict__ out = a.ws + ((static_cast<size_t>(ct) * a.ws_steps + si) * L + lvl) * K1 * K1 * N + tl;
    double cpa[16], ca[16], cb[16];
    uint32_t dg0[16], dg1[16];
    {
        // the body, rotated, plus the gadget's 1 of this level, whole: cPrime.b,
        // stored at once; it also starts c.a
        uint32_t u[16];
        rotated(row + N, u);
        const uint32_t one = uint32_t{1} << (32 - (lvl + 1) * rb);
        double x[16];
#pragma unroll
        for (int r = 0; r < 16; r++) x[r] = static_cast<double>(static_cast<int32_t>((tl == 0 && r == 0) ? u[r] + one : u[r]));
        ntt_forward<LOGN, 1>(x, scr, tl, a.tw, pr);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const double b = centre(mulmod(x[r], a.ninv, pr), pr);
            out[3 * N + r * T] = b;
            ca[r] = b;
            cpa[r] = 0.0;
            cb[r] = 0.0;
        }
        // the mask: D digits of each rotated coefficient (gadgetDecomposeTrlweA)
        rotated(row, u);
#pragma unroll
        for (int r = 0; r < 16; r++) digits_runtime_packed<32>(static_cast<int32_t>(u[r]), D, rb, dg0[r], dg1[r]);
    }
    for (int l = 0; l < D; l++) {
        const double g = a.g[l];
        double x[16];
#pragma unroll
        for (int r = 0; r < 16; r++) x[r] = static_cast<double>(digit_runtime(dg0[r], dg1[r], l, rb));
        ntt_forward<LOGN, 1>(x, scr, tl, a.tw, pr);
        const double* __restrict__ s2a = a.s2e + (static_cast<size_t>(l) * K1 + 0) * N + tl;
        const double* __restrict__ s2b = a.s2e + (static_cast<size_t>(l) * K1 + 1) * N + tl;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            cpa[r] = __dadd_rn(cpa[r], mulmod(x[r], g, pr));
            ca[r] = __dadd_rn(ca[r], mulmod(x[r], s2a[r * T], pr));
            cb[r] = __dadd_rn(cb[r], mulmod(x[r], s2b[r * T], pr));
        }
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        out[0 * N + r * T] = centre(reduce(ca[r], pr), pr);
        out[1 * N + r * T] = centre(cb[r], pr);
        out[2 * N + r * T] = centre(cpa[r], pr);
    }
}

bool lazy_supported(int tbits, int logn, int k) { return tbits == 32 && logn == LZ_LOGN && k == 1; }

hipError_t launch_lazy_alt_first(const LazyArgs& a, int count_v, hipStream_t s) {
    hipLaunchKernelGGL(k_lazy_alt_first, dim3(static_cast<unsigned>(count_v)), dim3(64), 0, s, a, count_v);
    return hipGetLastError();
}

hipError_t launch_lazy_alt_derive(const LazyArgs& a, hipStream_t s) {
    const int64_t items = static_cast<int64_t>(a.count) * a.steps * a.level;
    const int64_t groups = (items + LZ_WAVES - 1) / LZ_WAVES;
    if (groups < 1 || groups > (int64_t{1} << 31) - 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_lazy_alt_derive, dim3(static_cast<unsigned>(groups)), dim3(64 * LZ_WAVES), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_lazy_prepare(const LazyArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_lazy_prepare, dim3(static_cast<unsigned>(a.count)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_lazy_derive(const LazyArgs& a, hipStream_t s) {
    const int64_t items = static_cast<int64_t>(a.count) * a.steps * a.level;
    const int64_t groups = (items + LZ_WAVES - 1) / LZ_WAVES;
    if (groups < 1 || groups > (int64_t{1} << 31) - 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_lazy_derive, dim3(static_cast<unsigned>(groups)), dim3(64 * LZ_WAVES), 0, s, a);
    return hipGetLastError();
}

}  // namespace ygpu
