// ygpu_kernels_lat.hip -- latency form of the TORUS32 blind rotation.
//
// One ciphertext per workgroup, one wavefront per (accumulator row r, gadget
// level l): 2 x L waves.  The throughput kernel gives each row one wavefront
// that runs all L forward transforms and multiply-accumulates in turn.  Here
// the levels run side by side, and the per-step chain gets shorter:
//   every wave: (X^a - 1) row r and its digits (all levels: the carry chain
//               needs them) at a third of the row's positions, into LDS;
//   barrier;
//   every wave: the forward transform of digit level l of row r, and the
//               products with the BSK entries of (level l, row r) for both
//               output columns;
//   barrier;
//   waves (c, 0) and (c, 2): each sums the 2 L partial products of column c
//               over one half of the bit-reversed array, runs stages 9 .. 1
//               of the inverse transform on it, and the two swap half their
//               values for stage 0; then each centres, reduces to the torus
//               and accumulates its half of row c (held in LDS);
//   barrier.
// This leaves most waves idle during the inverse transform, so it is the form
// for batches too small to fill the GPU (the plain drop-in call, short serving
// batches), not for throughput.  The arithmetic is the same exact mod-q arithmetic as
// k_blind_rotate: partial sums differ only in representative, and the inverse
// transform reduces them, so the results are bit-identical.
//
// Its own translation unit: the throughput kernels' register allocation is not
// perturbed.
#include "ygpu_kernels.h"

namespace ygpu {

namespace {

constexpr int LAT_LOGN = 10, LAT_N = 1 << LAT_LOGN, LAT_T = LAT_N / 16, LAT_L = 3, LAT_RB = 7, LAT_K1 = 2;
constexpr int LAT_WAVES = LAT_K1 * LAT_L;
constexpr int LAT_THREADS = 64 * LAT_WAVES;
constexpr int LAT_SCR = Scratch<LAT_LOGN>::DOUBLES;

// The inverse transform of one output column, split between two wavefronts.
// Gentleman-Sande stages 9 .. 1 pair positions within one half of the
// bit-reversed array (j and j + 2^(9 - st), both below or both above 512);
// only stage 0 pairs j with j + 512.  Wave h takes half h: 512 values, 8 per
// lane, in three passes of three stages, then the two waves swap half their
// values for stage 0.  Index j' within the half of lane L, register e:
//   pass 1 (stages 9, 8, 7 on bits 0-2):  j' = L << 3 | e
//   pass 2 (stages 6, 5, 4 on bits 3-5):  j' = (L & 7) | e << 3 | (L >> 3) << 6
//   pass 3 (stages 3, 2, 1 on bits 6-8):  j' = L | e << 6
// Register bit k is the bit of stage (first stage of the pass) - k, so the
// stages run at register distance 1, 2, 4 as in inv_pass_t.
__device__ __forceinline__ int half_idx(int pass, int lane, int e) {
    return pass == 1 ? (lane << 3 | e) : pass == 2 ? ((lane & 7) | e << 3 | (lane >> 3) << 6) : (lane | e << 6);
}
// Padded LDS addresses of the two exchanges: each 16-lane group of a 64-bit
// access touches 16 distinct 8-byte slots (mod 32) on the writing and the
// reading side (575 and 568 doubles, inside a wave's scratch).
__device__ __forceinline__ int half_xa(int j) { return j + (j >> 3); }
__device__ __forceinline__ int half_xb(int j) { return j + 8 * (j >> 6); }

// One pass of three Gentleman-Sande stages on a half; `itw` is the flat table
// of inverse twiddles (index (1 << st) + (j >> (10 - st)) for the pair's lower
// position j).  With reduce_sums the first stage's four sums are reduced after
// it (inputs below 8p: those sums are below 16p); the later stages' bounds
// are then at most 8p (pass 1 from 8p inputs), 4p (pass 2 from reduced
// inputs) and 5p (pass 3 from 4p inputs), all below 16p <= 2^53.
template <int PASS>
__device__ __forceinline__ void half_inv_pass(double (&x)[8], int lane, int h, const double* itw, const PrimeF64& pr, bool reduce_sums) {
    constexpr int ST0 = 9 - 3 * (PASS - 1);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int st = ST0 - k;
        const int half = 1 << k;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            if (e & half) continue;
            const int j = 512 * h + half_idx(PASS, lane, e);
            const double w = itw[(1 << st) + (j >> (10 - st))];
            const double U = x[e];
            const double V = x[e + half];
            x[e] = __dadd_rn(U, V);
            x[e + half] = mulmod(__dadd_rn(U, -V), w, pr);
        }
        if (k == 0 && reduce_sums) {
#pragma unroll
            for (int e = 0; e < 8; e += 2) x[e] = reduce(x[e], pr);
        }
    }
}

}  // namespace

__global__ void __launch_bounds__(LAT_THREADS) k_blind_rotate_lat(BrArgs a) {
    constexpr int N = LAT_N, T = LAT_T, L = LAT_L, K1 = LAT_K1, LOGN = LAT_LOGN;
    // per wave: transform scratch (also its partial sum of column r), and its
    // partial sum of column 1 - r
    __shared__ double scratch[LAT_WAVES][LAT_SCR];
    __shared__ double other[LAT_WAVES][N];
    __shared__ double tws[2 * N];
    __shared__ int32_t acc[K1][N];
    __shared__ int32_t abars[KS_MAX_N];  // the ciphertext's rotation amounts, mod-switched once
    __shared__ uint32_t digs[K1][N];      // packed digits of (X^a - 1) row r, all levels
    // the first two forward stages of a digit polynomial by table lookup
    // (fwd_first_two_stages): w * d mod p for every signed-byte digit d
    __shared__ double firstTab[FIRST_TABLE_ENTRIES];
    // digit-share barrier words of each row's waves, then the failure flag,
    // then the words of the two inverse waves of each column
    __shared__ uint32_t rowCount[K1 * L + 1 + K1 * 2];

    const int ct = map_workgroup(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
    if (ct >= a.count) return;  // whole workgroup
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / 64);
    const int r = w / L;      // accumulator row whose digits this wave transforms
    const int lvl = w % L;    // gadget level
    const int tl = static_cast<int>(threadIdx.x) % 64;
    double* scr = scratch[w];
    if (threadIdx.x < K1 * L + 1 + K1 * 2) rowCount[threadIdx.x] = 0u;
    TeamBarrier<L> rowSync{&rowCount[r * L], &rowCount[K1 * L], 0u, a.barrier_polls, lvl};
    // waves (r, 0) and (r, 2) share column r's inverse transform: half h = lvl / 2
    // (on four different SIMDs with the round-robin wave placement)
    const int ih = lvl >> 1;
    TeamBarrier<2> pairSync{&rowCount[K1 * L + 1 + r * 2], &rowCount[K1 * L], 0u, a.barrier_polls, ih};

    const int n = a.n;
    const int slot = a.ct_slot ? a.ct_slot[ct] : 0;
    const double* __restrict__ bsk = a.bsk_table[slot];
    const int32_t* __restrict__ lwe = a.lwe_in + static_cast<size_t>(ct) * (n + 1);
    const int32_t* __restrict__ v = static_cast<const int32_t*>(a.v) + static_cast<size_t>(a.v_stride) * ct;
    const PrimeF64 pr = a.prime[0];

    for (int j = static_cast<int>(threadIdx.x); j < 2 * N; j += LAT_THREADS) tws[j] = a.tw[j];
    // pass 0's twiddles of both directions stay in (scalar) registers for the
    // whole launch
    TwPass0Reg<TwFlat> fw0{}, iw0{};
    {
        const double* __restrict__ g = a.tw;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            fw0.f[e] = g[e];
            iw0.f[e] = g[N + e];
        }
        fw0.base = TwFlat{tws};
        iw0.base = TwFlat{tws + N};
    }
    build_first_stage_tables(firstTab, a.tw, pr, static_cast<int>(threadIdx.x), LAT_THREADS);
    for (int j = static_cast<int>(threadIdx.x); j < n; j += LAT_THREADS) abars[j] = a.prescaled ? lwe[j] : mod_switch_2n(lwe[j], LOGN + 1);
    // accumulator: caller's, or genNoiselessTrlweSample (trlwe.cpp:175-182)
    // after rescaleTlweToNewMod (tlwe.cpp:75-81): (0, X^{-bbar} v)
    if (lvl == 0) {
        if (a.acc_in) {
            const int32_t* ai = static_cast<const int32_t*>(a.acc_in) + static_cast<size_t>(ct) * K1 * N + r * N;
            for (int q = 0; q < 16; q++) acc[r][tl + T * q] = ai[tl + T * q];
        } else {
            const int bbar = a.prescaled ? lwe[n] : mod_switch_2n(lwe[n], LOGN + 1);
            int aT, wrap;
            validate_rotator(-bbar, N, aT, wrap);
            for (int q = 0; q < 16; q++) {
                const int j = tl + T * q;
                const int src = j - aT;
                const int64_t val = src < 0 ? -static_cast<int64_t>(v[src + N]) * wrap : static_cast<int64_t>(v[src]) * wrap;
                acc[r][j] = r == 0 ? 0 : static_cast<int32_t>(sext_t<32>(val));
            }
        }
    }
    __syncthreads();

    const size_t stepStride = static_cast<size_t>(L) * K1 * K1 * N;  // doubles per BSK row
    using GD = const __attribute__((address_space(1))) double;
    // this wave's BSK entries (level lvl, input row r) of step i, for output
    // column r (own) and 1 - r (oth): kp + i * stepStride.  Step i + 1's are
    // loaded while step i's barriers and inverse transform run.
    const double* kp = bsk + static_cast<size_t>(lvl * K1 + r) * K1 * N + tl;
    double kown[16], koth[16];
    {
        GD* __restrict__ ko = (GD*)(kp + static_cast<size_t>(r) * N);
        GD* __restrict__ kt = (GD*)(kp + static_cast<size_t>(1 - r) * N);
#pragma unroll
        for (int q = 0; q < 16; q++) { kown[q] = ko[q * T]; koth[q] = kt[q * T]; }
    }
    for (int i = 0; i < n; i++) {
        int aT, wrap;
        validate_rotator(abars[i], N, aT, wrap);
        // (X^a - 1) row r (rotateTorusPolynomialMinusOne, polynomial.cpp:213-238)
        // and its digits (gadget_decomposition.h:75-95): the three waves of
        // row r each decompose a share of the positions into LDS, registers
        // q of the lane layout in [0, 4), [4, 8), [8, 16) for levels 0, 1, 2
        // of row 0 and [0, 8), [8, 12), [12, 16) of row 1.  Waves w and
        // w + 4 share a SIMD (wave w = 3 r + level), so waves 0 and 4, 1 and
        // 5 take four registers each and the waves alone on their SIMDs
        // (2 and 3) take eight: every SIMD decomposes eight.
        const bool flip = wrap < 0;
        {
            const int q0 = r == 0 ? (lvl == 0 ? 0 : lvl == 1 ? 4 : 8) : (lvl == 0 ? 0 : lvl == 1 ? 8 : 12);
            const int q1 = r == 0 ? (lvl == 0 ? 4 : lvl == 1 ? 8 : 16) : (lvl == 0 ? 8 : lvl == 1 ? 12 : 16);
            for (int q = q0; q < q1; q++) {
                const int j = tl + T * q;
                const int src = (j - aT) & (N - 1);
                const bool negs = (j < aT) != flip;
                const uint32_t av = static_cast<uint32_t>(acc[r][src]);
                const uint32_t rot = (negs ? 0u - av : av) - static_cast<uint32_t>(acc[r][j]);
                uint32_t d0, d1;
                digits_packed_add32<L, LAT_RB>(rot, d0, d1);
                digs[r][j] = d0;
            }
        }
        // only the three waves of row r read its digits
        rowSync.sync();
        // the first two forward stages of level lvl's digit polynomial by
        // table lookup (1 % per plain call over computing them)
        double x[16];
        {
            uint32_t dg[16];
#pragma unroll
            for (int q = 0; q < 16; q++) dg[q] = digs[r][tl + T * q];
            fwd_first_two_stages(x, dg, lvl, firstTab);
            WorkgroupSync<1> sy;
            ntt_forward_from_stage2_t<LOGN>(x, scr, tl, fw0, pr, sy);
        }
        // products with BSK row (level lvl, input row r): column r and 1 - r
        double po[16], pt[16];
#pragma unroll
        for (int q = 0; q < 16; q++) po[q] = mulmod(x[q], kown[q], pr);
#pragma unroll
        for (int q = 0; q < 16; q++) pt[q] = mulmod(x[q], koth[q], pr);
        if (i + 1 < n) {
            const double* kn = kp + static_cast<size_t>(i + 1) * stepStride;
            GD* __restrict__ ko = (GD*)(kn + static_cast<size_t>(r) * N);
            GD* __restrict__ kt = (GD*)(kn + static_cast<size_t>(1 - r) * N);
#pragma unroll
            for (int q = 0; q < 16; q++) { kown[q] = ko[q * T]; koth[q] = kt[q * T]; }
        }
        // the forward transform's last exchange reads are this wave's own
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int q = 0; q < 16; q++) {
            scr[q * T + tl] = po[q];
            other[w][q * T + tl] = pt[q];
        }
        __syncthreads();
        if (lvl != 1) {
            // column r, half ih of the bit-reversed array: the own products
            // of the waves of row r and the other products of the waves of
            // row 1 - r, read from the forward transform's last-pass layout
            // (position j at register (j & 3) | (j >> 8) << 2, lane (j >> 2) & 63);
            // sums of six products (each at most 1.32p): below 8p
            double s[8];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int j = 512 * ih + half_idx(1, tl, e);
                const int ad = ((j & 3) | (j >> 8) << 2) * T + ((j >> 2) & 63);
                double t = scratch[r * L][ad];
#pragma unroll
                for (int l2 = 0; l2 < L; l2++) {
                    if (l2 > 0) t = __dadd_rn(t, scratch[r * L + l2][ad]);
                    t = __dadd_rn(t, other[(1 - r) * L + l2][ad]);
                }
                s[e] = t;
            }
            const double* itw = tws + N;
            half_inv_pass<1>(s, tl, ih, itw, pr, true);
            // the partner wave has read its share of this wave's scratch
            // before the exchanges rewrite it
            pairSync.sync();
#pragma unroll
            for (int e = 0; e < 8; e++) scr[half_xa(half_idx(1, tl, e))] = s[e];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int e = 0; e < 8; e++) s[e] = reduce(scr[half_xa(half_idx(2, tl, e))], pr);
            half_inv_pass<2>(s, tl, ih, itw, pr, false);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int e = 0; e < 8; e++) scr[half_xb(half_idx(2, tl, e))] = s[e];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int e = 0; e < 8; e++) s[e] = scr[half_xb(half_idx(3, tl, e))];
            half_inv_pass<3>(s, tl, ih, itw, pr, true);
            // stage 0 pairs j' of half 0 (U) with j' of half 1 (V) at the same
            // lane and register (j' = tl + 64 e): wave ih keeps e in
            // [4 ih, 4 ih + 4) and hands the other four values to its partner
            // through the start of its own scratch (its exchange reads are
            // complete: wave order)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int k = 0; k < 4; k++) scr[k * T + tl] = ih ? s[k] : s[4 + k];
            pairSync.sync();
            const double* theirs = scratch[r * L + 2 - lvl];
            const double w1 = iw0.f[1];
            double o[8];  // o[k]: position tl + 64 (4 ih + k); o[4 + k]: that position + 512
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const double mine = ih ? s[4 + k] : s[k];
                const double got = theirs[k * T + tl];
                const double U = ih ? got : mine;
                const double V = ih ? mine : got;
                o[k] = __dadd_rn(U, V);
                o[4 + k] = mulmod(__dadd_rn(U, -V), w1, pr);
            }
            // centre mod q: the sums of registers e = 0, 1 (wave 0) can reach
            // 4p and 10p and take a second reduction; every other output is
            // below 3p (one reduction is exact below 4p, centre16_after_inverse)
#pragma unroll
            for (int k = 0; k < 8; k++) o[k] = reduce(o[k], pr);
            o[0] = reduce(o[0], pr);
            o[1] = reduce(o[1], pr);
            // reduce to the torus, accumulate (mod 2^32)
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int j = tl + 64 * (4 * ih + (k & 3)) + 512 * (k >> 2);
                acc[r][j] = static_cast<int32_t>(static_cast<uint32_t>(acc[r][j]) + f64_low32(o[k]));
            }
        }
        __syncthreads();
    }

    // a row barrier that timed out (TeamBarrier) leaves the step order
    // unproven: the accumulators read INT32_MIN instead and the launch's
    // status word reports it (the step loop ends with a workgroup barrier, so
    // the flag is final here)
    const bool failed = rowCount[K1 * L] != 0u;
    report_barrier_failure(failed, a.status);
    if (lvl == 0) {
        int32_t* __restrict__ outAcc = static_cast<int32_t*>(a.acc_out) + static_cast<size_t>(ct) * K1 * N + r * N;
        for (int q = 0; q < 16; q++) outAcc[tl + T * q] = failed ? INT32_MIN : acc[r][tl + T * q];
    }
}

bool blind_rotate_lat_supported(int tbits, int logn, int level, int radix_bits) {
    return tbits == 32 && logn == LAT_LOGN && level == LAT_L && radix_bits == LAT_RB;
}

hipError_t launch_blind_rotate_lat(const BrArgs& a, hipStream_t s) {
    if (a.n < 1 || a.n > KS_MAX_N) return hipErrorInvalidValue;  // abars[] holds at most KS_MAX_N
    const unsigned grid = static_cast<unsigned>((a.count + 7) / 8 * 8);  // map_workgroup needs a multiple of 8
    hipLaunchKernelGGL(k_blind_rotate_lat, dim3(grid), dim3(LAT_THREADS), 0, s, a);
    return hipGetLastError();
}

}  // namespace ygpu
