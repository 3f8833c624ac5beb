// ygpu_kernels_lat.hip -- latency form of the TORUS32 blind rotation.
//
// One ciphertext per workgroup of eight wavefronts.  The throughput kernel
// gives each accumulator row one wavefront that runs all L forward transforms
// and multiply-accumulates in turn; here the 2 L = 6 forward transforms of a
// step run side by side.  Waves 0-3 transform whole digit polynomials (row
// w >> 1, gadget level w & 1); waves 4-7 each take one half of a level-2
// polynomial (row (w - 4) >> 1, half (w - 4) & 1) and then one half of row
// r's inverse transform.  Waves w and w + 4 share a SIMD (the dispatcher
// places a workgroup's waves round-robin over the four SIMDs, in an order
// that varies between launches), so every SIMD holds one whole and one half
// forward transform, and the four inverse halves run on four SIMDs.  Per step:
//   every wave: (X^a - 1) row r and its digits (all levels: the carry chain
//               needs them) at a quarter of the row's positions, into LDS;
//   row barrier (the four waves of row r);
//   every wave: its forward transform (whole or half) and the products with
//               the BSK entries of (level, row r) for both output columns;
//   barrier;
//   waves 4-7 (column c = row r, half h): the sum of the 2 L partial products
//               of column c over half h of the bit-reversed array, stages
//               9 .. 1 of the inverse transform on it, and a swap of half
//               their values with the partner for stage 0; then each centres,
//               reduces to the torus and accumulates its half of row c (held
//               in LDS);
//   barrier.
// Most waves idle during the inverse transform, so this is the form for
// batches too small to fill the GPU (the plain drop-in call, short serving
// batches), not for throughput.  The arithmetic is the same exact mod-q
// arithmetic as k_blind_rotate: partial sums differ only in representative,
// and the inverse transform reduces them, so the results are bit-identical.
//
// Its own translation unit: the throughput kernels' register allocation is not
// perturbed.
#include "ygpu_kernels.h"

namespace ygpu {

namespace {

constexpr int LAT_LOGN = 10, LAT_N = 1 << LAT_LOGN, LAT_T = LAT_N / 16, LAT_L = 3, LAT_RB = 7, LAT_K1 = 2;
constexpr int LAT_FULL = 4;  // waves transforming whole digit polynomials
constexpr int LAT_WAVES = 8;
constexpr int LAT_THREADS = 64 * LAT_WAVES;
// A half wave's exchange region inside its pair's block (the exchanges use at
// most 575 doubles).
constexpr int LAT_HSTRIDE = 576;
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

// One pass of three Cooley-Tukey stages on a half of a forward transform
// (stage 0, pairing j with j + 512, comes from the digit table first): layout
// 3 runs stages 1, 2, 3 (bits 8, 7, 6), layout 2 stages 4, 5, 6, layout 1
// stages 7, 8, 9, each at register distance 4, 2, 1; `tw` is the flat table
// of forward twiddles.  From the table stage's |x| <= p/2 + 65 the nine stages
// keep every value below 7.8p, no reduction needed (a stage takes a bound b p
// to at most (1.09375 b + 0.5) p).
template <int PASS>
__device__ __forceinline__ void half_fwd_pass(double (&x)[8], int lane, int h, const double* tw, const PrimeF64& pr) {
    constexpr int ST0 = 10 - 3 * PASS;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int st = ST0 + k;
        const int half = 4 >> k;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            if (e & half) continue;
            const int j = 512 * h + half_idx(PASS, lane, e);
            const double w = tw[(1 << st) + (j >> (10 - st))];
            const double U = x[e];
            const double V = mulmod(x[e + half], w, pr);
            x[e] = __dadd_rn(U, V);
            x[e + half] = __dadd_rn(U, -V);
        }
    }
}
// Position j = 512 h + (lane << 3 | e) of a half's last forward pass (layout
// 1) in the whole transform's last-pass layout, register (j & 3) | (j >> 8) << 2
// of lane (j >> 2) & 63: where the BSK entry and the partial products of that
// position live.
__device__ __forceinline__ int half_pos(int h, int lane, int e) {
    const int j = 512 * h + half_idx(1, lane, e);
    return ((j & 3) | (j >> 8) << 2) * LAT_T + ((j >> 2) & 63);
}
// Wave-local LDS ordering between one wave's writes and its own later reads.
__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace

__global__ void __launch_bounds__(LAT_THREADS) k_blind_rotate_lat(BrArgs a) {
    constexpr int N = LAT_N, T = LAT_T, L = LAT_L, K1 = LAT_K1, LOGN = LAT_LOGN;
    // whole-transform waves: transform scratch, then their partial products
    // of column r
    __shared__ double fscr[LAT_FULL][LAT_SCR];
    // each row's pair of half waves: their exchange regions, then (after both
    // forward halves) the pair's partial products of column r, and in the
    // inverse again their exchange regions
    __shared__ double hblk[LAT_K1][2 * LAT_HSTRIDE];
    // partial products of column 1 - r of each (row, level)
    __shared__ double other[LAT_K1 * LAT_L][N];
    __shared__ double tws[2 * N];
    __shared__ int32_t acc[K1][N];
    __shared__ int32_t abars[KS_MAX_N];  // the ciphertext's rotation amounts, mod-switched once
    __shared__ uint32_t digs[K1][N];      // packed digits of (X^a - 1) row r, all levels
    // the first two forward stages of a digit polynomial by table lookup
    // (fwd_first_two_stages): w * d mod p for every signed-byte digit d
    __shared__ double firstTab[FIRST_TABLE_ENTRIES];
    // digit-share barrier words of each row's four waves, the failure flag,
    // then the words of each row's half-wave pair
    constexpr int ROWW = K1 * 4, PAIRW = ROWW + 1;
    __shared__ uint32_t syncWords[PAIRW + K1 * 2];

    const int ct = map_workgroup(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
    if (ct >= a.count) return;  // whole workgroup
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / 64);
    const bool halfw = w >= LAT_FULL;
    const int r = halfw ? (w - LAT_FULL) >> 1 : w >> 1;  // accumulator row whose digits this wave transforms
    const int lvl = halfw ? 2 : (w & 1);                 // gadget level
    const int hh = halfw ? ((w - LAT_FULL) & 1) : 0;     // a half wave's half
    const int rank = halfw ? 2 + hh : lvl;               // this wave's share of row r's digits
    const int tl = static_cast<int>(threadIdx.x) % 64;
    double* fs = fscr[halfw ? 0 : w];
    double* hs = hblk[r] + hh * LAT_HSTRIDE;
    if (threadIdx.x < PAIRW + K1 * 2) syncWords[threadIdx.x] = 0u;
    TeamBarrier<4> rowSync{&syncWords[r * 4], &syncWords[ROWW], 0u, a.barrier_polls, rank};
    // the two half waves of row r: level 2's forward transform and column r's
    // inverse transform
    TeamBarrier<2> pairSync{&syncWords[PAIRW + r * 2], &syncWords[ROWW], 0u, a.barrier_polls, hh};

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
    if (!halfw && lvl == 0) {
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
    // column r (own) and 1 - r (oth): kp + i * stepStride, all sixteen
    // positions of the lane for a whole-transform wave, the eight of its half
    // for a half wave.  Step i + 1's are loaded while step i's barriers and
    // inverse transform run.
    const double* kp = bsk + static_cast<size_t>(lvl * K1 + r) * K1 * N;
    double kown[16], koth[16];
    auto load_bsk = [&](int step, int lane) {
        const double* kb = kp + static_cast<size_t>(step) * stepStride;
        GD* __restrict__ ko = (GD*)(kb + static_cast<size_t>(r) * N);
        GD* __restrict__ kt = (GD*)(kb + static_cast<size_t>(1 - r) * N);
        if (!halfw) {
#pragma unroll
            for (int q = 0; q < 16; q++) { kown[q] = ko[q * T + lane]; koth[q] = kt[q * T + lane]; }
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int ad = half_pos(hh, lane, e);
                kown[e] = ko[ad];
                koth[e] = kt[ad];
            }
        }
    };
    load_bsk(0, tl);
    for (int i = 0; i < n; i++) {
        // the half waves' LDS and key addresses derive from this copy of the
        // lane index, which the empty asm statement makes opaque at every
        // step: they are recomputed (a few shifts) instead of being hoisted
        // out of the loop, where both roles' addresses held in registers
        // across the loop spilled 13-21 VGPRs, reloaded every step (4.6-5.4
        // against 4.1 ms per plain call)
        int tlv = tl;
        asm volatile("" : "+v"(tlv));
        int aT, wrap;
        validate_rotator(abars[i], N, aT, wrap);
        // (X^a - 1) row r (rotateTorusPolynomialMinusOne, polynomial.cpp:213-238)
        // and its digits (gadget_decomposition.h:75-95): the four waves of
        // row r each decompose registers [4 rank, 4 rank + 4) of the lane
        // layout into LDS; every SIMD decomposes eight
        const bool flip = wrap < 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int j = tl + T * (4 * rank + k);
            const int src = (j - aT) & (N - 1);
            const bool negs = (j < aT) != flip;
            const uint32_t av = static_cast<uint32_t>(acc[r][src]);
            const uint32_t rot = (negs ? 0u - av : av) - static_cast<uint32_t>(acc[r][j]);
            uint32_t d0, d1;
            digits_packed_add32<L, LAT_RB>(rot, d0, d1);
            digs[r][j] = d0;
        }
        // only the four waves of row r read its digits
        rowSync.sync();
        if (!halfw) {
            // the first two forward stages of level lvl's digit polynomial by
            // table lookup, then the whole transform and the products with
            // the BSK entries of (level lvl, input row r): column r and 1 - r
            double x[16];
            {
                uint32_t dg[16];
#pragma unroll
                for (int q = 0; q < 16; q++) dg[q] = digs[r][tl + T * q];
                fwd_first_two_stages(x, dg, lvl, firstTab);
                WorkgroupSync<1> sy;
                ntt_forward_from_stage2_t<LOGN>(x, fs, tl, fw0, pr, sy);
            }
            double po[16], pt[16];
#pragma unroll
            for (int q = 0; q < 16; q++) po[q] = mulmod(x[q], kown[q], pr);
#pragma unroll
            for (int q = 0; q < 16; q++) pt[q] = mulmod(x[q], koth[q], pr);
            if (i + 1 < n) load_bsk(i + 1, tl);
            // the forward transform's last exchange reads are this wave's own
            wave_fence();
#pragma unroll
            for (int q = 0; q < 16; q++) {
                fs[q * T + tl] = po[q];
                other[r * L + lvl][q * T + tl] = pt[q];
            }
        } else {
            // half hh of level 2's digit polynomial: stage 0 by lookup,
            // y = a + w1 c (half 0) or a - w1 c (half 1) for the digits a at
            // j' = tl + 64 e and c at j' + 512 (firstTab[0, 256) = w1 d), then
            // stages 1 .. 9 in three passes of three
            double x[8];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int j = tl + 64 * e;
                const double av = static_cast<double>(static_cast<int8_t>(static_cast<uint8_t>(digs[r][j] >> 16)));
                const double t = firstTab[(digs[r][j + 512] >> 16) & 255u];
                x[e] = __dadd_rn(av, hh ? -t : t);
            }
            half_fwd_pass<3>(x, tlv, hh, tws, pr);
#pragma unroll
            for (int e = 0; e < 8; e++) hs[half_xb(half_idx(3, tlv, e))] = x[e];
            wave_fence();
#pragma unroll
            for (int e = 0; e < 8; e++) x[e] = hs[half_xb(half_idx(2, tlv, e))];
            half_fwd_pass<2>(x, tlv, hh, tws, pr);
            wave_fence();
#pragma unroll
            for (int e = 0; e < 8; e++) hs[half_xa(half_idx(2, tlv, e))] = x[e];
            wave_fence();
#pragma unroll
            for (int e = 0; e < 8; e++) x[e] = hs[half_xa(half_idx(1, tlv, e))];
            half_fwd_pass<1>(x, tlv, hh, tws, pr);
            double po[8], pt[8];
#pragma unroll
            for (int e = 0; e < 8; e++) po[e] = mulmod(x[e], kown[e], pr);
#pragma unroll
            for (int e = 0; e < 8; e++) pt[e] = mulmod(x[e], koth[e], pr);
            if (i + 1 < n) load_bsk(i + 1, tl);
            // both halves' exchanges in the pair's block are complete before
            // the block takes the pair's partial products
            pairSync.sync();
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int ad = half_pos(hh, tlv, e);
                hblk[r][ad] = po[e];
                other[r * L + 2][ad] = pt[e];
            }
        }
        __syncthreads();
        if (halfw) {
            // column r, half hh of the bit-reversed array: the own products
            // of the waves of row r and the other products of the waves of
            // row 1 - r, in the whole transform's last-pass layout; sums of
            // six products (each at most 1.32p): below 8p
            const int ih = hh;
            double s8[8];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int ad = half_pos(ih, tlv, e);
                double t = fscr[r * 2][ad];
                t = __dadd_rn(t, fscr[r * 2 + 1][ad]);
                t = __dadd_rn(t, hblk[r][ad]);
#pragma unroll
                for (int l2 = 0; l2 < L; l2++) t = __dadd_rn(t, other[(1 - r) * L + l2][ad]);
                s8[e] = t;
            }
            const double* itw = tws + N;
            half_inv_pass<1>(s8, tlv, ih, itw, pr, true);
            // the partner has read its share of the pair's block before the
            // exchanges rewrite it
            pairSync.sync();
#pragma unroll
            for (int e = 0; e < 8; e++) hs[half_xa(half_idx(1, tlv, e))] = s8[e];
            wave_fence();
#pragma unroll
            for (int e = 0; e < 8; e++) s8[e] = reduce(hs[half_xa(half_idx(2, tlv, e))], pr);
            half_inv_pass<2>(s8, tlv, ih, itw, pr, false);
            wave_fence();
#pragma unroll
            for (int e = 0; e < 8; e++) hs[half_xb(half_idx(2, tlv, e))] = s8[e];
            wave_fence();
#pragma unroll
            for (int e = 0; e < 8; e++) s8[e] = hs[half_xb(half_idx(3, tlv, e))];
            half_inv_pass<3>(s8, tlv, ih, itw, pr, true);
            // stage 0 pairs j' of half 0 (U) with j' of half 1 (V) at the same
            // lane and register (j' = tl + 64 e): wave ih keeps e in
            // [4 ih, 4 ih + 4) and hands the other four values to its partner
            // through the start of its own exchange region (its exchange reads
            // are complete: wave order)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int k = 0; k < 4; k++) hs[k * T + tl] = ih ? s8[k] : s8[4 + k];
            pairSync.sync();
            const double* theirs = hblk[r] + (1 - ih) * LAT_HSTRIDE;
            const double w1 = iw0.f[1];
            double o[8];  // o[k]: position tl + 64 (4 ih + k); o[4 + k]: that position + 512
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const double mine = ih ? s8[4 + k] : s8[k];
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
    const bool failed = syncWords[ROWW] != 0u;
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
