// ygpu_kernels_crt.hip -- blind rotation for NTT moduli above binary64 range
// (TORUS56: q56 = 27 * 2^56 + 1, bootstrapping.cpp:71-80 at 56 bits).
//
// The external product runs modulo two ~49-bit primes whose product bounds
// the exact result, and the value mod q56 is recovered by Garner CRT
// (crt_mod_q).  Each prime gets its own teams: a ciphertext is four teams of
// T = N/16 lanes, team (row r, prime p) holding a copy of accumulator row r,
// transforming row r's digit polynomials mod p, owning output column r mod p.
// Per step and per team: digits of half of row r's positions, swapped with
// the other prime team of row r through LDS (both need all of them); l
// forward transforms and the multiply-accumulate against both BSK columns; the partner-column swap with
// team (1-r, p); the inverse transform of column r; then the two prime teams of
// row r swap half their residues and each reconstructs half the values mod q56
// and updates them in row r.  The accumulator rows live in LDS, read by both
// prime teams.
// Splitting the primes over teams keeps each team's registers within two waves
// per SIMD (no residue of one prime is held across the other), and the two
// primes proceed in parallel.
//
// Kept in its own translation unit so that the one-prime kernel's register
// allocation is not perturbed.
#include "ygpu_kernels.h"

namespace ygpu {

template <class C>
__global__ void __launch_bounds__(C::THREADS, C::MIN_WAVES) k_blind_rotate_crt(BrArgs a) {
    static_assert(C::NP == 2, "two transform primes");
    using TorusT = int64_t;
    constexpr int LOGN = C::LOGN, N = C::N, T = C::T, K1 = 2, NP = 2, TW = C::TEAM_WAVES, CPW = C::CPW;
    constexpr int SCR = Scratch<LOGN>::DOUBLES;
    constexpr int CT_THREADS = 4 * T;

    __shared__ double scratch[CPW][4 * SCR];
    // the accumulator rows live in LDS, shared by the two prime teams of a row
    // (registers are the binding resource); each updates half the positions
    __shared__ int64_t accbuf[CPW][K1][N];
    // forward twiddles of both primes (the forward transform runs l times per
    // step) in LDS unless the tile reads them through L1 (TWL == 0); the
    // inverse twiddles are always read through L1
    constexpr bool FTW_LDS = C::TWL != 0;
    __shared__ double ftw[FTW_LDS ? NP : 1][FTW_LDS ? N : 1];
    // barrier words of the four two-wave teams of each ciphertext (one per
    // wave), and the flag a timed-out team barrier sets
    __shared__ uint32_t teamCount[CPW * 4 * 2 + 1];
    // the specialised shapes may take the forward transform's stage 0 from a
    // table of w1 d mod p for every digit d, per prime (the tile chooses)
    constexpr int TAB = (C::L > 0 && C::TAB == 1) ? 1 : 0;
    constexpr int DH = TAB ? (1 << (C::RB - 1)) : 0;
    constexpr int DTE = TAB ? ((2 * DH + 1 + 7) & ~7) : 1;
    __shared__ double dtab[TAB ? NP : 1][DTE];

    const int grp = map_workgroup(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
    const int first = grp * CPW;
    if (first >= a.count) return;  // whole workgroup
    const int lc = static_cast<int>(threadIdx.x) / CT_THREADS;
    const int ctid = static_cast<int>(threadIdx.x) % CT_THREADS;
    const int tt = ctid / T;       // team: row tm = tt & 1, prime tp = tt >> 1
    const int tm = tt & 1;
    const int tp = tt >> 1;
    const int tl = ctid % T;
    const bool live = first + lc < a.count;
    const int ct = live ? first + lc : a.count - 1;
    double* scr = scratch[lc] + tt * SCR;
    int64_t* acc = accbuf[lc][tm];

    const int n = a.n;
    // the gadget depth: compile-time for the specialised shapes (the level
    // loop is then unrolled, and the digit fields have constant offsets)
    const int L = C::L > 0 ? C::L : a.level;
    const int rb = a.radix_bits;
    const int slot = a.ct_slot ? a.ct_slot[ct] : 0;
    const double* __restrict__ bsk = a.bsk_table[slot];
    const int32_t* __restrict__ lwe = a.lwe_in + static_cast<size_t>(ct) * (n + 1);
    const TorusT* __restrict__ v = static_cast<const TorusT*>(a.v) + static_cast<size_t>(a.v_stride) * ct;
    const PrimeF64 pr = a.prime[tp];
    if (threadIdx.x < CPW * 4 * 2 + 1) teamCount[threadIdx.x] = 0u;
    if constexpr (TAB) {
#pragma unroll
        for (int q = 0; q < NP; q++)
            build_stage0_table_wide(dtab[q], 2 * DH + 1, DH, a.tw[static_cast<size_t>(q) * 2 * N + 1], a.prime[q], static_cast<int>(threadIdx.x),
                                    C::THREADS);
    }
    if constexpr (FTW_LDS) {
        for (int e = static_cast<int>(threadIdx.x); e < NP * N; e += C::THREADS)
            ftw[e / N][e % N] = a.tw[static_cast<size_t>(e / N) * 2 * N + e % N];
    }
    const double* tw = FTW_LDS ? static_cast<const double*>(ftw[tp]) : a.tw + static_cast<size_t>(tp) * 2 * N;
    const double* __restrict__ itw = a.tw + static_cast<size_t>(tp) * 2 * N + N;
    // pass 0's twiddles of this team's prime, both directions, stay in
    // (scalar) registers for the whole launch (the specialised shapes; the
    // run-time shapes have no scalar registers to spare); the prime index is
    // wave-uniform, which readfirstlane makes visible to the compiler
    TwPass0Reg<TwFlat> fw0{}, iw0{};
    {
        const double* __restrict__ g = a.tw + static_cast<size_t>(__builtin_amdgcn_readfirstlane(tp)) * 2 * N;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            fw0.f[e] = g[e];
            iw0.f[e] = g[N + e];
        }
        fw0.base = TwFlat{tw};
        iw0.base = TwFlat{itw};
    }

    // rescaleTlweToNewMod (tlwe.cpp:75-81) of the body, then
    // genNoiselessTrlweSample (trlwe.cpp:175-182): acc = (0, X^{-bbar} v).
    if (tp == 0) {
        if (a.acc_in) {
            const TorusT* ai = static_cast<const TorusT*>(a.acc_in) + static_cast<size_t>(ct) * K1 * N + tm * N;
#pragma unroll
            for (int r = 0; r < 16; r++) acc[tl + T * r] = ai[tl + T * r];
        } else {
            const int bbar = a.prescaled ? lwe[n] : mod_switch_2n(lwe[n], LOGN + 1);
            int aT, wrap;
            validate_rotator(-bbar, N, aT, wrap);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int j = tl + T * r;
                const int src = j - aT;
                const int64_t val = src < 0 ? -static_cast<int64_t>(v[src + N]) * wrap : static_cast<int64_t>(v[src]) * wrap;
                acc[j] = tm == 0 ? TorusT(0) : sext_t<C::TBITS>(val);
            }
        }
    }
    __syncthreads();
    // the transforms and the team's own scratch reuse synchronise only the
    // team's two waves; hand-overs between teams keep workgroup barriers
    using TeamSync = typename std::conditional<TW == 2, TeamBarrier<2>, WorkgroupSync<TW>>::type;
    TeamSync ts{};
    if constexpr (TW == 2)
        ts = TeamSync{&teamCount[(lc * 4 + tt) * 2], &teamCount[CPW * 4 * 2], 0u, a.barrier_polls,
                      __builtin_amdgcn_readfirstlane(tl / 64)};

    const size_t stepStride = static_cast<size_t>(L) * K1 * K1 * NP * N;  // doubles per BSK row
    // the next step's rotation amount is loaded a step ahead, so no step
    // starts with a memory round trip
    int32_t rawNext = lwe[0];
    for (int i = 0; i < n; i++) {
        const int32_t raw = rawNext;
        if (i + 1 < n) rawNext = lwe[i + 1];
        const int abar = a.prescaled ? raw : mod_switch_2n(raw, LOGN + 1);
        int aT, wrap;
        validate_rotator(abar, N, aT, wrap);

        // (X^a - 1) acc_r (rotateTorusPolynomialMinusOne, polynomial.cpp:213-238)
        // at this lane's pass-0 positions, and its digits: the two prime teams
        // of row tm each decompose half the positions (registers [8 tp, 8 tp + 8))
        // and swap the packed digits through their scratch
        uint32_t dg0[16], dg1[16];
        const bool flip = wrap < 0;
        auto decompose = [&](int r) {
            const int j = tl + T * r;
            const int src = (j - aT) & (N - 1);
            const bool negs = (j < aT) != flip;
            const int64_t av = acc[src];
            const int64_t rot = sext_t<C::TBITS>((negs ? -av : av) - acc[j]);
            if constexpr (C::L > 0) {
                if constexpr (C::L * C::RB <= 32) digits_packed_wide<C::TBITS, C::L, C::RB>(rot, dg0[r], dg1[r]);
                else digits_packed<C::TBITS, C::L, C::RB>(rot, dg0[r], dg1[r]);
            } else {
                digits_runtime_packed<C::TBITS>(rot, L, rb, dg0[r], dg1[r]);
            }
        };
        {
            uint32_t* mine = reinterpret_cast<uint32_t*>(scr);
            const uint32_t* theirs = reinterpret_cast<const uint32_t*>(scratch[lc] + (tt ^ 2) * SCR);
            if (tp == 0) {
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    decompose(r);
                    mine[r * T + tl] = dg0[r];
                    mine[N + r * T + tl] = dg1[r];
                }
            } else {
#pragma unroll
                for (int r = 8; r < 16; r++) {
                    decompose(r);
                    mine[r * T + tl] = dg0[r];
                    mine[N + r * T + tl] = dg1[r];
                }
            }
            __syncthreads();
            if (tp == 0) {
#pragma unroll
                for (int r = 8; r < 16; r++) {
                    dg0[r] = theirs[r * T + tl];
                    dg1[r] = theirs[N + r * T + tl];
                }
            } else {
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    dg0[r] = theirs[r * T + tl];
                    dg1[r] = theirs[N + r * T + tl];
                }
            }
            // the partner has read this team's half before its transforms reuse the scratch
            __syncthreads();
        }

        const double* __restrict__ bskRow = bsk + static_cast<size_t>(i) * stepStride;
        double own[16], oth[16];
        if constexpr (C::L == 0) {
#pragma unroll
            for (int r = 0; r < 16; r++) { own[r] = 0.0; oth[r] = 0.0; }
        }
        // Issue priority falls as a step advances, so that of the two waves
        // on one SIMD the one behind issues first.  Waves w and w + 4 of the
        // workgroup share a SIMD: a prime-0 wave and a prime-1 wave, which
        // meet again at the next workgroup barrier.  At equal priority the
        // arbiter favours the older wave, and the prime-0 waves ran ahead
        // and then idled at the hand-over while the prime-1 waves finished
        // alone on their SIMDs.  Between two workgroup barriers the priority
        // only decreases: 3, 2, 1 over the levels, then 0 to the hand-over;
        // 2 for the inverse, 1 for centring and reconstruction, 0 after.
        auto level = [&](const int lvl) {
            if (lvl == 0) __builtin_amdgcn_s_setprio(3);
            else if (lvl == 1) __builtin_amdgcn_s_setprio(2);
            else __builtin_amdgcn_s_setprio(1);
            const int d = lvl * K1 + tm;
            const double* __restrict__ kp = bskRow + (static_cast<size_t>(d) * K1 * NP + tp) * N + tl;
            // global address space: the compiler cannot infer it through the
            // key table and would issue flat loads (which also count against
            // the LDS wait counter)
            using GD = const __attribute__((address_space(1))) double;
            GD* __restrict__ kpo = (GD*)(kp + static_cast<size_t>(tm) * NP * N);
            GD* __restrict__ kpt = (GD*)(kp + static_cast<size_t>(1 - tm) * NP * N);
            double x[16];
            int dv[16];
#pragma unroll
            for (int r = 0; r < 16; r++) {
                if constexpr (C::L > 0 && C::RB <= 7) {
                    dv[r] = __builtin_amdgcn_sbfe(static_cast<int>(dg0[r]), 8 * lvl, 8);
                } else if constexpr (C::L > 0) {
                    const uint32_t w = lvl < 2 ? dg0[r] : dg1[r];
                    dv[r] = __builtin_amdgcn_sbfe(static_cast<int>(w), 16 * (lvl & 1), 16);
                } else {
                    dv[r] = digit_runtime(dg0[r], dg1[r], lvl, rb);
                }
            }
            if constexpr (TAB) {
                const double* tb = dtab[tp];
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    const double av = static_cast<double>(dv[r]);
                    const double t = tb[dv[r + 8] + DH];
                    x[r] = __dadd_rn(av, t);
                    x[r + 8] = __dadd_rn(av, -t);
                }
                ntt_forward_from_stage1_t<LOGN>(x, scr, tl, fw0, pr, ts);
            } else {
#pragma unroll
                for (int r = 0; r < 16; r++) x[r] = static_cast<double>(dv[r]);
                if constexpr (C::L > 0) ntt_forward_sync_t<LOGN>(x, scr, tl, fw0, pr, ts);
                else ntt_forward_sync<LOGN>(x, scr, tl, tw, pr, ts);
            }
            // one column at a time: the scheduling barrier keeps the second
            // column's loads from being hoisted over the first column's
            // products, which would hold 32 more registers live; the first
            // level of an unrolled loop assigns instead of adding
            const bool firstLvl = C::L > 0 && lvl == 0;
            double kb[16];
#pragma unroll
            for (int r = 0; r < 16; r++) kb[r] = kpo[r * T];
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const double pm = mulmod(x[r], kb[r], pr);
                own[r] = firstLvl ? pm : __dadd_rn(own[r], pm);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < 16; r++) kb[r] = kpt[r * T];
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const double pm = mulmod(x[r], kb[r], pr);
                oth[r] = firstLvl ? pm : __dadd_rn(oth[r], pm);
            }
            if constexpr (C::L == 0) {
                // more than four levels: reduce the partial sums once after
                // four, so that with the partner's they stay below 16p
                if (L > 4 && lvl == 3) {
#pragma unroll
                    for (int r = 0; r < 16; r++) { own[r] = reduce(own[r], pr); oth[r] = reduce(oth[r], pr); }
                }
            }
        };
        if constexpr (C::L > 0) {
#pragma unroll
            for (int lvl = 0; lvl < C::L; lvl++) level(lvl);
        } else {
            for (int lvl = 0; lvl < L; lvl++) level(lvl);
        }

        __builtin_amdgcn_s_setprio(0);
        // partner sums (same prime): hand column 1-tm to team (1-tm, tp)
        ts.sync();
#pragma unroll
        for (int r = 0; r < 16; r++) scr[r * T + tl] = oth[r];
        __syncthreads();
        double x[16];
        {
            const double* other = scratch[lc] + (tt ^ 1) * SCR;
#pragma unroll
            for (int r = 0; r < 16; r++) x[r] = __dadd_rn(own[r], other[r * T + tl]);
        }
        // the partner's reads must finish before this team's scratch is
        // rewritten (the transform's own barriers order only this team)
        __syncthreads();

        __builtin_amdgcn_s_setprio(2);
        if constexpr (C::L > 0) ntt_inverse_sync_t<LOGN>(x, scr, tl, iw0, pr, ts);
        else ntt_inverse_sync<LOGN>(x, scr, tl, itw, pr, ts);

        // residues of column tm mod both primes: the two prime teams of row tm
        // split the reconstruction, team tp taking registers [8 tp, 8 tp + 8);
        // each hands the other team the half it needs
        __builtin_amdgcn_s_setprio(1);
        centre16_after_inverse(x, pr);
        ts.sync();
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int rr = r + 8 * (1 - tp);
            scr[rr * T + tl] = tp ? x[r] : x[r + 8];
        }
        __syncthreads();
        {
            const double* other = scratch[lc] + (tt ^ 2) * SCR;
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int rr = r + 8 * tp;
                const double mine = tp ? x[r + 8] : x[r];
                const double theirs = other[rr * T + tl];
                const int j = tl + T * rr;
                const int64_t val = crt_mod_q(tp ? theirs : mine, tp ? mine : theirs, a.prime[1], a.crt);
                acc[j] = sext_t<C::TBITS>(acc[j] + val);
            }
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
    }

    // a team barrier that timed out (TeamBarrier) leaves the step order
    // unproven: every accumulator of the workgroup reads INT64_MIN instead and
    // the launch's status word reports it (the step loop ends with a workgroup
    // barrier, so the flag is final here)
    const bool failed = teamCount[CPW * 4 * 2] != 0u;
    report_barrier_failure(failed, a.status);
    if (live && tp == 0) {
        TorusT* __restrict__ outAcc = static_cast<TorusT*>(a.acc_out) + static_cast<size_t>(ct) * K1 * N + tm * N;
#pragma unroll
        for (int r = 0; r < 16; r++) outAcc[tl + T * r] = failed ? INT64_MIN : acc[tl + T * r];
    }
}

template <class C>
hipError_t launch_br_crt(const BrArgs& a, hipStream_t s) {
    const int groups = (a.count + C::CPW - 1) / C::CPW;
    const unsigned grid = static_cast<unsigned>((groups + 7) / 8 * 8);  // map_workgroup needs a multiple of 8
    hipLaunchKernelGGL(k_blind_rotate_crt<C>, dim3(grid), dim3(C::THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_blind_rotate_crt(int logn, const BrTile& t, int radix_bits, int level, const BrArgs& a, hipStream_t s) {
    if (logn == 11 && level == 3 && radix_bits == 10) {
        if (t.tab == 1)
            return t.twl ? launch_br_crt<BrCfg<56, 11, 2, 1, false, 3, 10, 1, 1>>(a, s) : launch_br_crt<BrCfg<56, 11, 2, 1, false, 3, 10, 0, 1>>(a, s);
        return t.twl ? launch_br_crt<BrCfg<56, 11, 2, 1, false, 3, 10, 1>>(a, s) : launch_br_crt<BrCfg<56, 11, 2, 1, false, 3, 10, 0>>(a, s);
    }
    const bool p8 = radix_bits <= 8;
    if (logn == 10) return p8 ? launch_br_crt<BrCfg<56, 10, 2, 1, true>>(a, s) : launch_br_crt<BrCfg<56, 10, 2, 1, false>>(a, s);
    if (logn == 11) return p8 ? launch_br_crt<BrCfg<56, 11, 2, 1, true>>(a, s) : launch_br_crt<BrCfg<56, 11, 2, 1, false>>(a, s);
    return hipErrorInvalidValue;
}

}  // namespace ygpu
