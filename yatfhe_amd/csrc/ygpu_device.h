// ygpu_device.h -- device arithmetic and the wave-level negacyclic NTT for gfx950.
//
// Exact modular arithmetic in binary64.  A residue mod a prime p < 2^49 is held
// as a double that is an integer of magnitude < 16p <= 2^53, so every sum,
// difference and fused product below is exact.  The 64-bit integer multiplier
// on CDNA is built from quarter-rate 32-bit multiplies; FP64 FMA is full rate,
// so a Barrett-style product in FP64 (a handful of full-rate ops) is the
// fastest exact modmul this VALU offers.
//
// mulmod(x, w):  h = fl(x*w), l = x*w - h (exact, by FMA), t = rint(h / p),
//                r = h - t*p (exact, by FMA), return r + l.
//   x*w - (r + l) = t*p exactly, so the result is congruent to x*w, and
//   |r + l| = p |x*w/p - t| <= p (0.5 + 3 * 2^-53 |x*w| / p): with |w| <= p/2
//   the magnitude is at most p (0.5 + 1.5 |x| 2^-53), e.g. 1.34p for |x| <= 9p.
// reduce(x):     x - rint(x / p) * p, magnitude <= p/2 + 1.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

namespace ygpu {

struct PrimeF64 {
    double p;      // the prime, exact
    double pinv;   // fl(1/p)
    double half;   // (p - 1) / 2, exact
};

__device__ __forceinline__ double mulmod(double x, double w, const PrimeF64& P) {
    const double h = __dmul_rn(x, w);
    const double l = __fma_rn(x, w, -h);
    const double t = __builtin_rint(__dmul_rn(h, P.pinv));
    const double r = __fma_rn(-t, P.p, h);
    return __dadd_rn(r, l);
}

__device__ __forceinline__ double reduce(double x, const PrimeF64& P) {
    const double t = __builtin_rint(__dmul_rn(x, P.pinv));
    return __fma_rn(-t, P.p, x);
}

// x[r] = reduce(x[r]) for all 16 values, four at a time (one instruction of
// each of four independent reductions in turn).
__device__ __forceinline__ void reduce16(double (&x)[16], const PrimeF64& P) {
#pragma unroll
    for (int g = 0; g < 16; g += 4) {
        double t[4];
#pragma unroll
        for (int b = 0; b < 4; b++) t[b] = __dmul_rn(x[g + b], P.pinv);
#pragma unroll
        for (int b = 0; b < 4; b++) t[b] = __builtin_rint(t[b]);
#pragma unroll
        for (int b = 0; b < 4; b++) x[g + b] = __fma_rn(-t[b], P.p, x[g + b]);
    }
}
// Exact centred representative in [-(p-1)/2, (p-1)/2] of a value |x| <= 9p.
// The first reduction leaves |r| <= (p+1)/2: x * fl(1/p) is within
// 9p 2^-102 + 2^-50 <= 17 2^-53 of x/p (p < 2^49), so |r| <= p/2 + 1.07; the
// second maps exactly the two out-of-range values
// +-(p+1)/2 to -+(p-1)/2: r * fl(1/p) stays below 1/2 - 2^-51 in magnitude for
// |r| <= (p-1)/2 and above 1/2 + 2^-51 for |r| = (p+1)/2, so its rint is 0 or
// +-1 as required.  Six instructions, against ten with two compare-selects.
__device__ __forceinline__ double centre(double x, const PrimeF64& P) {
    return reduce(reduce(x, P), P);
}

// ---------------------------------------------------------------------------
// Register layout of the transform.  A team of T = N/16 lanes holds one
// polynomial, 16 values per lane.  The log2(N) stages run in three passes of
// 4, 4 and log2(N)-8 stages; within a pass every butterfly is lane-local.
// Pass P covers the index bits [lo, lo+S) (lo = log2N - 4P - S); lane tl,
// register r = g*2^S + e holds index  low(G) | e << lo | high(G) << (lo+S)
// with G = tl + T*g.  Pass 0 is natural order j = tl + T*r.
template <int LOGN, int P>
struct Pass {
    static constexpr int N = 1 << LOGN;
    static constexpr int T = N / 16;
    static constexpr int S = (P < 2) ? 4 : (LOGN - 8);
    static constexpr int S0 = 4 * P;
    static constexpr int LO = LOGN - S0 - S;
    __host__ __device__ __forceinline__ static int idx(int tl, int r) {
        const int g = r >> S;
        const int e = r & ((1 << S) - 1);
        const int G = tl + T * g;
        return (G & ((1 << LO) - 1)) | (e << LO) | ((G >> LO) << (LO + S));
    }
};

// LDS addresses of the two exchanges (0: passes 0<->1, 1: passes 1<->2), padded
// so that each 16-lane group of a 64-bit access touches 16 distinct 8-byte
// slots on both the writing and the reading side.
template <int LOGN, int X>
struct Xch {
    static constexpr int S2 = LOGN - 8;
    static constexpr int P12 = 272 + (16 >> S2);
    static constexpr int SIZE = (X == 0) ? ((1 << LOGN) + (1 << (LOGN - 8)) * 16) : ((1 << S2) * P12);
    __device__ __forceinline__ static int addr(int j) {
        if constexpr (X == 0) {
            return j + (1 << (LOGN - 8)) * (j >> (LOGN - 4));
        } else {
            const int c = j & ((1 << S2) - 1);
            const int G = j >> S2;
            return c * P12 + G + (G >> 4);
        }
    }
};

template <int LOGN>
struct Scratch {
    static constexpr int A = Xch<LOGN, 0>::SIZE;
    static constexpr int B = Xch<LOGN, 1>::SIZE;
    static constexpr int DOUBLES = A > B ? A : B;
};

// Synchronisation of the lanes of one team: a single wavefront executes its
// LDS instructions in order, so only the compiler needs fencing; a two-wave
// team uses the workgroup barrier (every team of a workgroup runs the same
// sequence of exchanges).
template <int TEAM_WAVES>
__device__ __forceinline__ void team_sync() {
    if constexpr (TEAM_WAVES == 1) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

template <int TEAM_WAVES>
struct WorkgroupSync {
    __device__ __forceinline__ void sync() { team_sync<TEAM_WAVES>(); }
};

// Bounded wait of a wavefront until the LDS word *ctr reaches `target`
// (another wave of the workgroup raises it).  `polls` bounds the wait as in
// TeamBarrier: with no polls left the wave stores 1 into *fail and stops
// waiting, and a set *fail ends every later wait at once (polls = 0 fails
// without waiting).  The loop is one asm statement: as compiler-visible
// control flow in the middle of the blind-rotation step it cost that kernel
// a large register spill.  Its LDS reads complete inside the statement
// (s_waitcnt lgkmcnt(0) before any use), and it writes only LDS.
__device__ __forceinline__ void lds_wait_at_least(const uint32_t* ctr, uint32_t target, int polls, uint32_t* fail) {
    using L3 = __attribute__((address_space(3))) const uint32_t;
    using L3w = __attribute__((address_space(3))) uint32_t;
    const uint32_t ca = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((L3*)ctr));
    const uint32_t fa = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((L3w*)fail));
    uint32_t v, sv, cnt;
    asm volatile(
        "s_mov_b32 %[cnt], %[polls]\n"
        "1:\n"
        "s_cmp_le_i32 %[cnt], 0\n"
        "s_cbranch_scc1 2f\n"
        "ds_read_b32 %[v], %[ca]\n"
        "s_waitcnt lgkmcnt(0)\n"
        "v_readfirstlane_b32 %[sv], %[v]\n"
        "s_nop 4\n"
        "s_cmp_ge_u32 %[sv], %[target]\n"
        "s_cbranch_scc1 3f\n"
        "ds_read_b32 %[v], %[fa]\n"
        "s_waitcnt lgkmcnt(0)\n"
        "v_readfirstlane_b32 %[sv], %[v]\n"
        "s_nop 4\n"
        "s_cmp_lg_u32 %[sv], 0\n"
        "s_cbranch_scc1 3f\n"
        "s_sub_i32 %[cnt], %[cnt], 1\n"
        "s_sleep 1\n"
        "s_branch 1b\n"
        "2:\n"
        "v_mov_b32 %[v], 1\n"
        "ds_write_b32 %[fa], %[v]\n"
        "s_waitcnt lgkmcnt(0)\n"
        "3:\n"
        : [v] "=&v"(v), [sv] "=&s"(sv), [cnt] "=&s"(cnt)
        : [ca] "v"(ca), [fa] "v"(fa), [target] "s"(target), [polls] "s"(polls)
        : "memory", "scc");
}

// Barrier of the W wavefronts of one team through LDS, for teams smaller than
// the workgroup: other teams do not wait.  Each wave of the team owns one
// word, the number of barriers it has passed: it stores the new count (all
// lanes store the same value, so no atomic is needed) once its earlier LDS
// writes have completed, then waits until every other wave's word has
// reached it.  Every team member passes the same sequence of barriers, so the
// counts agree.  The waits are bounded (lds_wait_at_least): a wave that has
// polled `polls` times without the counts arriving sets *fail, and from then
// on no wait blocks, so a miscount ends in wrong results and a set flag, never
// in a hang (polls = 0 fails every barrier at once).  The kernel reports the
// flag (report_barrier_failure).
constexpr int TEAM_BARRIER_POLLS = 1 << 16;
template <int W>
struct TeamBarrier {
    uint32_t* words;  // the team's W words, one per wave
    uint32_t* fail;
    uint32_t passed;
    int polls;
    int me;           // this wave's index in the team (wave-uniform)
    __device__ __forceinline__ void sync() {
        passed += 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __hip_atomic_store(&words[me], passed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
        for (int o = 1; o < W; o++) {
            const int other = me + o < W ? me + o : me + o - W;
            lds_wait_at_least(&words[other], passed, polls, fail);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
};

// A workgroup whose team barrier timed out reports it to the host: one lane
// stores 1 into the launch's status word (host-visible memory; every failing
// workgroup stores the same value, so no atomic is needed).
__device__ __forceinline__ void report_barrier_failure(bool failed, uint32_t* status) {
    if (failed && threadIdx.x == 0 && status) *reinterpret_cast<volatile uint32_t*>(status) = 1u;
}

// Where a pass finds its twiddles.  TwFlat: the bit-reversed table of psi
// powers (forward [N], inverse [N] after it), index (1 << st) + (j >> (LOGN - st)).
// TwLane (N = 1024): passes 1 and 2 read lane-major tables instead, in which
// the twiddles a lane needs for one stage sit at k * 64 + tl for its k-th
// distinct twiddle, so the 64 lanes of a read touch 64 consecutive doubles:
// no bank conflicts (the flat table's reads in these passes are strided, up to
// four-way conflicting).  Pass 0's twiddles are uniform across lanes and stay
// in the flat table.
struct TwFlat {
    const double* tw;
    template <int LOGN, int P>
    __device__ __forceinline__ double get(int st, int tl, int r) const {
        const int j = Pass<LOGN, P>::idx(tl, r);
        return tw[(1 << st) + (j >> (LOGN - st))];
    }
};
// Lane-table geometry (LOGN = 10, T = 64): pass 1 has 1 + 2 + 4 + 8 = 15
// distinct twiddles per lane over its four stages, pass 2 has 4 + 8 = 12.
constexpr int TW_LANE_P1 = 15, TW_LANE_P2 = 12, TW_LANE_PER_DIR = (TW_LANE_P1 + TW_LANE_P2) * 64;
// Offset (in units of 64 doubles) of stage ls's block, and the lane's k for
// register r: pass 1 (j = (tl & 3) | r << 2 | (tl >> 2) << 6): the twiddle of
// stage 4 + ls is indexed by (tl >> 2) << ls | r >> (4 - ls), so k = r >> (4 - ls);
// pass 2 (r = 4g + e, j = e | (tl + 64 g) << 2): stage 8 + ls is indexed by
// (tl + 64 g) << ls | e >> (2 - ls), so k = g (ls = 0) or 2g + (e >> 1) (ls = 1).
__host__ __device__ constexpr int tw_lane_block(int P, int ls) {
    return P == 1 ? (1 << ls) - 1 : TW_LANE_P1 + (ls == 0 ? 0 : 4);
}
__host__ __device__ constexpr int tw_lane_k(int P, int ls, int r) {
    return P == 1 ? (r >> (4 - ls)) : (ls == 0 ? (r >> 2) : 2 * (r >> 2) + ((r & 3) >> 1));
}
struct TwLane {
    const double* flat;  // the flat table of this direction (pass 0)
    const double* lane;  // this direction's lane tables
    template <int LOGN, int P>
    __device__ __forceinline__ double get(int st, int tl, int r) const {
        static_assert(LOGN == 10, "lane tables: N = 1024");
        if constexpr (P == 0) {
            const int j = Pass<LOGN, 0>::idx(tl, r);
            return flat[(1 << st) + (j >> (LOGN - st))];
        } else {
            const int ls = st - Pass<LOGN, P>::S0;
            return lane[(tw_lane_block(P, ls) + tw_lane_k(P, ls, r)) * 64 + tl];
        }
    }
};
// Pass 0's twiddles held in registers (wave-uniform values, so scalar
// registers), loaded once per launch; the other passes read `base`.  In pass 0
// lane tl's register r holds position j = tl + T r with T = N/16 > tl, so the
// stage-st twiddle index (1 << st) + (j >> (LOGN - st)) is (1 << st) + (r >> (4 - st)).
template <class BASE>
struct TwPass0Reg {
    double f[16];  // the direction's flat table entries 0..15
    BASE base;
    template <int LOGN, int P>
    __device__ __forceinline__ double get(int st, int tl, int r) const {
        if constexpr (P == 0) return f[(1 << st) + (r >> (4 - st))];
        else return base.template get<LOGN, P>(st, tl, r);
    }
};

// Fills one direction's lane tables from its flat table (all threads of the
// workgroup take part; the caller synchronises before use).
__device__ __forceinline__ void build_tw_lane(double* lane, const double* flat, int tid, int nthreads) {
    for (int e = tid; e < TW_LANE_PER_DIR; e += nthreads) {
        const int tl = e & 63, blk = e >> 6;
        int P, ls, k;
        if (blk < TW_LANE_P1) {
            P = 1;
            ls = blk < 1 ? 0 : blk < 3 ? 1 : blk < 7 ? 2 : 3;
            k = blk - ((1 << ls) - 1);
        } else {
            P = 2;
            ls = blk - TW_LANE_P1 < 4 ? 0 : 1;
            k = blk - TW_LANE_P1 - (ls == 0 ? 0 : 4);
        }
        // a register r with this k (the lowest one)
        int r = 0;
        if (P == 1) r = k << (4 - ls);
        else r = ls == 0 ? 4 * k : 4 * (k >> 1) + 2 * (k & 1);
        const int st = (P == 1 ? 4 : 8) + ls;
        const int j = P == 1 ? Pass<10, 1>::idx(tl, r) : Pass<10, 2>::idx(tl, r);
        lane[e] = flat[(1 << st) + (j >> (10 - st))];
    }
}

// Cooley-Tukey pass (forward): U, V -> U + wV, U - wV; stages LS0.. of the pass.
template <int LOGN, int P, int LS0 = 0, class TW>
__device__ __forceinline__ void fwd_pass_t(double (&x)[16], int tl, const TW& tw, const PrimeF64& pr) {
    using PS = Pass<LOGN, P>;
#pragma unroll
    for (int ls = LS0; ls < PS::S; ls++) {
        const int st = PS::S0 + ls;
        const int half = 1 << (PS::S - 1 - ls);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            if (r & half) continue;
            const double w = tw.template get<LOGN, P>(st, tl, r);
            const double U = x[r];
            const double V = mulmod(x[r + half], w, pr);
            x[r] = __dadd_rn(U, V);
            x[r + half] = __dadd_rn(U, -V);
        }
    }
}
template <int LOGN, int P, int LS0 = 0>
__device__ __forceinline__ void fwd_pass(double (&x)[16], int tl, const double* __restrict__ tw, const PrimeF64& pr) {
    fwd_pass_t<LOGN, P, LS0>(x, tl, TwFlat{tw}, pr);
}
// Gentleman-Sande pass (inverse): U, V -> U + V, (U - V) w.
template <int LOGN, int P, class TW>
__device__ __forceinline__ void inv_pass_t(double (&x)[16], int tl, const TW& itw, const PrimeF64& pr) {
    using PS = Pass<LOGN, P>;
#pragma unroll
    for (int ls = PS::S - 1; ls >= 0; ls--) {
        const int st = PS::S0 + ls;
        const int half = 1 << (PS::S - 1 - ls);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            if (r & half) continue;
            const double w = itw.template get<LOGN, P>(st, tl, r);
            const double U = x[r];
            const double V = x[r + half];
            x[r] = __dadd_rn(U, V);
            x[r + half] = mulmod(__dadd_rn(U, -V), w, pr);
        }
    }
}
template <int LOGN, int P>
__device__ __forceinline__ void inv_pass(double (&x)[16], int tl, const double* __restrict__ itw, const PrimeF64& pr) {
    inv_pass_t<LOGN, P>(x, tl, TwFlat{itw}, pr);
}

// A four-stage Gentleman-Sande pass on inputs |x| <= 4p (+ a few units: the
// outputs of a pass that began reduced) without first reducing all sixteen:
// the first stage's sums (below 8p) are reduced after it, its differences go
// through mulmod (below 1.25p), and the worst case over the registers after
// the three later stages is 10p (register 1), so every value stays below
// 16p <= 2^53.  Eight reductions instead of sixteen.
template <int LOGN, int P, class TW>
__device__ __forceinline__ void inv_pass_from_4p(double (&x)[16], int tl, const TW& itw, const PrimeF64& pr) {
    using PS = Pass<LOGN, P>;
    static_assert(PS::S == 4, "a four-stage pass");
#pragma unroll
    for (int r = 0; r < 16; r += 2) {
        const double w = itw.template get<LOGN, P>(PS::S0 + PS::S - 1, tl, r);
        const double U = x[r];
        const double V = x[r + 1];
        x[r] = __dadd_rn(U, V);
        x[r + 1] = mulmod(__dadd_rn(U, -V), w, pr);
    }
#pragma unroll
    for (int g = 0; g < 16; g += 8) {
        double t[4];
#pragma unroll
        for (int b = 0; b < 4; b++) t[b] = __dmul_rn(x[g + 2 * b], pr.pinv);
#pragma unroll
        for (int b = 0; b < 4; b++) t[b] = __builtin_rint(t[b]);
#pragma unroll
        for (int b = 0; b < 4; b++) x[g + 2 * b] = __fma_rn(-t[b], pr.p, x[g + 2 * b]);
    }
#pragma unroll
    for (int ls = PS::S - 2; ls >= 0; ls--) {
        const int st = PS::S0 + ls;
        const int half = 1 << (PS::S - 1 - ls);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            if (r & half) continue;
            const double w = itw.template get<LOGN, P>(st, tl, r);
            const double U = x[r];
            const double V = x[r + half];
            x[r] = __dadd_rn(U, V);
            x[r + half] = mulmod(__dadd_rn(U, -V), w, pr);
        }
    }
}

// Exact centred representatives of the outputs of an inverse transform
// (ntt_inverse_sync_t's reduction schedule: pass 0 from inputs below 3p with
// its first stage's sums reduced).  Register r's worst case after the last
// pass is 4p for r = 0, 8.41p for r = 1 and at most 2.79p for the others
// (with a last pass that began reduced: 8p, 4.75p and 2.75p).  One reduction
// is exact below 4p, two below 9p (centre): with
// p < 2^49, the computed x * fl(1/p) is within 2^-51 (from fl(1/p)) plus half
// an ulp of a value below 4, 2^-52, of x/p, which is less than x/p's distance
// 1/(2p) > 2^-50 from the nearest half-integer (x is an integer, p odd), so
// its rint is the nearest integer to x/p.  Registers 0 and 1 get centre()'s
// second reduction.
__device__ __forceinline__ void centre16_after_inverse(double (&x)[16], const PrimeF64& pr) {
    reduce16(x, pr);
    x[0] = reduce(x[0], pr);
    x[1] = reduce(x[1], pr);
}

// The inverse transform's first pass for N = 1024 (stages 9 and 8) on inputs
// |x| <= 8p, without the reduction of all sixteen values that a general input
// needs first: stage 9's sums stay below 16p <= 2^53 (exact) and its
// differences (below 16p) go through mulmod (below 2p); only the eight sums
// are reduced (to p/2 + 1) before stage 8, after which every value is below 4p.
// The external product's sums of six products of at most 1.32p (centred key
// entries, forward outputs below 8.7p) are below 8p.
template <class TW>
__device__ __forceinline__ void inv_pass2_from_8p(double (&x)[16], int tl, const TW& itw, const PrimeF64& pr) {
#pragma unroll
    for (int r = 0; r < 16; r += 2) {
        const double w = itw.template get<10, 2>(9, tl, r);
        const double U = x[r];
        const double V = x[r + 1];
        x[r] = __dadd_rn(U, V);
        x[r + 1] = mulmod(__dadd_rn(U, -V), w, pr);
    }
#pragma unroll
    for (int g = 0; g < 16; g += 8) {
        double t[4];
#pragma unroll
        for (int b = 0; b < 4; b++) t[b] = __dmul_rn(x[g + 2 * b], pr.pinv);
#pragma unroll
        for (int b = 0; b < 4; b++) t[b] = __builtin_rint(t[b]);
#pragma unroll
        for (int b = 0; b < 4; b++) x[g + 2 * b] = __fma_rn(-t[b], pr.p, x[g + 2 * b]);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        if (r & 2) continue;
        const double w = itw.template get<10, 2>(8, tl, r);
        const double U = x[r];
        const double V = x[r + 2];
        x[r] = __dadd_rn(U, V);
        x[r + 2] = mulmod(__dadd_rn(U, -V), w, pr);
    }
}

template <int LOGN, int PW, int PR, int X, class Sync>
__device__ __forceinline__ void exchange(double (&x)[16], double* scr, int tl, Sync& sy) {
    sy.sync();
#pragma unroll
    for (int r = 0; r < 16; r++) scr[Xch<LOGN, X>::addr(Pass<LOGN, PW>::idx(tl, r))] = x[r];
    sy.sync();
#pragma unroll
    for (int r = 0; r < 16; r++) x[r] = scr[Xch<LOGN, X>::addr(Pass<LOGN, PR>::idx(tl, r))];
}

// Forward negacyclic NTT: natural order in pass-0 layout -> bit-reversed order
// in pass-2 layout.  Inputs are digits (|x| <= 2^15); no reduction is needed:
// a stage adds at most p(0.5 + 1.5|x| 2^-53) to the bound (mulmod above), so
// after 10 (11) stages every value is below 7.75p (8.98p) < 16p <= 2^53.
template <int LOGN, class Sync, class TW>
__device__ __forceinline__ void ntt_forward_sync_t(double (&x)[16], double* scr, int tl, const TW& tw, const PrimeF64& pr, Sync& sy) {
    fwd_pass_t<LOGN, 0>(x, tl, tw, pr);
    exchange<LOGN, 0, 1, 0>(x, scr, tl, sy);
    fwd_pass_t<LOGN, 1>(x, tl, tw, pr);
    exchange<LOGN, 1, 2, 1>(x, scr, tl, sy);
    fwd_pass_t<LOGN, 2>(x, tl, tw, pr);
}
template <int LOGN, class Sync>
__device__ __forceinline__ void ntt_forward_sync(double (&x)[16], double* scr, int tl, const double* __restrict__ tw, const PrimeF64& pr, Sync& sy) {
    ntt_forward_sync_t<LOGN>(x, scr, tl, TwFlat{tw}, pr, sy);
}
template <int LOGN, int TEAM_WAVES>
__device__ __forceinline__ void ntt_forward(double (&x)[16], double* scr, int tl, const double* __restrict__ tw, const PrimeF64& pr) {
    WorkgroupSync<TEAM_WAVES> sy;
    ntt_forward_sync<LOGN>(x, scr, tl, tw, pr, sy);
}

// The first two forward stages on digit inputs by table lookup.  Both
// stages of pass 0 use fixed twiddles: stage 0 pairs (r, r + 8) with w1 =
// tw[1]; stage 1 pairs (r, r + 4) with w2 = tw[2] for r < 4 and w3 = tw[3] for
// r in [8, 12).  For the digits a, b, c, d of registers g, g + 4, g + 8, g + 12
// the two stages give
//   x[g]      = a + w1 c + (w2 b + w2 w1 d)     x[g + 4]  = a + w1 c - (w2 b + w2 w1 d)
//   x[g + 8]  = a - w1 c + (w3 b - w3 w1 d)     x[g + 12] = a - w1 c - (w3 b - w3 w1 d)
// so with tables of w * d mod p for every digit d no product is computed.
// FirstStageTables holds them: [w1, w2, w2 w1, w3, w3 w1][256], indexed by the
// digit's byte (two's complement), each entry of magnitude <= p/2 + 1.  The
// outputs are below 1.5p + 2^7; a stage takes a bound b p to (1.09375 b + 0.5) p
// (the mulmod bound above with p < 2^49), so after log2(N) stages every value
// is below 8.7p (N = 1024) or 10p (N = 2048), against 7.75p / 8.98p with
// computed products: all < 16p, exact.
// With stage 0 alone from a table (fwd_first_stage) only [w1] is used.
constexpr int FIRST_TABLE_ENTRIES = 5 * 256;
__host__ __device__ constexpr int first_table_entries(int stages) { return stages == 2 ? FIRST_TABLE_ENTRIES : stages == 1 ? 256 : 1; }

// Fills `entries` table entries (all threads of the workgroup take part; the
// caller synchronises before use).  tw: the prime's forward twiddles.
__device__ __forceinline__ void build_first_stage_tables(double* tab, const double* tw, const PrimeF64& pr, int tid, int nthreads,
                                                         int entries = FIRST_TABLE_ENTRIES) {
    const double w1 = tw[1], w2 = tw[2], w3 = tw[3];
    const double w21 = reduce(mulmod(w2, w1, pr), pr), w31 = reduce(mulmod(w3, w1, pr), pr);
    for (int e = tid; e < entries; e += nthreads) {
        const int t = e >> 8;
        const double d = static_cast<double>(static_cast<int8_t>(static_cast<uint8_t>(e & 255)));
        const double w = t == 0 ? w1 : t == 1 ? w2 : t == 2 ? w21 : t == 3 ? w3 : w31;
        tab[e] = reduce(mulmod(d, w, pr), pr);
    }
}

// x = stages 0 and 1 of the forward transform of the digit polynomial whose
// signed-byte digits are byte `lvl` of dg[r] (pass-0 layout).
__device__ __forceinline__ void fwd_first_two_stages(double (&x)[16], const uint32_t (&dg)[16], int lvl, const double* tab) {
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const uint32_t ba = (dg[g] >> (8 * lvl)) & 255u, bb = (dg[g + 4] >> (8 * lvl)) & 255u;
        const uint32_t bc = (dg[g + 8] >> (8 * lvl)) & 255u, bd = (dg[g + 12] >> (8 * lvl)) & 255u;
        const double a = static_cast<double>(static_cast<int8_t>(static_cast<uint8_t>(ba)));
        const double t1c = tab[bc], t2b = tab[256 + bb], t21d = tab[512 + bd], t3b = tab[768 + bb], t31d = tab[1024 + bd];
        const double P = __dadd_rn(a, t1c), Q = __dadd_rn(a, -t1c);
        const double R = __dadd_rn(t2b, t21d), S = __dadd_rn(t3b, -t31d);
        x[g] = __dadd_rn(P, R);
        x[g + 4] = __dadd_rn(P, -R);
        x[g + 8] = __dadd_rn(Q, S);
        x[g + 12] = __dadd_rn(Q, -S);
    }
}

// Stage 0 alone by lookup: x[r] = a + w1 c, x[r + 8] = a - w1 c for the
// digits a, c of registers r, r + 8 (r < 8); tab[0..256) = w1 * d.
__device__ __forceinline__ void fwd_first_stage(double (&x)[16], const uint32_t (&dg)[16], int lvl, const double* tab) {
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const uint32_t ba = (dg[r] >> (8 * lvl)) & 255u, bc = (dg[r + 8] >> (8 * lvl)) & 255u;
        const double a = static_cast<double>(static_cast<int8_t>(static_cast<uint8_t>(ba)));
        const double t = tab[bc];
        x[r] = __dadd_rn(a, t);
        x[r + 8] = __dadd_rn(a, -t);
    }
}

// The forward transform after fwd_first_stage.
template <int LOGN, class Sync, class TW>
__device__ __forceinline__ void ntt_forward_from_stage1_t(double (&x)[16], double* scr, int tl, const TW& tw, const PrimeF64& pr, Sync& sy) {
    fwd_pass_t<LOGN, 0, 1>(x, tl, tw, pr);
    exchange<LOGN, 0, 1, 0>(x, scr, tl, sy);
    fwd_pass_t<LOGN, 1>(x, tl, tw, pr);
    exchange<LOGN, 1, 2, 1>(x, scr, tl, sy);
    fwd_pass_t<LOGN, 2>(x, tl, tw, pr);
}
template <int LOGN, class Sync>
__device__ __forceinline__ void ntt_forward_from_stage1(double (&x)[16], double* scr, int tl, const double* __restrict__ tw, const PrimeF64& pr,
                                                        Sync& sy) {
    ntt_forward_from_stage1_t<LOGN>(x, scr, tl, TwFlat{tw}, pr, sy);
}

// Stage 0 by lookup for wide digits (|d| <= H, H = B/2 up to 2^15): tab[d + H]
// = w1 d mod p, centred (|entry| <= p/2 + 1), w1 = tw[1] the stage-0 twiddle.
// x[r] = a + tab[c + H], x[r + 8] = a - tab[c + H] for the digits a, c of
// registers r, r + 8 (r < 8).  The values differ from the computed stage only
// by multiples of p, and stay within the same bounds (|x| <= p/2 + 1 + H).
__device__ __forceinline__ void build_stage0_table_wide(double* tab, int entries, int half, double w1, const PrimeF64& pr, int tid,
                                                        int nthreads) {
    for (int e = tid; e < entries; e += nthreads) tab[e] = reduce(mulmod(static_cast<double>(e - half), w1, pr), pr);
}

// The forward transform after fwd_first_two_stages.
template <int LOGN, class Sync, class TW>
__device__ __forceinline__ void ntt_forward_from_stage2_t(double (&x)[16], double* scr, int tl, const TW& tw, const PrimeF64& pr, Sync& sy) {
    fwd_pass_t<LOGN, 0, 2>(x, tl, tw, pr);
    exchange<LOGN, 0, 1, 0>(x, scr, tl, sy);
    fwd_pass_t<LOGN, 1>(x, tl, tw, pr);
    exchange<LOGN, 1, 2, 1>(x, scr, tl, sy);
    fwd_pass_t<LOGN, 2>(x, tl, tw, pr);
}
template <int LOGN, class Sync>
__device__ __forceinline__ void ntt_forward_from_stage2(double (&x)[16], double* scr, int tl, const double* __restrict__ tw, const PrimeF64& pr,
                                                        Sync& sy) {
    ntt_forward_from_stage2_t<LOGN>(x, scr, tl, TwFlat{tw}, pr, sy);
}

// Inverse negacyclic NTT without the 1/N factor (folded into the key):
// pass-2 layout (bit-reversed) -> pass-0 layout (natural).  Pass 2 begins
// with every value reduced, so its outputs are below 4p (at most three
// stages); pass 1 reduces only its first stage's sums (inv_pass_from_4p), and
// then its registers 0 and 1, the two whose worst case (4p and 10p) is not
// below 2.94p; pass 0 starts below 2.94p and again reduces only its first
// stage's sums.  Every value stays below 10p; the outputs are below 8.41p
// (centre16_after_inverse gives their exact centred values).
template <int LOGN, class Sync, class TW>
__device__ __forceinline__ void ntt_inverse_sync_t(double (&x)[16], double* scr, int tl, const TW& itw, const PrimeF64& pr, Sync& sy) {
    reduce16(x, pr);
    inv_pass_t<LOGN, 2>(x, tl, itw, pr);
    exchange<LOGN, 2, 1, 1>(x, scr, tl, sy);
    inv_pass_from_4p<LOGN, 1>(x, tl, itw, pr);
    x[0] = reduce(x[0], pr);
    x[1] = reduce(x[1], pr);
    exchange<LOGN, 1, 0, 0>(x, scr, tl, sy);
    inv_pass_from_4p<LOGN, 0>(x, tl, itw, pr);
}
template <int LOGN, class Sync>
__device__ __forceinline__ void ntt_inverse_sync(double (&x)[16], double* scr, int tl, const double* __restrict__ itw, const PrimeF64& pr, Sync& sy) {
    ntt_inverse_sync_t<LOGN>(x, scr, tl, TwFlat{itw}, pr, sy);
}
template <int LOGN, int TEAM_WAVES>
__device__ __forceinline__ void ntt_inverse(double (&x)[16], double* scr, int tl, const double* __restrict__ itw, const PrimeF64& pr) {
    WorkgroupSync<TEAM_WAVES> sy;
    ntt_inverse_sync<LOGN>(x, scr, tl, itw, pr, sy);
}

// ---------------------------------------------------------------------------
// Torus helpers.
template <int TBITS>
__host__ __device__ __forceinline__ int64_t sext_t(int64_t x) {
    return static_cast<int64_t>(static_cast<uint64_t>(x) << (64 - TBITS)) >> (64 - TBITS);
}

// Zero-mean balanced digits (gadget_decomposition.h:75-95): all `l` digits of
// one torus value, least significant level first; out[lvl] for lvl < l.
template <int TBITS>
__device__ __forceinline__ void digits_of(int64_t x, int l, int rb, int (&out)[4]) {
    const uint64_t mask = (TBITS >= 64) ? ~0ULL : ((1ULL << TBITS) - 1);
    const uint64_t u = static_cast<uint64_t>(x) & mask;
    const bool neg = (u >> (TBITS - 1)) & 1;
    const int dropped = TBITS - l * rb;
    const uint64_t bias = dropped > 0 ? (1ULL << (dropped - 1)) : 0;
    const uint64_t m = (neg ? ((0ULL - u) & mask) : u) + bias;
    const int B = 1 << rb;
    const int halfB = B >> 1;
    int carry = 0;
#pragma unroll
    for (int lvl = 3; lvl >= 0; lvl--) {
        if (lvl < l) {
            const int shift = TBITS - (lvl + 1) * rb;
            const int window = static_cast<int>((m >> shift) & static_cast<uint64_t>(B - 1));
            const int digit = window + carry;
            const int tie = lvl > 0 ? static_cast<int>((m >> (shift + 2 * rb - 1)) & 1) : 0;
            carry = (digit > halfB) | ((digit == halfB) & tie);
            const int bal = carry ? digit - B : digit;
            out[lvl] = neg ? -bal : bal;
        }
    }
}

// All l balanced digits of a TBITS-bit torus value for a gadget shape known
// only at run time, packed for the level loop: 8-bit fields when rb <= 7
// (levels 0-3 in w0, 4-7 in w1, so up to 8 levels), else 16-bit fields
// (levels 0-1 in w0, 2-3 in w1, up to 4 levels).  Same digits as digits_of.
template <int TBITS>
__device__ __forceinline__ void digits_runtime_packed(int64_t x, int l, int rb, uint32_t& w0, uint32_t& w1) {
    using U = typename std::conditional<TBITS == 32, uint32_t, uint64_t>::type;
    const U mask = (TBITS >= static_cast<int>(sizeof(U) * 8)) ? ~U{0} : static_cast<U>((uint64_t{1} << (TBITS % 64)) - 1);
    const U u = static_cast<U>(x) & mask;
    const bool neg = (u >> (TBITS - 1)) & 1;
    const int dropped = TBITS - l * rb;
    const U bias = dropped > 0 ? (U{1} << (dropped - 1)) : U{0};
    const U m = (neg ? ((U{0} - u) & mask) : u) + bias;
    const int B = 1 << rb, halfB = B >> 1;
    const bool p8 = rb <= 7;
    int carry = 0;
    w0 = 0;
    w1 = 0;
#pragma unroll
    for (int lvl = 7; lvl >= 0; lvl--) {
        if (lvl < l) {
            const int shift = TBITS - (lvl + 1) * rb;
            const int window = static_cast<int>((m >> shift) & static_cast<U>(B - 1));
            const int digit = window + carry;
            const int tie = lvl > 0 ? static_cast<int>((m >> (shift + 2 * rb - 1)) & 1) : 0;
            carry = (digit + tie) > halfB;
            const int bal = carry ? digit - B : digit;
            const uint32_t sd = static_cast<uint32_t>(neg ? -bal : bal);
            if (p8) {
                if (lvl < 4) w0 |= (sd & 0xFFu) << (8 * lvl);
                else w1 |= (sd & 0xFFu) << (8 * (lvl - 4));
            } else {
                if (lvl < 2) w0 |= (sd & 0xFFFFu) << (16 * lvl);
                else w1 |= (sd & 0xFFFFu) << (16 * (lvl - 2));
            }
        }
    }
}

// Digit `lvl` of digits_runtime_packed's words, sign-extended by a shift and a
// narrowing cast (not __builtin_amdgcn_sbfe: with a wave-uniform but
// non-constant offset hipcc has converted its result to double as unsigned).
__device__ __forceinline__ int digit_runtime(uint32_t w0, uint32_t w1, int lvl, int rb) {
    if (rb <= 7) return static_cast<int>(static_cast<int8_t>((lvl < 4 ? w0 : w1) >> (8 * (lvl & 3))));
    return static_cast<int>(static_cast<int16_t>((lvl < 2 ? w0 : w1) >> (16 * (lvl & 1))));
}

// Exact conversions of an integer-valued double: adding 1.5 * 2^52 keeps the
// sum inside one binade, where the mantissa field is the integer plus 2^51.
__device__ __forceinline__ int64_t f64_to_i64(double v) {  // |v| < 2^51
    return static_cast<int64_t>(__double_as_longlong(__dadd_rn(v, 0x1.8p52))) - 0x4338000000000000LL;
}
__device__ __forceinline__ uint32_t f64_low32(double v) {  // |v| < 2^51: v mod 2^32
    return static_cast<uint32_t>(__double_as_longlong(__dadd_rn(v, 0x1.8p52)));
}

// Packing of the balanced digits below: 8-bit fields (RB <= 7, where every
// digit in [-B/2, B/2] fits a signed byte) into w0, or 16-bit fields into w0
// (levels 0, 1) and w1 (levels 2, 3).
template <int RB>
__host__ __device__ __forceinline__ void pack_digit(int lvl, int sd, uint32_t& w0, uint32_t& w1) {
    if constexpr (RB <= 7) {
        w0 |= (static_cast<uint32_t>(sd) & 0xFFu) << (8 * lvl);
    } else {
        if (lvl < 2) w0 |= (static_cast<uint32_t>(sd) & 0xFFFFu) << (16 * lvl);
        else w1 |= (static_cast<uint32_t>(sd) & 0xFFFFu) << (16 * (lvl - 2));
    }
}

// The balanced digits of a 32-bit torus value with the carry chain folded into
// one addition (same digits as digits_packed<32, L, RB>).  With the windows
// w_l of m (level 0 most significant), the tie bit t_l (top bit of window
// l - 1, zero for level 0) and the carry c_l into window l, the rule "carry
// out when w_l + c_l + t_l > B/2" is the binary carry out of
// w_l + c_l + (B/2 - 1) + t_l.  So m + K + T, with K = (B/2 - 1) in every
// window and T = the tie bits moved down to their window's base, carries
// exactly like the reference's loop, and digit l = window l of the sum
// - (B/2 - 1) - t_l (checked against digits_of over all 2^32 inputs for
// 2^7 x 3, 2^10 x 3, 2^8 x 2 and 2^8 x 4).
template <int L, int RB>
struct DigitsAdd32 {
    static constexpr uint32_t window_constant() {
        uint32_t k = 0;
        for (int l = 0; l < L; l++) k += static_cast<uint32_t>((1 << (RB - 1)) - 1) << (32 - (l + 1) * RB);
        return k;
    }
    static constexpr uint32_t tie_mask() {
        uint32_t t = 0;
        for (int l = 1; l < L; l++) t |= 1u << (32 - (l + 1) * RB);
        return t;
    }
};

// (The digit functions from here to digits_packed_wide are also compiled for
// the host, where tests compare them with the oracle's decomposition.)
// The carry-chain core of digits_packed_add32: m holds the biased magnitude
// with the gadget's L windows at bits 32 - (l + 1) RB, s is all ones for a
// negative value.
template <int L, int RB>
__host__ __device__ __forceinline__ void digits_add32_core(uint32_t m, uint32_t s, uint32_t& w0, uint32_t& w1) {
    constexpr int B = 1 << RB, HB = B >> 1;
    constexpr uint32_t K = DigitsAdd32<L, RB>::window_constant();
    constexpr uint32_t TM = DigitsAdd32<L, RB>::tie_mask();
    const uint32_t t = (m >> (2 * RB - 1)) & TM;
    const uint32_t sum = m + K + t;
    const int sneg = static_cast<int>(s);
    w0 = 0;
    w1 = 0;
#pragma unroll
    for (int lvl = 0; lvl < L; lvl++) {
        const int shift = 32 - (lvl + 1) * RB;
        const int tl = lvl > 0 ? static_cast<int>((t >> shift) & 1u) : 0;
        const int bal = static_cast<int>((sum >> shift) & static_cast<uint32_t>(B - 1)) - (HB - 1) - tl;
        pack_digit<RB>(lvl, (bal ^ sneg) - sneg, w0, w1);
    }
}

template <int L, int RB>
__host__ __device__ __forceinline__ void digits_packed_add32(uint32_t u, uint32_t& w0, uint32_t& w1) {
    constexpr int DROP = 32 - L * RB;
    static_assert(DROP >= 0, "gadget wider than the torus");
    constexpr uint32_t BIAS = DROP > 0 ? (1u << (DROP - 1)) : 0u;
    const uint32_t s = static_cast<uint32_t>(static_cast<int32_t>(u) >> 31);  // all ones when negative
    digits_add32_core<L, RB>(((u ^ s) - s) + BIAS, s, w0, w1);
}

// The same digits of a TBITS-bit torus value (TBITS > 32, L RB <= 32): the
// bias is added to the magnitude at full width, and only its top 32 bits,
// which hold every window and tie bit, go through the 32-bit carry chain
// (same digits as digits_packed<TBITS, L, RB>).
template <int TBITS, int L, int RB>
__host__ __device__ __forceinline__ void digits_packed_wide(int64_t x, uint32_t& w0, uint32_t& w1) {
    constexpr int DROP = TBITS - L * RB;
    static_assert(TBITS > 32 && TBITS <= 63 && L * RB <= 32 && DROP > 0, "wide torus, gadget within 32 bits");
    const int64_t xs = sext_t<TBITS>(x);
    const uint32_t s = static_cast<uint32_t>(static_cast<int32_t>(static_cast<uint64_t>(xs) >> 32) >> 31);
    const uint64_t mag = static_cast<uint64_t>(xs < 0 ? -xs : xs) + (uint64_t{1} << (DROP - 1));
    digits_add32_core<L, RB>(static_cast<uint32_t>(mag >> (TBITS - 32)), s, w0, w1);
}

// All L balanced digits (gadget_decomposition.h:75-95) of a value of a
// TBITS-bit torus, with the gadget shape known at compile time, packed as
// pack_digit describes.
template <int TBITS, int L, int RB>
__device__ __forceinline__ void digits_packed(int64_t x, uint32_t& w0, uint32_t& w1) {
    using U = typename std::conditional<TBITS == 32, uint32_t, uint64_t>::type;
    constexpr int DROP = TBITS - L * RB;
    constexpr U BIAS = DROP > 0 ? (U{1} << (DROP - 1)) : U{0};
    constexpr U MASK = (TBITS >= static_cast<int>(sizeof(U) * 8)) ? ~U{0} : static_cast<U>((uint64_t{1} << (TBITS % 64)) - 1);
    constexpr int B = 1 << RB, HB = B >> 1;
    const U u = static_cast<U>(x) & MASK;
    const bool neg = (u >> (TBITS - 1)) & 1;
    const U m = (neg ? ((U{0} - u) & MASK) : u) + BIAS;
    const int sneg = neg ? -1 : 0;
    int carry = 0;
    w0 = 0;
    w1 = 0;
#pragma unroll
    for (int lvl = L - 1; lvl >= 0; lvl--) {
        const int shift = TBITS - (lvl + 1) * RB;
        const int window = static_cast<int>((m >> shift) & static_cast<U>(B - 1));
        const int digit = window + carry;
        const int tie = lvl > 0 ? static_cast<int>((m >> (shift + 2 * RB - 1)) & 1) : 0;
        // digit > B/2, or digit == B/2 with the tie bit set
        carry = (digit + tie) > HB;
        const int bal = carry ? digit - B : digit;
        pack_digit<RB>(lvl, (bal ^ sneg) - sneg, w0, w1);
    }
}

}  // namespace ygpu
