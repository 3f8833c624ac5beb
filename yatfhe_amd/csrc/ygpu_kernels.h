// ygpu_kernels.h -- kernel configurations, argument blocks and launch entry points.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>
#include <algorithm>
#include "ygpu_device.h"

namespace ygpu {

// Blind-rotation kernel configuration (k = 1, two teams per ciphertext).
//   TBITS  torus width (32 or 56)
//   LOGN   log2 ring degree (10 or 11)
//   NP     transform primes (1: the reference prime itself; 2: CRT pair)
//   CPW    ciphertexts per workgroup
//   PACK8  digits packed 8 bits each (radix_bits <= 8), else 16 bits
//   L, RB  gadget levels and base log2 when fixed at compile time (0: runtime)
//   TWL    twiddles in LDS (1), global memory through L1 (0), or automatic (-1);
//          2: lane-major LDS tables for transform passes 1 and 2 (TwLane;
//          TORUS32 N = 1024 with TAB = 2)
//   TAB    forward stages taken from digit tables (0, 1 or 2; fixed 8-bit shapes)
template <int TBITS_, int LOGN_, int NP_, int CPW_, bool PACK8_, int L_ = 0, int RB_ = 0, int TWL_ = -1, int TAB_ = 0, bool LAZY_ = false>
struct BrCfg {
    static constexpr int TAB = TAB_;
    // the lazy-pipe step: digits of the accumulator itself (no rotation), the
    // accumulator replaced by the product, steps with abar = 0 left unchanged
    static constexpr bool LAZY = LAZY_;
    static constexpr int L = L_;
    static constexpr int RB = RB_;
    static constexpr int TWL = TWL_;
    static constexpr int TBITS = TBITS_;
    using TorusT = typename std::conditional<TBITS == 32, int32_t, int64_t>::type;
    static constexpr int LOGN = LOGN_;
    static constexpr int N = 1 << LOGN;
    static constexpr int T = N / 16;
    static constexpr int TEAM_WAVES = T / 64;
    static constexpr int NP = NP_;
    static constexpr int CPW = CPW_;
    static constexpr bool PACK8 = PACK8_;
    // teams per ciphertext: one per accumulator row, times one per transform prime
    static constexpr int TEAMS = 2 * NP;
    static constexpr int THREADS = TEAMS * T * CPW;
    // Waves per SIMD the register allocation must allow: two (256 VGPRs).
    static constexpr int MIN_WAVES = 2;
    // twiddles in LDS when they fit beside the per-ciphertext state
    static constexpr bool TW_LDS = TWL_ < 0 ? (NP * 2 * N * 8 <= 16384) : (TWL_ >= 1);
    static constexpr bool TW_LANE = TWL_ == 2;
};

// Garner reconstruction constants for NP == 2.
struct CrtConst {
    double p0;         // first prime (double, exact)
    double inv01;      // p0^{-1} mod p1, centred
    int64_t p0i;       // first prime (integer)
    uint64_t q;        // target modulus (the reference NTT prime)
    double qd;         // q as double (rounded)
    double qinv;       // 1/q (rounded)
    int64_t qhalf;     // (q - 1) / 2
};

struct BrArgs {
    int n, level, radix_bits, count;
    int prescaled;                    // lwe_in already holds abar (mod 2N) values
    const void* acc_in;               // optional initial accumulators Torus[count][2][N] (replaces v)
    const int32_t* lwe_in;            // [count][n+1]
    const void* v;                    // Torus[N] (v_stride 0) or Torus[count][N]
    int v_stride;
    void* acc_out;                    // Torus[count][2][N]
    const double* const* bsk_table;   // device array: BSK base per key slot
    const int32_t* ct_slot;           // per-ciphertext key slot (null: slot 0)
    const double* tw;                 // [NP][2][N]: forward then inverse powers
    PrimeF64 prime[2];
    CrtConst crt;
    // Failure report of the bounded team barriers: a workgroup whose barrier
    // wait ran out writes 1 here (host-visible memory) and poisons its
    // accumulators; barrier_polls bounds every wait (0: every team barrier
    // fails at once without waiting, which exercises the report path).
    uint32_t* status;
    int barrier_polls;
    int lwe_stride;  // ints between ciphertexts of lwe_in (0: n + 1)
};

// The lazy-pipe key's per-ciphertext work (k_lazy_prepare, k_lazy_derive).
struct LazyArgs {
    int count, n;
    int level;               // R: gadget levels of the rows (param.lApprox)
    int dlevel;              // D: stored digit levels (param.l), R <= D <= 8
    int radix_bits;
    int prescaled;           // lwe_in already holds abar (mod 2N) values
    const int32_t* lwe_in;   // [count][n+1]
    const int32_t* v;        // Torus[N] (v_stride 0) or Torus[count][N]
    int v_stride;
    const int32_t* first;    // bskFirst[0] [k+1][N] (Alt: the derived first component, per v)
    int first_stride;        // ints between ciphertexts' first components (0: one for all)
    int32_t* acc;            // [count][k+1][N]
    int32_t* avals;          // [count][n-1]: the rotation amounts of steps 1 .. n-1
    const int8_t* da;        // [n-1][R][D][k][N]
    const int8_t* db;        // [n-1][R][D][N]
    const double* s2e;       // s^2 TRLEV in the engine form [D][k+1][N]
    double g[8];             // 2^(32 - (l+1) b) N^-1 mod p, centred, l < D
    double* ws;              // rows [count][ws_steps][R][k+1][k+1][N]
    int s0, steps, ws_steps; // the window: components s0 .. s0 + steps - 1
    const double** row_tab;  // [count]: each ciphertext's rows in ws (k_lazy_prepare fills it)
    int32_t* row_slot;       // [count]: c -> c
    // the Alt form (BootstrappingKeyMPLazyPipeAlt): torus rows instead of digits
    const int32_t* rows_t;   // [n][R][k+1][N]: cPrime[lvl] of every s_i, torus
    const double* row0e;     // component 0's rows in the engine form [R][k+1][N]
    int32_t* first_out;      // [count_v][k+1][N]: k_lazy_alt_first's output
    double ninv;             // N^-1 mod p, centred
    const double* tw;        // [2][N]
    PrimeF64 prime;
};
bool lazy_supported(int tbits, int logn, int k);
hipError_t launch_lazy_prepare(const LazyArgs& a, hipStream_t s);
hipError_t launch_lazy_derive(const LazyArgs& a, hipStream_t s);
hipError_t launch_lazy_alt_first(const LazyArgs& a, int count_v, hipStream_t s);
hipError_t launch_lazy_alt_derive(const LazyArgs& a, hipStream_t s);
// The batched blind rotation in its lazy form (no rotation, the accumulator
// replaced by the product, steps with a rotation amount of 0 skipped), TORUS32
// N = 1024: gadget 2^7 x 3 specialised, other gadgets up to 2^7 and 8 levels.
hipError_t launch_blind_rotate_lazy(const BrArgs& a, hipStream_t s);

constexpr int KS_THREADS = 256;
constexpr int KS_TC = 64;    // ciphertexts per keyswitch tile
constexpr int KS_TK = 64;    // output columns per keyswitch tile
constexpr int KS_MAX_LEVEL = 16;
constexpr int KS_MAX_N = 1023;

struct KsArgs {
    int n, N, k, count, ks_level, ks_radix_bits, ks_width_bits;
    const void* acc;                   // Torus[count][k+1][N]
    const int32_t* const* ksk_table;   // device array: KSK byte planes per key slot
    const int32_t* ct_slot;
    int32_t* lwe_out;                  // [count][n+1]
    int32_t* partial;                  // [count][n+1] (slices > 1)
    uint32_t* digits;                  // [groups][count]: the keyswitch digits of four rows per word
    int extract_index;                 // coefficient the sample is extracted at (param.driftPhase), in [0, N)
};

// Words of the keyswitch workspace of `count` ciphertexts: the sliced partial
// sums ([count][n+1]) and the packed digits ([(k N t + 3) / 4][count]).
inline size_t ks_workspace_words(int n, int k, int N, int ks_level, int count) {
    const size_t partial = static_cast<size_t>(count) * (n + 1);
    const size_t groups = (static_cast<size_t>(k) * N * ks_level + 3) / 4;
    return (partial + 63) / 64 * 64 + groups * static_cast<size_t>(count);
}

// ---------------------------------------------------------------- helpers
// Workgroup b runs on XCD b % 8 (round-robin dispatch); map logical work so each
// XCD gets a contiguous range (same keys share that XCD's L2).  The grid is a
// multiple of 8.
__device__ __forceinline__ int map_workgroup(int b, int grid) {
    const int per = grid / 8;
    return (b % 8) * per + b / 8;
}

// modSwitchFromTorusGeneral(x, 2N, 2^32) (numeric.cpp:189-197): round half away
// from zero of x / 2^(32 - log2(2N)), then sign extension mod 2N.
__device__ __forceinline__ int mod_switch_2n(int32_t x, int log2_2n) {
    const int sb = 32 - log2_2n;
    const int64_t h = int64_t{1} << (sb - 1);
    const int64_t v = x;
    const int64_t r = v >= 0 ? ((v + h) >> sb) : -((-v + h) >> sb);
    return static_cast<int>(static_cast<int64_t>(static_cast<uint64_t>(r) << (64 - log2_2n)) >> (64 - log2_2n));
}

// validateRotator (polynomial.cpp:22-29).
__device__ __forceinline__ void validate_rotator(int a, int N, int& aT, int& wrap) {
    int t = a % (2 * N);
    if (t < 0) t += 2 * N;
    wrap = t < N ? 1 : -1;
    aT = t < N ? t : t - N;
}

// Torus (2^TBITS) -> LWE (2^32) modswitch of extractTlweFromTrlwe
// (trlwe.cpp:247,250 via numeric.cpp:189-197).
template <int TBITS>
__device__ __forceinline__ int32_t torus_to_lwe32(int64_t x) {
    if constexpr (TBITS == 32) {
        return static_cast<int32_t>(x);
    } else {
        constexpr int sb = TBITS - 32;
        const int64_t h = int64_t{1} << (sb - 1);
        const int64_t r = x >= 0 ? ((x + h) >> sb) : -((-x + h) >> sb);
        return static_cast<int32_t>(static_cast<uint32_t>(static_cast<uint64_t>(r)));
    }
}

// Keyswitch digits: zeroMeanDigits(x, t, b_ks, width) (gadget_decomposition.h:75-95
// via signedGadgetDecompositionKs, gadget_decomposition.cpp:53-56).
__device__ __forceinline__ void ks_digits(int32_t x, int t, int b, int width, int8_t* out) {
    const uint64_t mask = width >= 64 ? ~0ULL : ((1ULL << width) - 1);
    const uint64_t u = static_cast<uint64_t>(static_cast<int64_t>(x)) & mask;
    const bool neg = (u >> (width - 1)) & 1;
    const int dropped = width - t * b;
    const uint64_t bias = dropped > 0 ? (1ULL << (dropped - 1)) : 0;
    const uint64_t m = (neg ? ((0ULL - u) & mask) : u) + bias;
    const int B = 1 << b, halfB = B >> 1;
    int carry = 0;
    for (int lvl = t - 1; lvl >= 0; lvl--) {
        const int shift = width - (lvl + 1) * b;
        const int window = static_cast<int>((m >> shift) & static_cast<uint64_t>(B - 1));
        const int digit = window + carry;
        const int tie = lvl > 0 ? static_cast<int>((m >> (shift + 2 * b - 1)) & 1) : 0;
        carry = (digit > halfB) | ((digit == halfB) & tie);
        const int bal = carry ? digit - B : digit;
        out[lvl] = static_cast<int8_t>(neg ? -bal : bal);
    }
}

// Exact value mod q, centred in [-(q-1)/2, (q-1)/2], of the integer C with
// C = c0 mod p0, C = c1 mod p1 and |C| < p0 p1 / 2 (Garner), from the centred
// residues c0, c1.  C = c0 + p0 * y with y = (c1 - c0) p0^{-1} mod p1 centred;
// the quotient t = C / q is estimated in binary64 (off by at most one), so the
// remainder R = C - t q lies in (-3q/2, 3q/2), inside int64 for q < 2^61: it is
// formed exactly by wrapping 64-bit arithmetic (C itself may exceed 64 bits;
// its low 64 bits suffice) and corrected into the centred range.  y takes
// one reduction: a product from mulmod is below 2 p1 in magnitude (inputs
// below 16p), and one reduction is exact below 4p (centre16_after_inverse).
__device__ __forceinline__ int64_t crt_mod_q(double c0, double c1, const PrimeF64& P1, const CrtConst& k) {
    const double y = reduce(mulmod(__dadd_rn(c1, -c0), k.inv01, P1), P1);
    const double Cd = __fma_rn(k.p0, y, c0);
    const int64_t qt = static_cast<int64_t>(__builtin_rint(__dmul_rn(Cd, k.qinv)));
    const uint64_t lowC = static_cast<uint64_t>(k.p0i) * static_cast<uint64_t>(f64_to_i64(y)) + static_cast<uint64_t>(f64_to_i64(c0));
    int64_t R = static_cast<int64_t>(lowC - static_cast<uint64_t>(qt) * k.q);
    if (R > k.qhalf) R -= static_cast<int64_t>(k.q);
    if (R < -k.qhalf) R += static_cast<int64_t>(k.q);
    return R;
}

// Tile choice of the blind-rotation launch: ciphertexts per workgroup and
// where the transform twiddles live.
struct BrTile {
    int cpw = 4;
    int twl = 1;
    int lat = 0;  // the latency form: one wavefront per (row, gadget level)
    int tab = 0;  // forward stages from digit tables (0: none; 1: stage 0; 2: stages 0 and 1)
};
constexpr int MAX_TILES = 32;
hipError_t launch_blind_rotate(int tbits, int logn, const BrTile& tile, int radix_bits, const BrArgs& a, hipStream_t s);
// The latency form of the TORUS32 N = 1024 gadget-2^7x3 blind rotation (one
// ciphertext per workgroup, a wavefront per row and level), in its own
// translation unit.
bool blind_rotate_lat_supported(int tbits, int logn, int level, int radix_bits);
hipError_t launch_blind_rotate_lat(const BrArgs& a, hipStream_t s);
// The two-prime (CRT) blind rotation, in its own translation unit.
hipError_t launch_blind_rotate_crt(int logn, const BrTile& tile, int radix_bits, int level, const BrArgs& a, hipStream_t s);
// Tiles instantiated for a parameter set (the autotuner's candidates).
int blind_rotate_tiles(int tbits, int logn, int level, int radix_bits, BrTile* out, int max);
hipError_t launch_extract_ks(int tbits, const KsArgs& a, int slices, hipStream_t s);
hipError_t launch_extract(int tbits, const KsArgs& a, int32_t* out, hipStream_t s);
hipError_t launch_modmul_peak(int kind, const PrimeF64& pr, uint64_t q, int blocks, int iters, void* sink, hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
hipError_t launch_copy_words(const void* src, void* dst, size_t words, hipStream_t s);
constexpr int MODMUL_CHAINS = 8;

// Device-side key preparation: bootstrapping-key residues mod q, bit-packed
// `bits` per value (polynomial after polynomial, natural coefficient order),
// into the engine's BSK form: per transform prime, the forward transform of
// the centred lift times N^-1, centred, as doubles in last-pass order.
struct BskExpandArgs {
    const uint64_t* packed;
    int bits;
    uint64_t q;
    int polys;
    int np;
    double* out;
    const double* tw;       // per prime: forward twiddles at np*2N
    PrimeF64 prime[2];
    uint64_t p[2];
    double ninv[2];         // N^-1 mod p, centred
};
hipError_t launch_bsk_expand(int logn, const BskExpandArgs& a, hipStream_t s);

// Derivation of the full key from a WWL+24 succinct key: the packed body-digit
// rows ([i][lvl][col][N] residues, `bits` per value) and the TRLEV of s^2 in
// engine form ([l2][col][prime][N]) -> the engine-form BSK [i][lvl][row][col][prime][N].
struct BskDeriveArgs {
    const uint64_t* half;
    const double* s2;
    int bits;
    uint64_t q;
    int n, level, s2_level, radix_bits;
    int np;
    double* out;
    const double* tw;       // per prime: forward then inverse twiddles at np*2N
    PrimeF64 prime[2];
    uint64_t p[2];
    double ninv[2];
    CrtConst crt;
};
hipError_t launch_bsk_derive(int logn, int tbits, const BskDeriveArgs& a, hipStream_t s);

}  // namespace ygpu
