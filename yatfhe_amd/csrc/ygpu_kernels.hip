// ygpu_kernels.hip -- CDNA4 kernels of the bootstrapping path.
//
//  k_blind_rotate   rescale to 2N, accumulator init, and all n CMUX steps of the
//                   blind rotation in one persistent launch; CPW ciphertexts per
//                   workgroup, two one-wavefront teams per ciphertext, the
//                   accumulator resident in registers for all n steps.
//  k_extract_ks     sample extract (index 0) fused into the LWE keyswitch, a
//                   GEMM-shaped integer reduction over the KSK, tiled over
//                   ciphertexts so each KSK row is read once per tile.
//  k_extract        the extracted LWE alone (for callers that want it).
//  k_ks_finalize    negation / body of the sliced keyswitch.
#include "ygpu_kernels.h"
#include <limits>

namespace ygpu {

// ============================================================ blind rotation
// One CMUX step (cmux.cpp:16-21 / blind_rotate.cpp:1560-1562), per ciphertext:
//   tmp = (X^{a_i} - 1) acc                      rotateTrlweMinusOne
//   tmp = ExtProd(BSK_i, tmp)                    externalProductTrgswMPNttImpl
//   acc += tmp                                   accumulateTrlwe
// ExtProd: digits of each row (decomposeRow), forward NTT of every digit
// polynomial, multiply-accumulate against the NTT-form BSK row, inverse NTT of
// each output column, centre mod q, reduce to the torus (ntt_hexl.cpp:153-175).
//
// Work split: a workgroup holds CPW ciphertexts, each served by two teams of
// T = N/16 lanes (T = 64: one wavefront).  Team r transforms the digit
// polynomials of accumulator row r (d = lvl*2 + r, lvl < level) and owns output
// column r, so it reads and writes only accumulator row r; the two teams swap
// their partial sums for the other column once per step.  The workgroup's
// ciphertexts run in lockstep (the swap is a workgroup barrier), so the BSK row
// loads of one step hit in L1 after the first ciphertext; the transform
// twiddles sit in LDS, shared by the workgroup.
template <class C>
__global__ void __launch_bounds__(C::THREADS, C::MIN_WAVES) k_blind_rotate(BrArgs a) {
    static_assert(C::NP == 1, "two-prime moduli use k_blind_rotate_crt");
    using TorusT = typename C::TorusT;
    constexpr int LOGN = C::LOGN, N = C::N, T = C::T, K1 = 2, NP = C::NP, TW = C::TEAM_WAVES, CPW = C::CPW;
    constexpr int SCR = Scratch<LOGN>::DOUBLES;
    static_assert(C::TBITS != 32 || 2 * N * sizeof(int32_t) <= SCR * sizeof(double), "the rotation stage [acc, -acc] fits the scratch");
    constexpr int CT_THREADS = 2 * T;
    // lane-major twiddle tables (TwLane): passes 1 and 2 read them; the flat
    // table then only holds pass 0's sixteen twiddles of each direction
    constexpr bool LANE = C::TW_LANE;
    static_assert(!LANE || (LOGN == 10 && NP == 1 && C::TAB == 2), "lane tables: TORUS32 N = 1024, TAB = 2");
    constexpr int TWN = LANE ? 32 : C::TW_LDS ? NP * 2 * N : 1;

    __shared__ double tws[TWN];
    __shared__ double twLane[LANE ? 2 * TW_LANE_PER_DIR : 1];
    __shared__ double scratch[CPW][2 * SCR];
    // the specialised 8-bit-digit shapes may run the first one or two forward
    // stages by table lookup (fwd_first_stage / fwd_first_two_stages): fewer
    // products for more LDS reads; the tile chooses
    constexpr int TAB = (C::L > 0 && C::RB <= 7 && NP == 1) ? C::TAB : 0;
    constexpr int TAB_ENTRIES = first_table_entries(TAB);
    __shared__ double firstTab[TAB_ENTRIES];
    // NP == 2: the first prime's residues of each team's column wait here while
    // the second prime is processed (registers are the binding resource)
    constexpr int RESN = NP > 1 ? 2 * N : 1;
    __shared__ double resbuf[CPW][RESN];
    // One-wave teams (N = 1024, one prime) hand their partial sums over with
    // one workgroup barrier per step instead of two: a team counts its reads
    // of the partner's values (readDone), and the partner waits for that
    // count only where it next writes its scratch, an inverse transform pass
    // later, by which time the reads have long completed.
    constexpr bool HANDOFF = TW == 1 && NP == 1;
    __shared__ uint32_t readDone[HANDOFF ? CPW * 2 + 1 : 1];

    const int grp = map_workgroup(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
    const int first = grp * CPW;
    if (first >= a.count) return;  // whole workgroup
    const int lc = static_cast<int>(threadIdx.x) / CT_THREADS;
    const int ctid = static_cast<int>(threadIdx.x) % CT_THREADS;
    const int tm = ctid / T;
    const int tl = ctid % T;
    // tail slots of the last workgroup repeat its last ciphertext and write nothing
    const bool live = first + lc < a.count;
    const int ct = live ? first + lc : a.count - 1;
    double* scr = scratch[lc] + tm * SCR;
    // the team's accumulator row lives in registers at its pass-0 positions
    // j = tl + T*r; it is staged through the team's scratch for the rotation
    TorusT* stage = reinterpret_cast<TorusT*>(scr);

    const int n = a.n;
    // the gadget depth: compile-time for the specialised shapes (the level
    // loop below is then unrolled)
    const int L = C::L > 0 ? C::L : a.level;
    const int rb = a.radix_bits;
    const int slot = a.ct_slot ? a.ct_slot[ct] : 0;
    const double* __restrict__ bsk = a.bsk_table[slot];
    const int32_t* __restrict__ lwe = a.lwe_in + static_cast<size_t>(ct) * (a.lwe_stride > 0 ? a.lwe_stride : n + 1);
    const TorusT* __restrict__ v = static_cast<const TorusT*>(a.v) + static_cast<size_t>(a.v_stride) * ct;

    if constexpr (LANE) {
        for (int j = static_cast<int>(threadIdx.x); j < 32; j += C::THREADS) tws[j] = a.tw[(j >> 4) * N + (j & 15)];
        build_tw_lane(twLane, a.tw, static_cast<int>(threadIdx.x), C::THREADS);
        build_tw_lane(twLane + TW_LANE_PER_DIR, a.tw + N, static_cast<int>(threadIdx.x), C::THREADS);
    } else if constexpr (C::TW_LDS) {
        for (int j = static_cast<int>(threadIdx.x); j < TWN; j += C::THREADS) tws[j] = a.tw[j];
    }
    if constexpr (TAB > 0) build_first_stage_tables(firstTab, a.tw, a.prime[0], static_cast<int>(threadIdx.x), C::THREADS, TAB_ENTRIES);
    // the hand-over counts, then the failure flag of their bounded waits
    if constexpr (HANDOFF) {
        if (threadIdx.x < CPW * 2 + 1) readDone[threadIdx.x] = 0u;
    }
    uint32_t* const handFail = &readDone[HANDOFF ? CPW * 2 : 0];
    // rescaleTlweToNewMod (tlwe.cpp:75-81) of the body, then
    // genNoiselessTrlweSample (trlwe.cpp:175-182): acc = (0, X^{-bbar} v).
    TorusT accr[16];
    if (a.acc_in) {
        // caller-supplied accumulator (blindRotateMP21Ntt's in/out argument)
        const TorusT* ai = static_cast<const TorusT*>(a.acc_in) + static_cast<size_t>(ct) * K1 * N + tm * N;
#pragma unroll
        for (int r = 0; r < 16; r++) accr[r] = ai[tl + T * r];
    } else {
        const int bbar = a.prescaled ? lwe[n] : mod_switch_2n(lwe[n], LOGN + 1);
        int aT, wrap;
        validate_rotator(-bbar, N, aT, wrap);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int j = tl + T * r;
            const int src = j - aT;
            const int64_t val = src < 0 ? -static_cast<int64_t>(v[src + N]) * wrap : static_cast<int64_t>(v[src]) * wrap;
            accr[r] = tm == 0 ? TorusT(0) : static_cast<TorusT>(sext_t<C::TBITS>(val));
        }
    }
    __syncthreads();
    const double* twBase = C::TW_LDS ? static_cast<const double*>(tws) : a.tw;
    // pass 0's twiddles of both directions stay in (scalar) registers for the
    // whole launch instead of being read from LDS at every step
    constexpr bool P0REG = NP == 1 && C::TW_LDS;
    TwPass0Reg<TwFlat> fw0{}, iw0{};
    TwPass0Reg<TwLane> fl0{}, il0{};
    if constexpr (P0REG) {
        const double* __restrict__ g = a.tw;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            fw0.f[e] = g[e];
            iw0.f[e] = g[N + e];
        }
        fw0.base = TwFlat{twBase};
        iw0.base = TwFlat{twBase + N};
        if constexpr (LANE) {
#pragma unroll
            for (int e = 0; e < 16; e++) {
                fl0.f[e] = fw0.f[e];
                il0.f[e] = iw0.f[e];
            }
            fl0.base = TwLane{tws, twLane};
            il0.base = TwLane{tws + 16, twLane + TW_LANE_PER_DIR};
        }
    }

    const size_t stepStride = static_cast<size_t>(L) * K1 * K1 * NP * N;  // doubles per BSK row
    // the next step's rotation amount is loaded a step ahead, so no step
    // starts with a memory round trip
    int32_t rawNext = lwe[0];
    for (int i = 0; i < n; i++) {
        // (blind_rotate.cpp:1557-1559 skips a_i == 0; ExtProd(0) = 0 exactly,
        // so the lockstep workgroup simply computes it.)
        const int32_t raw = rawNext;
        if (i + 1 < n) rawNext = lwe[i + 1];
        const int abar = a.prescaled ? raw : mod_switch_2n(raw, LOGN + 1);
        int aT, wrap;
        validate_rotator(abar, N, aT, wrap);

        // (X^a - 1) acc_r at this lane's pass-0 positions (rotateTorusPolynomialMinusOne,
        // polynomial.cpp:213-238) and its digits, packed 8 or 16 bits per digit.
        uint32_t dg0[16], dg1[16];
        if constexpr (C::LAZY) {
            // the lazy-pipe step decomposes the accumulator itself
            // (externalProductTrgswMPNttInPlace, trgsw.cpp:774-776)
            static_assert(C::TBITS == 32, "lazy form: TORUS32");
#pragma unroll
            for (int r = 0; r < 16; r++) {
                if constexpr (C::L > 0) digits_packed_add32<C::L, C::RB>(static_cast<uint32_t>(accr[r]), dg0[r], dg1[r]);
                else digits_runtime_packed<32>(static_cast<int32_t>(accr[r]), L, rb, dg0[r], dg1[r]);
            }
        }
        if constexpr (!C::LAZY && C::TBITS == 32) {
        // the row and its negation side by side, G = [acc, -acc]: then
        // (X^t acc)[j] = G[(j - t) mod 2N] for the full rotation amount
        // t = abar mod 2N, with no per-position sign or wrap test
        team_sync<TW>();
#pragma unroll
        for (int r = 0; r < 16; r++) {
            stage[tl + T * r] = accr[r];
            stage[N + tl + T * r] = static_cast<TorusT>(0u - static_cast<uint32_t>(accr[r]));
        }
        team_sync<TW>();
        const int base = tl - (abar & (2 * N - 1));
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const uint32_t av = static_cast<uint32_t>(stage[(base + T * r) & (2 * N - 1)]);
            const uint32_t rot = av - static_cast<uint32_t>(accr[r]);
            if constexpr (C::L > 0) {
                // the carry chain as one 32-bit addition (2.3 % per
                // launch over the level loop of digits_packed)
                digits_packed_add32<C::L, C::RB>(rot, dg0[r], dg1[r]);
            } else {
                digits_runtime_packed<32>(static_cast<int32_t>(rot), L, rb, dg0[r], dg1[r]);
            }
        }
        } else if constexpr (!C::LAZY) {
        team_sync<TW>();
#pragma unroll
        for (int r = 0; r < 16; r++) stage[tl + T * r] = accr[r];
        team_sync<TW>();
        // negate the rotated source where j < aT (wrap-around), times the
        // rotator's own sign
        const bool flip = wrap < 0;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int j = tl + T * r;
            const int src = (j - aT) & (N - 1);
            const bool negs = (j < aT) != flip;
            {
                const int64_t av = static_cast<int64_t>(stage[src]);
                const int64_t rot = sext_t<C::TBITS>((negs ? -av : av) - static_cast<int64_t>(accr[r]));
                if constexpr (C::L > 0) {
                    digits_packed<C::TBITS, C::L, C::RB>(rot, dg0[r], dg1[r]);
                } else {
                    digits_runtime_packed<C::TBITS>(rot, L, rb, dg0[r], dg1[r]);
                }
            }
        }
        }

        const double* __restrict__ bskRow = bsk + static_cast<size_t>(i) * stepStride;

#pragma unroll
        for (int np = 0; np < NP; np++) {
            const PrimeF64 pr = a.prime[np];
            const double* tw = twBase + static_cast<size_t>(np) * 2 * N;
            const double* itw = tw + N;
            // own: output column tm (this team's); oth: column 1 - tm.  The
            // team index only moves addresses, never register-array indices.
            double own[16], oth[16];
            if constexpr (C::L == 0) {
#pragma unroll
                for (int r = 0; r < 16; r++) { own[r] = 0.0; oth[r] = 0.0; }
            }

            auto level = [&](const int lvl) {
                const int d = lvl * K1 + tm;
                // BSK values of this digit polynomial: the own column is issued
                // before the transform so its latency overlaps it, the other
                // column after it (registers are the binding resource; the
                // workgroup's ciphertexts run in lockstep, so these loads hit
                // in L1 after the first ciphertext of the workgroup)
                double kbo[16], kbt[16];
                const double* __restrict__ kp = bskRow + (static_cast<size_t>(d) * K1 * NP + np) * N + tl;
                // global address space: through the key table the compiler
                // cannot infer it and issues flat loads, which also count
                // against the LDS wait counter, so every transform exchange
                // would wait for these loads to land
                using GD = const __attribute__((address_space(1))) double;
                GD* __restrict__ kpo = (GD*)(kp + static_cast<size_t>(tm) * NP * N);
                GD* __restrict__ kpt = (GD*)(kp + static_cast<size_t>(1 - tm) * NP * N);
                if constexpr (NP == 1) {
#pragma unroll
                    for (int r = 0; r < 16; r++) kbo[r] = kpo[r * T];
                }
                double x[16];
                if constexpr (TAB == 2) {
                    WorkgroupSync<TW> sy;
                    fwd_first_two_stages(x, dg0, lvl, firstTab);
                    if constexpr (LANE) ntt_forward_from_stage2_t<LOGN>(x, scr, tl, fl0, pr, sy);
                    else if constexpr (P0REG) ntt_forward_from_stage2_t<LOGN>(x, scr, tl, fw0, pr, sy);
                    else ntt_forward_from_stage2<LOGN>(x, scr, tl, tw, pr, sy);
                } else if constexpr (TAB == 1) {
                    WorkgroupSync<TW> sy;
                    fwd_first_stage(x, dg0, lvl, firstTab);
                    ntt_forward_from_stage1<LOGN>(x, scr, tl, tw, pr, sy);
                } else {
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        int dv;
                        if constexpr (C::L > 0 && C::RB <= 7) {
                            dv = __builtin_amdgcn_sbfe(static_cast<int>(dg0[r]), 8 * lvl, 8);
                        } else if constexpr (C::L > 0) {
                            const uint32_t w = lvl < 2 ? dg0[r] : dg1[r];
                            dv = __builtin_amdgcn_sbfe(static_cast<int>(w), 16 * (lvl & 1), 16);
                        } else {
                            dv = digit_runtime(dg0[r], dg1[r], lvl, rb);
                        }
                        x[r] = static_cast<double>(dv);
                    }
                    ntt_forward<LOGN, TW>(x, scr, tl, tw, pr);
                }
                if constexpr (NP > 1) {
#pragma unroll
                    for (int r = 0; r < 16; r++) kbo[r] = kpo[r * T];
                }
                // the other column's loads are issued before the own column's
                // products so their latency overlaps them (1.9 % per launch
                // over issuing them after); the scheduling barrier keeps the
                // two columns' products apart, which bounds the live
                // registers; the first level of an unrolled loop assigns
                // instead of adding
                const bool first = C::L > 0 && lvl == 0;
#pragma unroll
                for (int r = 0; r < 16; r++) kbt[r] = kpt[r * T];
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const double pm = mulmod(x[r], kbo[r], pr);
                    own[r] = first ? pm : __dadd_rn(own[r], pm);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const double pm = mulmod(x[r], kbt[r], pr);
                    oth[r] = first ? pm : __dadd_rn(oth[r], pm);
                }
                if constexpr (C::L == 0) {
                    // more than four levels: reduce the partial sums once after
                    // four, so that with the partner's they stay below 16p
                    // (each product is at most 1.34p)
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

            // partner sums: hand column 1-tm to the other team, gather the full
            // sum of column tm
            team_sync<TW>();
#pragma unroll
            for (int r = 0; r < 16; r++) scr[r * T + tl] = oth[r];
            __syncthreads();
            const double* other = scratch[lc] + (1 - tm) * SCR;
            double x[16];
#pragma unroll
            for (int r = 0; r < 16; r++) x[r] = __dadd_rn(own[r], other[r * T + tl]);
            if constexpr (HANDOFF) {
                // the reads of the partner's scratch are complete: publish the
                // step count (one writer per word: this team's wave)
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __hip_atomic_store(&readDone[lc * 2 + tm], static_cast<uint32_t>(i + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                // the inverse transform; before its first write to this
                // team's scratch, wait until the partner has read it (bounded:
                // a wait that runs out sets the failure flag, and no later
                // wait blocks)
                WorkgroupSync<TW> sy;
                auto inverse = [&](const auto& twsrc) {
                    if constexpr (LOGN == 10 && C::L > 0 && C::L <= 3) {
                        // sums of at most six products: below 8p
                        inv_pass2_from_8p(x, tl, twsrc, pr);
                    } else {
                        reduce16(x, pr);
                        inv_pass_t<LOGN, 2>(x, tl, twsrc, pr);
                    }
                    lds_wait_at_least(&readDone[lc * 2 + 1 - tm], static_cast<uint32_t>(i + 1), a.barrier_polls, handFail);
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    exchange<LOGN, 2, 1, 1>(x, scr, tl, sy);
                    // pass 2's outputs are below 4p; the reduction schedule
                    // of ntt_inverse_sync_t from here on
                    inv_pass_from_4p<LOGN, 1>(x, tl, twsrc, pr);
                    x[0] = reduce(x[0], pr);
                    x[1] = reduce(x[1], pr);
                    exchange<LOGN, 1, 0, 0>(x, scr, tl, sy);
                    inv_pass_from_4p<LOGN, 0>(x, tl, twsrc, pr);
                };
                if constexpr (LANE) inverse(il0);
                else if constexpr (P0REG) inverse(iw0);
                else inverse(TwFlat{itw});
            } else {
                __syncthreads();
                if constexpr (LANE) {
                    WorkgroupSync<TW> sy;
                    ntt_inverse_sync_t<LOGN>(x, scr, tl, TwLane{tws + 16, twLane + TW_LANE_PER_DIR}, pr, sy);
                } else {
                    ntt_inverse<LOGN, TW>(x, scr, tl, itw, pr);
                }
            }
            if constexpr (NP == 1) {
                // centre mod q, reduce to the torus, accumulate (mod 2^w); the
                // lazy-pipe step replaces the accumulator, and leaves it where
                // abar = 0 (blind_rotate.cpp:893-895)
                centre16_after_inverse(x, pr);
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const double c = x[r];
                    if constexpr (C::LAZY) {
                        accr[r] = abar == 0 ? accr[r] : static_cast<TorusT>(f64_low32(c));
                    } else if constexpr (C::TBITS == 32) {
                        accr[r] = static_cast<TorusT>(static_cast<uint32_t>(accr[r]) + f64_low32(c));
                    } else {
                        accr[r] = static_cast<TorusT>(sext_t<C::TBITS>(static_cast<int64_t>(accr[r]) + f64_to_i64(c)));
                    }
                }
            } else {
                // prime-0 residues of the owned column at this lane's
                // positions; each lane reads back only what it wrote, so no
                // barrier is needed
                double* res = resbuf[lc] + tm * N;
                if (np == 0) {
#pragma unroll
                    for (int r = 0; r < 16; r++) res[r * T + tl] = centre(x[r], pr);
                } else {
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        const int64_t val = crt_mod_q(res[r * T + tl], centre(x[r], pr), pr, a.crt);
                        accr[r] = static_cast<TorusT>(sext_t<C::TBITS>(static_cast<int64_t>(accr[r]) + val));
                    }
                }
            }
        }
    }

    // a hand-over wait that ran out leaves results that are not proven: the
    // accumulators read the most negative torus value instead, and the
    // launch's status word reports it
    bool failed = false;
    if constexpr (HANDOFF) {
        __syncthreads();
        failed = __hip_atomic_load(handFail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u;
        report_barrier_failure(failed, a.status);
    }
    if (live) {
        TorusT* __restrict__ outAcc = static_cast<TorusT*>(a.acc_out) + static_cast<size_t>(ct) * K1 * N + tm * N;
#pragma unroll
        for (int r = 0; r < 16; r++) outAcc[tl + T * r] = failed ? std::numeric_limits<TorusT>::min() : accr[r];
    }
}

// ============================================================ keyswitch
// switchKeyForTlwe (keyswitching.cpp:48-72) of the sample extracted at index
// idx (extractTlweFromTrlwe, trlwe.cpp:241-250; idx = param.driftPhase, 0 by
// default):
//   a'_j = a_{idx-j} (j <= idx), a'_j = -a_{N+idx-j} (j > idx), b' = b_idx,
//   each modswitched 2^w -> 2^32;
//   out.a[k] = -sum_{i,j} d_{i,j} KSK[i][j].a[k],  out.b = b' - sum d KSK[i][j].b
// (mod 2^32; the reference's int64 sum + longModP is the same residue).
//
// An integer GEMM on the matrix cores: digits D[ct][row] (row = i*t + j,
// signed 8-bit) times KSK[row][col].  The KSK is stored as four signed-byte
// planes, x = sum_b s_b 2^(8b) mod 2^32 with s_b in [-128, 127], so each plane
// is an int8 GEMM with exact int32 sums (|sum| <= rows * 2^(b_ks-1) * 2^7,
// 2^22 at most here) and the planes recombine mod 2^32.  Both operands are
// laid out so that one MFMA fragment (16 consecutive rows of one ciphertext
// or one column) is 16 contiguous bytes: digits [ct][group], planes
// [plane][col][group], a group being four rows packed one byte each.
// Coefficient j of the sample extracted at index `idx` from the polynomial at
// `base` (extractTlweFromTrlwe, trlwe.cpp:241-250): acc[idx - j] for j <= idx,
// else -acc[N + idx - j] (negated in the Torus type: INT32_MIN wraps at 32
// bits), then modswitched to 2^32.  The body is coefficient 0 of the body
// polynomial's form: acc_b[idx].
template <int TBITS>
__device__ __forceinline__ int32_t extracted_coeff(const void* accv, size_t base, int N, int j, int idx) {
    using TorusT = typename std::conditional<TBITS == 32, int32_t, int64_t>::type;
    const TorusT* acc = static_cast<const TorusT*>(accv) + base;
    const int64_t x = (j <= idx) ? static_cast<int64_t>(acc[idx - j]) : sext_t<TBITS == 32 ? 32 : 64>(-static_cast<int64_t>(acc[N + idx - j]));
    return torus_to_lwe32<TBITS>(x);
}

// Keyswitch digits of the extracted samples (ks_digits of every coefficient
// a'_i, b' excluded): word [ct][g] packs the signed 8-bit digits of rows
// 4g .. 4g + 3 (row = i t + j: level j of coefficient i).  One thread per
// word; a coefficient spanning two words is decomposed for each.
template <int TBITS>
__global__ void __launch_bounds__(256) k_ks_digits(KsArgs a) {
    const int N = a.N, t = a.ks_level;
    const int rows = a.k * N * t;
    const int groups = (rows + 3) / 4;
    const size_t total = static_cast<size_t>(groups) * a.count;
    for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total;
         idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int ct = static_cast<int>(idx / groups), g = static_cast<int>(idx % groups);
        const size_t accBase = static_cast<size_t>(ct) * (a.k + 1) * N;
        int8_t dig[KS_MAX_LEVEL];
        int lastI = -1;
        uint32_t w = 0;
        for (int r = 0; r < 4; r++) {
            const int row = g * 4 + r;
            if (row >= rows) break;
            const int i = row / t, j = row % t;
            if (i != lastI) {
                const int32_t x = extracted_coeff<TBITS>(a.acc, accBase + static_cast<size_t>(i / N) * N, N, i % N, a.extract_index);
                ks_digits(x, t, a.ks_radix_bits, a.ks_width_bits, dig);
                lastI = i;
            }
            w |= (static_cast<uint32_t>(dig[j]) & 0xFFu) << (8 * r);
        }
        a.digits[idx] = w;
    }
}

typedef int ks_v4i __attribute__((ext_vector_type(4)));

// 16 consecutive rows (four groups) of one ciphertext or one column, or
// zeros outside the matrix.
__device__ __forceinline__ ks_v4i ks_fragment(const uint32_t* __restrict__ base, bool valid) {
    if (!valid) return ks_v4i{0, 0, 0, 0};
    using GV = const __attribute__((address_space(1))) ks_v4i;
    return *(GV*)(base);
}

// Tile: KS_TC ciphertexts x KS_TK columns per workgroup of four waves, each
// wave 32 x 32 as 2 x 2 tiles of v_mfma_i32_16x16x64_i8 per plane.  Lane l
// supplies rows 16 (l / 16) .. + 16 of the 64-row step for ciphertext (A) or
// column (B) l % 16, and holds outputs rows 4 (l / 16) + v, column l % 16.
// grid.z slices the rows for small batches (partial sums meet in an atomic
// add, exact mod 2^32).  A tile's ciphertexts are processed in runs of one
// key (a key-grouped batch has contiguous jobs, so most tiles hold one run).
template <int TBITS>
__global__ void __launch_bounds__(KS_THREADS) k_extract_ks(KsArgs a) {
    const int n = a.n, N = a.N, cols = n + 1;
    const int groups = (a.k * N * a.ks_level + 3) / 4;
    const int steps = (groups + 15) / 16;  // 64-row MFMA steps
    const int ct0 = static_cast<int>(blockIdx.x) * KS_TC;
    const int ctEnd = min(ct0 + KS_TC, a.count);
    const int col0 = static_cast<int>(blockIdx.y) * KS_TK;
    const int slices = static_cast<int>(gridDim.z);
    const int per = (steps + slices - 1) / slices;
    const int s0 = static_cast<int>(blockIdx.z) * per;
    const int s1 = min(steps, s0 + per);
    const int lane = static_cast<int>(threadIdx.x) % 64;
    const int wave = static_cast<int>(threadIdx.x) / 64;
    const int li = lane % 16, lq = lane / 16;
    const int ctW = ct0 + (wave / 2) * 32, colW = col0 + (wave % 2) * 32;

    for (int segStart = ct0; segStart < ctEnd;) {
        const int slot = a.ct_slot ? a.ct_slot[segStart] : 0;
        int segEnd = segStart + 1;
        if (a.ct_slot) {
            while (segEnd < ctEnd && a.ct_slot[segEnd] == slot) segEnd++;
        } else {
            segEnd = ctEnd;
        }
        const uint32_t* __restrict__ planes = reinterpret_cast<const uint32_t*>(a.ksk_table[slot]);

        ks_v4i acc[4][2][2];
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) acc[b][x][y] = ks_v4i{0, 0, 0, 0};

        // this lane's fragment sources: ciphertexts of the run, columns of the matrix
        bool aok[2], bok[2];
        const uint32_t* ap[2];
        const uint32_t* bp[2];
#pragma unroll
        for (int x = 0; x < 2; x++) {
            const int ct = ctW + 16 * x + li;
            aok[x] = ct >= segStart && ct < segEnd;
            ap[x] = a.digits + static_cast<size_t>(aok[x] ? ct : 0) * groups + 4 * lq;
            const int col = colW + 16 * x + li;
            bok[x] = col < cols;
            bp[x] = planes + static_cast<size_t>(bok[x] ? col : 0) * groups + 4 * lq;
        }
        const size_t planeStride = static_cast<size_t>(cols) * groups;
        for (int st = s0; st < s1; st++) {
            const int g = st * 16;
            const bool inK = g + 4 * lq < groups;  // groups is a multiple of 4
            ks_v4i fa[2], fb[4][2];
#pragma unroll
            for (int x = 0; x < 2; x++) fa[x] = ks_fragment(ap[x] + g, aok[x] && inK);
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int y = 0; y < 2; y++) fb[b][y] = ks_fragment(bp[y] + b * planeStride + g, bok[y] && inK);
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int x = 0; x < 2; x++)
#pragma unroll
                    for (int y = 0; y < 2; y++)
                        acc[b][x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[x], fb[b][y], acc[b][x][y], 0, 0, 0);
        }

#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int ct = ctW + 16 * x + 4 * lq + v;
                    const int col = colW + 16 * y + li;
                    if (ct < segStart || ct >= segEnd || col >= cols) continue;
                    const uint32_t sum = static_cast<uint32_t>(acc[0][x][y][v]) + (static_cast<uint32_t>(acc[1][x][y][v]) << 8) +
                                         (static_cast<uint32_t>(acc[2][x][y][v]) << 16) + (static_cast<uint32_t>(acc[3][x][y][v]) << 24);
                    if (slices == 1) {
                        uint32_t o;
                        if (col < n) {
                            o = 0u - sum;
                        } else {
                            const int32_t b = extracted_coeff<TBITS>(a.acc, static_cast<size_t>(ct) * (a.k + 1) * N + static_cast<size_t>(a.k) * N, N, 0, a.extract_index);
                            o = static_cast<uint32_t>(b) - sum;
                        }
                        a.lwe_out[static_cast<size_t>(ct) * cols + col] = static_cast<int32_t>(o);
                    } else {
                        atomicAdd(reinterpret_cast<unsigned int*>(a.partial) + static_cast<size_t>(ct) * cols + col, sum);
                    }
                }
        segStart = segEnd;
    }
}

template <int TBITS>
__global__ void k_ks_finalize(KsArgs a) {
    const int cols = a.n + 1;
    const size_t total = static_cast<size_t>(a.count) * cols;
    for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total; idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int ct = static_cast<int>(idx / cols), col = static_cast<int>(idx % cols);
        const uint32_t s = reinterpret_cast<const uint32_t*>(a.partial)[idx];
        uint32_t o;
        if (col < a.n) {
            o = 0u - s;
        } else {
            const int32_t b = extracted_coeff<TBITS>(a.acc, static_cast<size_t>(ct) * (a.k + 1) * a.N + static_cast<size_t>(a.k) * a.N, a.N, 0, a.extract_index);
            o = static_cast<uint32_t>(b) - s;
        }
        a.lwe_out[idx] = static_cast<int32_t>(o);
    }
}

__global__ void k_zero_u32(uint32_t* p, size_t count) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < count; i += static_cast<size_t>(gridDim.x) * blockDim.x) p[i] = 0;
}

template <int TBITS>
__global__ void k_extract(KsArgs a, int32_t* out) {
    const int len = a.k * a.N + 1;
    const size_t total = static_cast<size_t>(a.count) * len;
    for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < total; idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int ct = static_cast<int>(idx / len), i = static_cast<int>(idx % len);
        const size_t base = static_cast<size_t>(ct) * (a.k + 1) * a.N;
        int32_t v;
        if (i < a.k * a.N) v = extracted_coeff<TBITS>(a.acc, base + static_cast<size_t>(i / a.N) * a.N, a.N, i % a.N, a.extract_index);
        else v = extracted_coeff<TBITS>(a.acc, base + static_cast<size_t>(a.k) * a.N, a.N, 0, a.extract_index);
        out[idx] = v;
    }
}

// ============================================================ measurement
// Modmul throughput microbenchmark (the roofline denominator): every lane runs
// MM_CHAINS independent chains x <- x * w mod p, either with the binary64 exact
// product the engine uses or with a 64-bit integer Shoup product.
constexpr int MM_CHAINS = 8;

__global__ void __launch_bounds__(256) k_modmul_peak_f64(PrimeF64 pr, int iters, double* sink) {
    double x[MM_CHAINS], w[MM_CHAINS];
#pragma unroll
    for (int c = 0; c < MM_CHAINS; c++) {
        x[c] = static_cast<double>((threadIdx.x * 131 + c * 977 + blockIdx.x) % 100000) + 1.0;
        w[c] = static_cast<double>(123456789 + c * 1000003);
    }
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int c = 0; c < MM_CHAINS; c++) x[c] = mulmod(x[c], w[c], pr);
    }
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < MM_CHAINS; c++) acc = __dadd_rn(acc, x[c]);
    if (acc == 12345.0) sink[blockIdx.x] = acc;  // keeps the chains live
}

__global__ void __launch_bounds__(256) k_modmul_peak_u64(uint64_t q, int iters, uint64_t* sink) {
    uint64_t x[MM_CHAINS], w[MM_CHAINS], wp[MM_CHAINS];
#pragma unroll
    for (int c = 0; c < MM_CHAINS; c++) {
        x[c] = (threadIdx.x * 131 + c * 977 + blockIdx.x) % 100000 + 1;
        w[c] = 123456789 + c * 1000003;
        wp[c] = static_cast<uint64_t>((static_cast<unsigned __int128>(w[c]) << 64) / q);
    }
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int c = 0; c < MM_CHAINS; c++) {
            const uint64_t hi = __umul64hi(x[c], wp[c]);
            uint64_t r = x[c] * w[c] - hi * q;
            x[c] = r >= q ? r - q : r;
        }
    }
    uint64_t acc = 0;
#pragma unroll
    for (int c = 0; c < MM_CHAINS; c++) acc += x[c];
    if (acc == 12345) sink[blockIdx.x] = acc;
}

// Plain streaming copy, the calibration kernel for HBM byte counters.
__global__ void __launch_bounds__(256) k_copy_u32x4(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

// Word copy for small host<->device staging through mapped pinned memory: a
// compute kernel, so it is not queued behind bulk DMA transfers.
__global__ void __launch_bounds__(256) k_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy_words(const void* src, void* dst, size_t words, hipStream_t s) {
    if (words == 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((words + 255) / 256, 1024));
    hipLaunchKernelGGL(k_copy_u32, dim3(blocks), dim3(256), 0, s, static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), words);
    return hipGetLastError();
}

hipError_t launch_modmul_peak(int kind, const PrimeF64& pr, uint64_t q, int blocks, int iters, void* sink, hipStream_t s) {
    if (kind == 0) hipLaunchKernelGGL(k_modmul_peak_f64, dim3(blocks), dim3(256), 0, s, pr, iters, static_cast<double*>(sink));
    else hipLaunchKernelGGL(k_modmul_peak_u64, dim3(blocks), dim3(256), 0, s, q, iters, static_cast<uint64_t*>(sink));
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 16;
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 8192));
    hipLaunchKernelGGL(k_copy_u32x4, dim3(blocks ? blocks : 1), dim3(256), 0, s, static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
    return hipGetLastError();
}

// ============================================================ launchers
template <class C>
hipError_t launch_br(const BrArgs& a, hipStream_t s) {
    const int groups = (a.count + C::CPW - 1) / C::CPW;
    const unsigned grid = static_cast<unsigned>((groups + 7) / 8 * 8);  // map_workgroup needs a multiple of 8
    hipLaunchKernelGGL(k_blind_rotate<C>, dim3(grid), dim3(C::THREADS), 0, s, a);
    return hipGetLastError();
}

// TORUS32 N=1024 2^7 x 3: ciphertexts per workgroup x twiddle placement x
// table-lookup forward stages.
template <int CPW, int TWL>
hipError_t launch_t32_b7x3_tab(int tab, const BrArgs& a, hipStream_t s) {
    if (tab == 2) return launch_br<BrCfg<32, 10, 1, CPW, true, 3, 7, TWL, 2>>(a, s);
    if (tab == 1) return launch_br<BrCfg<32, 10, 1, CPW, true, 3, 7, TWL, 1>>(a, s);
    return launch_br<BrCfg<32, 10, 1, CPW, true, 3, 7, TWL, 0>>(a, s);
}
template <int CPW>
hipError_t launch_t32_b7x3_twl(const BrTile& t, const BrArgs& a, hipStream_t s) {
    if (t.twl == 2) return launch_br<BrCfg<32, 10, 1, CPW, true, 3, 7, 2, 2>>(a, s);  // lane tables: TAB = 2 only
    return t.twl ? launch_t32_b7x3_tab<CPW, 1>(t.tab, a, s) : launch_t32_b7x3_tab<CPW, 0>(t.tab, a, s);
}
hipError_t launch_t32_b7x3(const BrTile& t, const BrArgs& a, hipStream_t s) {
    if (t.cpw >= 4) return launch_t32_b7x3_twl<4>(t, a, s);
    if (t.cpw == 2) return launch_t32_b7x3_twl<2>(t, a, s);
    return launch_t32_b7x3_twl<1>(t, a, s);
}

// Specialised gadget shapes (the headline parameter sets) come in several
// tiles; every other parameter set runs a runtime-gadget instantiation.
hipError_t launch_blind_rotate(int tbits, int logn, const BrTile& t, int radix_bits, const BrArgs& a, hipStream_t s) {
    if (t.lat && blind_rotate_lat_supported(tbits, logn, a.level, radix_bits)) return launch_blind_rotate_lat(a, s);
    const bool p8 = radix_bits <= 8;
    if (tbits == 32 && logn == 10 && a.level == 3 && radix_bits == 7) return launch_t32_b7x3(t, a, s);
    if (tbits == 56) return launch_blind_rotate_crt(logn, t, radix_bits, a.level, a, s);
    if (tbits == 32 && logn == 10) {
        if (t.cpw >= 4) return p8 ? launch_br<BrCfg<32, 10, 1, 4, true>>(a, s) : launch_br<BrCfg<32, 10, 1, 4, false>>(a, s);
        return p8 ? launch_br<BrCfg<32, 10, 1, 1, true>>(a, s) : launch_br<BrCfg<32, 10, 1, 1, false>>(a, s);
    }
    if (tbits == 32 && logn == 11) return p8 ? launch_br<BrCfg<32, 11, 1, 1, true>>(a, s) : launch_br<BrCfg<32, 11, 1, 1, false>>(a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_blind_rotate_lazy(const BrArgs& a, hipStream_t s) {
    if (a.n < 1 || a.level < 1 || a.level > 8 || a.radix_bits < 1 || a.radix_bits > 7 || !a.acc_in || !a.ct_slot)
        return hipErrorInvalidValue;
    if (a.level == 3 && a.radix_bits == 7) return launch_br<BrCfg<32, 10, 1, 2, true, 3, 7, 1, 2, true>>(a, s);
    return launch_br<BrCfg<32, 10, 1, 2, true, 0, 0, 1, 0, true>>(a, s);
}

int blind_rotate_tiles(int tbits, int logn, int level, int radix_bits, BrTile* out, int max) {
    BrTile c[MAX_TILES];
    int n = 0;
    if (tbits == 32 && logn == 10 && level == 3 && radix_bits == 7) {
        for (int cpw : {4, 2, 1}) {
            c[n++] = BrTile{cpw, 2, 0, 2};
            for (int twl : {1, 0})
                for (int tab : {0, 1, 2}) c[n++] = BrTile{cpw, twl, 0, tab};
        }
    } else if (tbits == 56 && logn == 11 && level == 3 && radix_bits == 10) {
        for (int tab : {1, 0})
            for (int twl : {1, 0}) c[n++] = BrTile{1, twl, 0, tab};
    } else if (tbits == 56) {
        c[n++] = BrTile{1, 1};
    } else if (tbits == 32 && logn == 10) {
        c[n++] = BrTile{4, 1};
        c[n++] = BrTile{1, 1};
    } else {
        c[n++] = BrTile{1, (tbits == 32 && logn == 10) ? 1 : 0};
    }
    for (int i = 0; i < n && i < max; i++) out[i] = c[i];
    return n < max ? n : max;
}

hipError_t launch_extract_ks(int tbits, const KsArgs& a, int slices, hipStream_t s) {
    // the MFMA fragments are 16-byte loads of four groups
    if (!a.digits || (static_cast<size_t>(a.k) * a.N * a.ks_level) % 16 != 0) return hipErrorInvalidValue;
    const dim3 grid((a.count + KS_TC - 1) / KS_TC, (a.n + 1 + KS_TK - 1) / KS_TK, slices);
    {
        const size_t words = (static_cast<size_t>(a.k) * a.N * a.ks_level + 3) / 4 * a.count;
        const unsigned blocks = static_cast<unsigned>(std::min<size_t>((words + 255) / 256, 65536));
        if (tbits == 32) hipLaunchKernelGGL(k_ks_digits<32>, dim3(blocks), dim3(256), 0, s, a);
        else hipLaunchKernelGGL(k_ks_digits<56>, dim3(blocks), dim3(256), 0, s, a);
    }
    if (slices > 1) {
        const size_t total = static_cast<size_t>(a.count) * (a.n + 1);
        hipLaunchKernelGGL(k_zero_u32, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s,
                           reinterpret_cast<uint32_t*>(a.partial), total);
    }
    if (tbits == 32) hipLaunchKernelGGL(k_extract_ks<32>, grid, dim3(KS_THREADS), 0, s, a);
    else hipLaunchKernelGGL(k_extract_ks<56>, grid, dim3(KS_THREADS), 0, s, a);
    if (slices > 1) {
        const size_t total = static_cast<size_t>(a.count) * (a.n + 1);
        const unsigned blocks = static_cast<unsigned>(std::min<size_t>((total + 255) / 256, 4096));
        if (tbits == 32) hipLaunchKernelGGL(k_ks_finalize<32>, dim3(blocks), dim3(256), 0, s, a);
        else hipLaunchKernelGGL(k_ks_finalize<56>, dim3(blocks), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_extract(int tbits, const KsArgs& a, int32_t* out, hipStream_t s) {
    const size_t total = static_cast<size_t>(a.count) * (a.k * a.N + 1);
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((total + 255) / 256, 4096));
    if (tbits == 32) hipLaunchKernelGGL(k_extract<32>, dim3(blocks), dim3(256), 0, s, a, out);
    else hipLaunchKernelGGL(k_extract<56>, dim3(blocks), dim3(256), 0, s, a, out);
    return hipGetLastError();
}

}  // namespace ygpu
