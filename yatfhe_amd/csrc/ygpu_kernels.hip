// ygpu_kernels.hip -- CDNA4 kernels of the bootstrapping path.
//
//  k_blind_rotate   rescale to 2N, accumulator init, and all n CMUX steps of the
//                   blind rotation in one persistent launch; one workgroup per
//                   ciphertext, the accumulator resident in LDS for all steps.
//  k_extract_ks     sample extract (index 0) fused into the LWE keyswitch, a
//                   GEMM-shaped integer reduction over the KSK, tiled over
//                   ciphertexts so each KSK row is read once per tile.
//  k_extract        the extracted LWE alone (for callers that want it).
//  k_ks_finalize    negation / body of the sliced keyswitch.
#include "ygpu_kernels.h"

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
__global__ void __launch_bounds__(C::THREADS) k_blind_rotate(BrArgs a) {
    using TorusT = typename C::TorusT;
    constexpr int LOGN = C::LOGN, N = C::N, T = C::T, K1 = 2, NP = C::NP, TW = C::TEAM_WAVES, CPW = C::CPW;
    constexpr int SCR = Scratch<LOGN>::DOUBLES;
    constexpr int CT_THREADS = 2 * T;
    constexpr int TWN = C::TW_LDS ? NP * 2 * N : 1;

    __shared__ double tws[TWN];
    __shared__ TorusT accs[CPW][K1 * N];
    __shared__ double scratch[CPW][2 * SCR];

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
    TorusT* acc = accs[lc];
    double* scr = scratch[lc] + tm * SCR;

    const int n = a.n;
    const int L = a.level;
    const int rb = a.radix_bits;
    const int slot = a.ct_slot ? a.ct_slot[ct] : 0;
    const double* __restrict__ bsk = a.bsk_table[slot];
    const int32_t* __restrict__ lwe = a.lwe_in + static_cast<size_t>(ct) * (n + 1);
    const TorusT* __restrict__ v = static_cast<const TorusT*>(a.v) + static_cast<size_t>(a.v_stride) * ct;

    if constexpr (C::TW_LDS) {
        for (int j = static_cast<int>(threadIdx.x); j < TWN; j += C::THREADS) tws[j] = a.tw[j];
    }
    // rescaleTlweToNewMod (tlwe.cpp:75-81) of the body, then
    // genNoiselessTrlweSample (trlwe.cpp:175-182): acc = (0, X^{-bbar} v).
    {
        const int bbar = mod_switch_2n(lwe[n], LOGN + 1);
        int aT, wrap;
        validate_rotator(-bbar, N, aT, wrap);
        for (int j = ctid; j < N; j += CT_THREADS) {
            acc[j] = 0;
            const int src = j - aT;
            const int64_t val = src < 0 ? -static_cast<int64_t>(v[src + N]) * wrap : static_cast<int64_t>(v[src]) * wrap;
            acc[N + j] = static_cast<TorusT>(sext_t<C::TBITS>(val));
        }
    }
    __syncthreads();
    const double* twBase = C::TW_LDS ? static_cast<const double*>(tws) : a.tw;

    const size_t stepStride = static_cast<size_t>(L) * K1 * K1 * NP * N;  // doubles per BSK row
    TorusT* accRow = acc + tm * N;
    for (int i = 0; i < n; i++) {
        // (blind_rotate.cpp:1557-1559 skips a_i == 0; ExtProd(0) = 0 exactly,
        // so the lockstep workgroup simply computes it.)
        const int abar = mod_switch_2n(lwe[i], LOGN + 1);
        int aT, wrap;
        validate_rotator(abar, N, aT, wrap);

        // (X^a - 1) acc_r at this lane's pass-0 positions (rotateTorusPolynomialMinusOne,
        // polynomial.cpp:213-238) and its digits, packed 8 or 16 bits per digit.
        uint32_t dg0[16], dg1[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int j = tl + T * r;
            int src = j - aT;
            int sgn = wrap;
            if (src < 0) { src += N; sgn = -wrap; }
            const int64_t rot = sext_t<C::TBITS>(static_cast<int64_t>(accRow[src]) * sgn - static_cast<int64_t>(accRow[j]));
            int d4[4] = {0, 0, 0, 0};
            digits_of<C::TBITS>(rot, L, rb, d4);
            if constexpr (C::PACK8) {
                dg0[r] = (static_cast<uint32_t>(d4[0]) & 0xFFu) | ((static_cast<uint32_t>(d4[1]) & 0xFFu) << 8) |
                         ((static_cast<uint32_t>(d4[2]) & 0xFFu) << 16) | (static_cast<uint32_t>(d4[3]) << 24);
                dg1[r] = 0;
            } else {
                dg0[r] = (static_cast<uint32_t>(d4[0]) & 0xFFFFu) | (static_cast<uint32_t>(d4[1]) << 16);
                dg1[r] = (static_cast<uint32_t>(d4[2]) & 0xFFFFu) | (static_cast<uint32_t>(d4[3]) << 16);
            }
        }

        const double* __restrict__ bskRow = bsk + static_cast<size_t>(i) * stepStride;
        double res[NP > 1 ? 16 : 1];  // prime-0 residues of the owned column (NP == 2)

#pragma unroll
        for (int np = 0; np < NP; np++) {
            const PrimeF64 pr = a.prime[np];
            const double* tw = twBase + static_cast<size_t>(np) * 2 * N;
            const double* itw = tw + N;
            // own: output column tm (this team's); oth: column 1 - tm.  The
            // team index only moves addresses, never register-array indices.
            double own[16], oth[16];
#pragma unroll
            for (int r = 0; r < 16; r++) { own[r] = 0.0; oth[r] = 0.0; }

            for (int lvl = 0; lvl < L; lvl++) {
                const int d = lvl * K1 + tm;
                // BSK values of this digit polynomial, both columns, issued
                // before the transform so their latency overlaps it.
                double kbo[16], kbt[16];
                const double* __restrict__ kp = bskRow + (static_cast<size_t>(d) * K1 * NP + np) * N + tl;
                const double* __restrict__ kpo = kp + static_cast<size_t>(tm) * NP * N;
                const double* __restrict__ kpt = kp + static_cast<size_t>(1 - tm) * NP * N;
#pragma unroll
                for (int r = 0; r < 16; r++) { kbo[r] = kpo[r * T]; kbt[r] = kpt[r * T]; }
                double x[16];
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    int dv;
                    if constexpr (C::PACK8) {
                        dv = static_cast<int32_t>(dg0[r] << (24 - 8 * lvl)) >> 24;
                    } else {
                        const uint32_t w = lvl < 2 ? dg0[r] : dg1[r];
                        dv = static_cast<int32_t>(w << (16 - 16 * (lvl & 1))) >> 16;
                    }
                    x[r] = static_cast<double>(dv);
                }
                ntt_forward<LOGN, TW>(x, scr, tl, tw, pr);
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    own[r] = __dadd_rn(own[r], mulmod(x[r], kbo[r], pr));
                    oth[r] = __dadd_rn(oth[r], mulmod(x[r], kbt[r], pr));
                }
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
            __syncthreads();

            ntt_inverse<LOGN, TW>(x, scr, tl, itw, pr);
            if constexpr (NP == 1) {
                // centre mod q, reduce to the torus, accumulate (mod 2^w)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int j = tl + T * r;
                    const int64_t val = static_cast<int64_t>(centre(x[r], pr));
                    accRow[j] = static_cast<TorusT>(sext_t<C::TBITS>(static_cast<int64_t>(accRow[j]) + val));
                }
            } else {
                if (np == 0) {
#pragma unroll
                    for (int r = 0; r < 16; r++) res[r] = centre(x[r], pr);
                } else {
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        const int j = tl + T * r;
                        const int64_t val = crt_mod_q(res[r], centre(x[r], pr), pr, a.crt);
                        accRow[j] = static_cast<TorusT>(sext_t<C::TBITS>(static_cast<int64_t>(accRow[j]) + val));
                    }
                }
            }
        }
        // each team reads back, at rotated positions, only the accumulator row
        // it wrote itself
        if constexpr (TW == 1) {
            team_sync<1>();
        } else {
            __syncthreads();
        }
    }

    __syncthreads();
    if (live) {
        TorusT* __restrict__ outAcc = static_cast<TorusT*>(a.acc_out) + static_cast<size_t>(ct) * K1 * N;
        for (int j = ctid; j < K1 * N; j += CT_THREADS) outAcc[j] = acc[j];
    }
}

// ============================================================ keyswitch
// switchKeyForTlwe (keyswitching.cpp:48-72) of the sample extracted at index 0
// (extractTlweFromTrlwe, trlwe.cpp:241-250):
//   a'_0 = a_0, a'_j = -a_{N-j}, b' = b_0, each modswitched 2^w -> 2^32;
//   out.a[k] = -sum_{i,t} d_{i,t} KSK[i][t].a[k],  out.b = b' - sum d KSK[i][t].b
// (mod 2^32; the reference's int64 sum + longModP is the same residue).
// Tile: KS_CT ciphertexts x all n+1 output columns per workgroup, rows of the
// extracted LWE split into `slices` along grid.y.  Digits of a chunk of rows
// are built once into LDS and broadcast; each KSK element is loaded once per
// tile and split into signed 16-bit halves so that every product is a
// full-rate 24-bit multiply-add: d*x = d*lo + (d*hi << 16) mod 2^32.
template <int TBITS>
__device__ __forceinline__ int32_t extracted_coeff(const void* accv, size_t base, int N, int j) {
    using TorusT = typename std::conditional<TBITS == 32, int32_t, int64_t>::type;
    const TorusT* acc = static_cast<const TorusT*>(accv) + base;
    const int64_t x = (j == 0) ? static_cast<int64_t>(acc[0]) : sext_t<TBITS == 32 ? 32 : 64>(-static_cast<int64_t>(acc[N - j]));
    return torus_to_lwe32<TBITS>(x);
}

template <int TBITS>
__global__ void __launch_bounds__(KS_THREADS) k_extract_ks(KsArgs a) {
    constexpr int CHUNK = 32;
    __shared__ int8_t dig[KS_CT][CHUNK][KS_MAX_LEVEL];
    const int n = a.n, N = a.N, t = a.ks_level, cols = n + 1;
    const int ct0 = static_cast<int>(blockIdx.x) * KS_CT;
    const int slices = static_cast<int>(gridDim.y);
    const int rows = a.k * N;
    const int per = (rows + slices - 1) / slices;
    const int i0 = static_cast<int>(blockIdx.y) * per;
    const int i1 = min(rows, i0 + per);
    const int tid = static_cast<int>(threadIdx.x);

    uint32_t accLo[KS_CT][KS_COLS_PER_THREAD], accHi[KS_CT][KS_COLS_PER_THREAD];
#pragma unroll
    for (int c = 0; c < KS_CT; c++)
#pragma unroll
        for (int m = 0; m < KS_COLS_PER_THREAD; m++) { accLo[c][m] = 0; accHi[c][m] = 0; }

    const int32_t* kskBase[KS_CT];
    size_t accBase[KS_CT];
#pragma unroll
    for (int c = 0; c < KS_CT; c++) {
        const int ct = min(ct0 + c, a.count - 1);
        const int slot = a.ct_slot ? a.ct_slot[ct] : 0;
        kskBase[c] = a.ksk_table[slot];
        accBase[c] = static_cast<size_t>(ct) * (a.k + 1) * N;
    }
    // Key-grouped tiles share one KSK; a tile spanning two keys falls back to
    // per-ciphertext rows.
    bool oneKey = true;
#pragma unroll
    for (int c = 1; c < KS_CT; c++) oneKey = oneKey && (kskBase[c] == kskBase[0]);

    for (int base = i0; base < i1; base += CHUNK) {
        __syncthreads();
        for (int idx = tid; idx < KS_CT * CHUNK; idx += KS_THREADS) {
            const int c = idx / CHUNK, ii = idx % CHUNK;
            const int i = base + ii;
            int8_t* dd = dig[c][ii];
            if (i < i1 && ct0 + c < a.count) {
                const int poly = i / N, j = i % N;
                const int32_t x = extracted_coeff<TBITS>(a.acc, accBase[c] + static_cast<size_t>(poly) * N, N, j);
                ks_digits(x, t, a.ks_radix_bits, a.ks_width_bits, dd);
            } else {
                for (int l = 0; l < KS_MAX_LEVEL; l++) dd[l] = 0;
            }
        }
        __syncthreads();
        const int iEnd = min(CHUNK, i1 - base);
        for (int ii = 0; ii < iEnd; ii++) {
            const int i = base + ii;
            for (int l = 0; l < t; l++) {
                const size_t rowOff = (static_cast<size_t>(i) * t + l) * cols;
                if (oneKey) {
                    int32_t lo[KS_COLS_PER_THREAD], hi[KS_COLS_PER_THREAD];
#pragma unroll
                    for (int m = 0; m < KS_COLS_PER_THREAD; m++) {
                        const int col = tid + m * KS_THREADS;
                        const int32_t x = col < cols ? kskBase[0][rowOff + col] : 0;
                        lo[m] = x & 0xFFFF;
                        hi[m] = x >> 16;
                    }
#pragma unroll
                    for (int c = 0; c < KS_CT; c++) {
                        const int d = dig[c][ii][l];
#pragma unroll
                        for (int m = 0; m < KS_COLS_PER_THREAD; m++) {
                            accLo[c][m] += static_cast<uint32_t>(__mul24(d, lo[m]));
                            accHi[c][m] += static_cast<uint32_t>(__mul24(d, hi[m]));
                        }
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < KS_CT; c++) {
                        const int d = dig[c][ii][l];
#pragma unroll
                        for (int m = 0; m < KS_COLS_PER_THREAD; m++) {
                            const int col = tid + m * KS_THREADS;
                            const int32_t x = col < cols ? kskBase[c][rowOff + col] : 0;
                            accLo[c][m] += static_cast<uint32_t>(__mul24(d, x & 0xFFFF));
                            accHi[c][m] += static_cast<uint32_t>(__mul24(d, x >> 16));
                        }
                    }
                }
            }
        }
    }

#pragma unroll
    for (int c = 0; c < KS_CT; c++) {
        const int ct = ct0 + c;
        if (ct >= a.count) continue;
#pragma unroll
        for (int m = 0; m < KS_COLS_PER_THREAD; m++) {
            const int col = tid + m * KS_THREADS;
            if (col >= cols) continue;
            const uint32_t s = accLo[c][m] + (accHi[c][m] << 16);
            if (slices == 1) {
                uint32_t o;
                if (col < n) {
                    o = 0u - s;
                } else {
                    const int32_t b = extracted_coeff<TBITS>(a.acc, accBase[c] + static_cast<size_t>(a.k) * N, N, 0);
                    o = static_cast<uint32_t>(b) - s;
                }
                a.lwe_out[static_cast<size_t>(ct) * cols + col] = static_cast<int32_t>(o);
            } else {
                atomicAdd(reinterpret_cast<unsigned int*>(a.partial) + static_cast<size_t>(ct) * cols + col, s);
            }
        }
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
            const int32_t b = extracted_coeff<TBITS>(a.acc, static_cast<size_t>(ct) * (a.k + 1) * a.N + static_cast<size_t>(a.k) * a.N, a.N, 0);
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
        if (i < a.k * a.N) v = extracted_coeff<TBITS>(a.acc, base + static_cast<size_t>(i / a.N) * a.N, a.N, i % a.N);
        else v = extracted_coeff<TBITS>(a.acc, base + static_cast<size_t>(a.k) * a.N, a.N, 0);
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

// cpw: ciphertexts per workgroup (1 for the plain form)
hipError_t launch_blind_rotate(int tbits, int logn, int cpw, int radix_bits, const BrArgs& a, hipStream_t s) {
    const bool p8 = radix_bits <= 8;
    if (tbits == 32 && logn == 10) {
        if (cpw >= 4) return p8 ? launch_br<BrCfg<32, 10, 1, 4, true>>(a, s) : launch_br<BrCfg<32, 10, 1, 4, false>>(a, s);
        return p8 ? launch_br<BrCfg<32, 10, 1, 1, true>>(a, s) : launch_br<BrCfg<32, 10, 1, 1, false>>(a, s);
    }
    if (tbits == 32 && logn == 11) return p8 ? launch_br<BrCfg<32, 11, 1, 1, true>>(a, s) : launch_br<BrCfg<32, 11, 1, 1, false>>(a, s);
    if (tbits == 56 && logn == 10) return p8 ? launch_br<BrCfg<56, 10, 2, 1, true>>(a, s) : launch_br<BrCfg<56, 10, 2, 1, false>>(a, s);
    if (tbits == 56 && logn == 11) return p8 ? launch_br<BrCfg<56, 11, 2, 1, true>>(a, s) : launch_br<BrCfg<56, 11, 2, 1, false>>(a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_extract_ks(int tbits, const KsArgs& a, int slices, hipStream_t s) {
    const dim3 grid((a.count + KS_CT - 1) / KS_CT, slices);
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
