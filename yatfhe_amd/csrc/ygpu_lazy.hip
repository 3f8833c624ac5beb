// ygpu_lazy.hip -- the lazy-pipe key format (BootstrappingKeyMPLazyPipe) on the
// engine: upload and bootstrapping.
//
// blindRotateLazyPipeNtt (blind_rotate.cpp:822-902) rebuilds every step's RGSW
// rows from the key's stored digits for each ciphertext.  Here a batch runs in
// windows of steps: k_lazy_derive writes the rows of the window's steps for
// every ciphertext into a workspace, then the lazy form of the batched blind
// rotation advances every accumulator through those steps, each ciphertext
// reading its own rows through the key-slot table.  The window is the largest
// number of steps whose rows fit LAZY_WS_BYTES for the batch (at least one).
#include "ygpu_internal.h"

namespace {

// 4 GiB of derived rows per batch window (32 KiB per ciphertext, step and row
// level at TORUS32 N = 1024): 8 steps for 4096 ciphertexts at 4 levels.
constexpr size_t LAZY_WS_BYTES = size_t{4} << 30;

inline size_t align_up(size_t x) { return (x + 255) & ~size_t{255}; }

// The device block of a lazy-pipe key: digits of A, digits of B, s^2 TRLEV in
// the engine form, bskFirst.
struct LazyLayout {
    size_t da = 0, db = 0, s2e = 0, first = 0, total = 0;
    size_t da_bytes = 0, db_bytes = 0, s2_bytes = 0, first_bytes = 0;
};

LazyLayout lazy_layout(const ygpu_params& p, int digits) {
    LazyLayout l;
    const size_t m = static_cast<size_t>(p.n - 1), R = p.level, D = digits, K = p.k, K1 = p.k + 1, N = p.N;
    l.da_bytes = m * R * D * K * N;
    l.db_bytes = m * R * D * N;
    l.s2_bytes = D * K1 * N * sizeof(double);
    l.first_bytes = K1 * N * sizeof(int32_t);
    l.da = 0;
    l.db = align_up(l.da + l.da_bytes);
    l.s2e = align_up(l.db + l.db_bytes);
    l.first = align_up(l.s2e + l.s2_bytes);
    l.total = align_up(l.first + l.first_bytes);
    return l;
}

// The device block of an Alt key: the torus rows of every s_i, component 0's
// rows in the engine form (the first-component derivation), s^2 TRLEV in the
// engine form.
struct AltLayout {
    size_t rows = 0, row0e = 0, s2e = 0, total = 0;
    size_t rows_bytes = 0, row0_bytes = 0, s2_bytes = 0;
};

AltLayout alt_layout(const ygpu_params& p, int digits) {
    AltLayout l;
    const size_t R = p.level, D = digits, K1 = p.k + 1, N = p.N;
    l.rows_bytes = static_cast<size_t>(p.n) * R * K1 * N * sizeof(int32_t);
    l.row0_bytes = R * K1 * N * sizeof(double);
    l.s2_bytes = D * K1 * N * sizeof(double);
    l.rows = 0;
    l.row0e = align_up(l.rows + l.rows_bytes);
    l.s2e = align_up(l.row0e + l.row0_bytes);
    l.total = align_up(l.s2e + l.s2_bytes);
    return l;
}

int check_lazy_params(const ygpu_ctx* ctx, int digits) {
    const ygpu_params& p = ctx->p;
    if (!lazy_supported(ctx->tbits, ctx->logn, p.k) || p.radix_bits > 7 || p.level > 8 || p.n < 2 || p.n - 1 > KS_MAX_N)
        return fail(YGPU_ERR_UNSUPPORTED, "lazy-pipe keys are instantiated for TORUS32, N = 1024, k = 1, gadget bases up to 2^7 "
                                          "with up to 8 levels and 2 <= n <= 1024");
    if (digits < p.level || digits > 8 || digits * p.radix_bits > p.torus_bits)
        return fail(YGPU_ERR_INVALID_ARG, "lazy-pipe digit levels (param.l) must satisfy lApprox <= l <= 8 and l * radixBits <= 32");
    return YGPU_OK;
}

// The lazy-pipe bootstrapping of `count` ciphertexts on device buffers
// (context lock held): accumulators into acc (or the context's buffer),
// keyswitched LWE into d_out, extracted LWE into d_ext when asked.
int run_lazy(ygpu_ctx* ctx, int32_t key_id, int count, const int32_t* d_lwe, bool prescaled, const void* d_v, int v_stride,
             int32_t* d_out, void* d_acc_out, int32_t* d_ext) {
    const ygpu_params& p = ctx->p;
    const KeyRec& k = ctx->keys[key_id];
    const int D = k.lazy_digits;
    const LazyLayout lay = lazy_layout(p, D);
    const AltLayout alay = alt_layout(p, D);
    const int n = p.n, L = p.level, K1 = p.k + 1, N = p.N;
    const size_t accBytes = sizeof(int32_t) * K1 * N * static_cast<size_t>(count);
    void* acc = d_acc_out;
    int rc = YGPU_OK;
    if (!acc) {
        rc = ensure(&ctx->d_lazy_acc, &ctx->lazy_acc_cap, accBytes);
        if (rc) return rc;
        acc = ctx->d_lazy_acc;
    }
    rc = ensure(&ctx->d_lazy_av, &ctx->lazy_av_cap, sizeof(int32_t) * static_cast<size_t>(n - 1) * count);
    if (rc) return rc;
    const size_t rowBytes = static_cast<size_t>(L) * K1 * K1 * N * sizeof(double);  // one ciphertext, one step
    int S = ctx->lazy_window > 0 ? ctx->lazy_window
                                 : static_cast<int>(std::min<size_t>(static_cast<size_t>(n - 1), std::max<size_t>(1, LAZY_WS_BYTES / (rowBytes * count))));
    S = std::max(1, std::min(S, n - 1));
    rc = ensure(&ctx->d_lazy_ws, &ctx->lazy_ws_cap, rowBytes * S * static_cast<size_t>(count));
    if (rc) return rc;
    rc = ensure(&ctx->d_lazy_tab, &ctx->lazy_tab_cap, (sizeof(double*) + sizeof(int32_t)) * static_cast<size_t>(count));
    if (rc) return rc;

    const char* block = static_cast<const char*>(k.d_lazy);
    LazyArgs la{};
    la.count = count;
    la.n = n;
    la.level = L;
    la.dlevel = D;
    la.radix_bits = p.radix_bits;
    la.prescaled = prescaled ? 1 : 0;
    la.lwe_in = d_lwe;
    la.v = static_cast<const int32_t*>(d_v);
    la.v_stride = v_stride;
    la.acc = static_cast<int32_t*>(acc);
    la.avals = static_cast<int32_t*>(ctx->d_lazy_av);
    if (k.lazy_alt) {
        la.rows_t = reinterpret_cast<const int32_t*>(block + alay.rows);
        la.row0e = reinterpret_cast<const double*>(block + alay.row0e);
        la.s2e = reinterpret_cast<const double*>(block + alay.s2e);
    } else {
        la.first = reinterpret_cast<const int32_t*>(block + lay.first);
        la.da = reinterpret_cast<const int8_t*>(block + lay.da);
        la.db = reinterpret_cast<const int8_t*>(block + lay.db);
        la.s2e = reinterpret_cast<const double*>(block + lay.s2e);
    }
    // getNttGadgetRecomper(l) (ntt_hexl.cpp:122-149) times N^-1, centred
    const HostNtt& t = ctx->ntt[0];
    for (int l = 0; l < D; l++) {
        const uint64_t g = (uint64_t{1} << (p.torus_bits - (l + 1) * p.radix_bits)) % t.p;
        la.g[l] = centred_double(static_cast<uint64_t>((static_cast<unsigned __int128>(g) * t.ninv) % t.p), t.p);
    }
    la.ws = static_cast<double*>(ctx->d_lazy_ws);
    la.ws_steps = S;
    la.row_tab = static_cast<const double**>(ctx->d_lazy_tab);
    la.row_slot = reinterpret_cast<int32_t*>(static_cast<char*>(ctx->d_lazy_tab) + sizeof(double*) * count);
    la.tw = ctx->d_tw;
    la.prime = ctx->prime[0];
    la.ninv = centred_double(t.ninv, t.p);
    hipStream_t st = ctx->stream;
    if (k.lazy_alt) {
        // the first component, GD(v) (x) RLEV(s_0), once per test polynomial
        const int countV = v_stride ? count : 1;
        rc = ensure(&ctx->d_lazy_first, &ctx->lazy_first_cap, sizeof(int32_t) * K1 * N * static_cast<size_t>(countV));
        if (rc) return rc;
        la.first_out = static_cast<int32_t*>(ctx->d_lazy_first);
        la.first = la.first_out;
        la.first_stride = v_stride ? K1 * N : 0;
        YG_HIP(timed(ctx, "k_lazy_alt_first", st, [&] { return launch_lazy_alt_first(la, countV, st); }));
    }
    YG_HIP(timed(ctx, "k_lazy_prepare", st, [&] { return launch_lazy_prepare(la, st); }));
    for (int s0 = 0; s0 < n - 1; s0 += S) {
        la.s0 = s0;
        la.steps = std::min(S, n - 1 - s0);
        if (k.lazy_alt) YG_HIP(timed(ctx, "k_lazy_alt_derive", st, [&] { return launch_lazy_alt_derive(la, st); }));
        else YG_HIP(timed(ctx, "k_lazy_derive", st, [&] { return launch_lazy_derive(la, st); }));
        BrArgs a = br_args(ctx, count, la.avals + s0, d_v, v_stride, acc, la.row_tab, la.row_slot);
        a.n = la.steps;
        a.prescaled = 1;
        a.lwe_stride = n - 1;
        a.acc_in = acc;
        YG_HIP(timed(ctx, "k_blind_rotate_lazy", st, [&] { return launch_blind_rotate_lazy(a, st); }));
    }
    return run_ks(ctx, count, acc, ctx->d_ksk_table + key_id, nullptr, d_out, d_ext);
}

}  // namespace

extern "C" {

int ygpu_key_upload_lazy(ygpu_ctx* ctx, int32_t digit_levels, const int32_t* first, const int8_t* digits_a, const int8_t* digits_b,
                         const uint64_t* s2, const int32_t* ksk, int32_t* key_id) {
    if (!first || !digits_a || !digits_b || !s2 || !ksk || !key_id) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    return with_ctx(ctx, [&]() -> int {
        int rc = check_lazy_params(ctx, digit_levels);
        if (rc) return rc;
        const ygpu_params& p = ctx->p;
        const LazyLayout lay = lazy_layout(p, digit_levels);
        const size_t s2Polys = static_cast<size_t>(digit_levels) * (p.k + 1);
        std::vector<uint64_t> packed(s2Polys * poly_words(p) + 1, 0);
        const int64_t bad = pack_residues(s2, s2Polys, p, packed.data());
        if (bad >= 0) return fail(YGPU_ERR_INVALID_ARG, "s2 coefficient " + std::to_string(bad) + " is not below q_ntt");
        std::vector<uint32_t> planes;
        ksk_planes(p, ksk, planes);
        KeyRec rec;
        rec.live = true;
        rec.lazy = true;
        rec.lazy_digits = digit_levels;
        rec.bytes = lay.total + planes.size() * sizeof(uint32_t);
        hipError_t e = hipMalloc(&rec.d_lazy, lay.total);
        if (e != hipSuccess) return fail(YGPU_ERR_OUT_OF_MEMORY, std::string("hipMalloc lazy-pipe key: ") + hipGetErrorString(e));
        e = hipMalloc(reinterpret_cast<void**>(&rec.d_ksk), planes.size() * sizeof(uint32_t));
        if (e != hipSuccess) {
            hipFree(rec.d_lazy);
            return fail(YGPU_ERR_OUT_OF_MEMORY, std::string("hipMalloc KSK: ") + hipGetErrorString(e));
        }
        rc = ensure(&ctx->d_stage, &ctx->stage_cap, packed.size() * sizeof(uint64_t));
        if (rc) {
            hipFree(rec.d_lazy);
            hipFree(rec.d_ksk);
            return rc;
        }
        char* block = static_cast<char*>(rec.d_lazy);
        hipStream_t st = ctx->stream;
        e = hipMemcpyAsync(block + lay.da, digits_a, lay.da_bytes, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(block + lay.db, digits_b, lay.db_bytes, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(block + lay.first, first, lay.first_bytes, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(ctx->d_stage, packed.data(), packed.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st);
        // the s^2 TRLEV into the engine form (forward transform times N^-1)
        if (e == hipSuccess)
            e = launch_bsk_expand(ctx->logn, expand_args(ctx, static_cast<const uint64_t*>(ctx->d_stage), static_cast<int>(s2Polys),
                                                         reinterpret_cast<double*>(block + lay.s2e)), st);
        if (e == hipSuccess) e = hipMemcpyAsync(rec.d_ksk, planes.data(), planes.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            hipStreamSynchronize(st);
            hipFree(rec.d_lazy);
            hipFree(rec.d_ksk);
            return fail(YGPU_ERR_HIP, std::string("lazy-pipe key upload: ") + hipGetErrorString(e));
        }
        return register_key(ctx, rec, key_id);
    });
}

int ygpu_key_upload_lazy_alt(ygpu_ctx* ctx, int32_t digit_levels, const int32_t* rows, const uint64_t* s2, const int32_t* ksk,
                             int32_t* key_id) {
    if (!rows || !s2 || !ksk || !key_id) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    return with_ctx(ctx, [&]() -> int {
        int rc = check_lazy_params(ctx, digit_levels);
        if (rc) return rc;
        const ygpu_params& p = ctx->p;
        const AltLayout lay = alt_layout(p, digit_levels);
        const size_t K1 = p.k + 1, N = p.N;
        const size_t s2Polys = static_cast<size_t>(digit_levels) * K1, row0Polys = static_cast<size_t>(p.level) * K1;
        // component 0's rows as residues (torus values lifted mod q) beside s^2
        std::vector<uint64_t> res(row0Polys * N);
        for (size_t i = 0; i < res.size(); i++) res[i] = rows[i] < 0 ? static_cast<uint64_t>(static_cast<int64_t>(rows[i]) + static_cast<int64_t>(p.q_ntt))
                                                                     : static_cast<uint64_t>(rows[i]);
        const size_t pw = poly_words(p);
        std::vector<uint64_t> packed((s2Polys + row0Polys) * pw + 1, 0);
        int64_t bad = pack_residues(s2, s2Polys, p, packed.data());
        if (bad >= 0) return fail(YGPU_ERR_INVALID_ARG, "s2 coefficient " + std::to_string(bad) + " is not below q_ntt");
        pack_residues(res.data(), row0Polys, p, packed.data() + s2Polys * pw);
        std::vector<uint32_t> planes;
        ksk_planes(p, ksk, planes);
        KeyRec rec;
        rec.live = true;
        rec.lazy = true;
        rec.lazy_alt = true;
        rec.lazy_digits = digit_levels;
        rec.bytes = lay.total + planes.size() * sizeof(uint32_t);
        hipError_t e = hipMalloc(&rec.d_lazy, lay.total);
        if (e != hipSuccess) return fail(YGPU_ERR_OUT_OF_MEMORY, std::string("hipMalloc lazy-pipe key: ") + hipGetErrorString(e));
        e = hipMalloc(reinterpret_cast<void**>(&rec.d_ksk), planes.size() * sizeof(uint32_t));
        if (e != hipSuccess) {
            hipFree(rec.d_lazy);
            return fail(YGPU_ERR_OUT_OF_MEMORY, std::string("hipMalloc KSK: ") + hipGetErrorString(e));
        }
        rc = ensure(&ctx->d_stage, &ctx->stage_cap, packed.size() * sizeof(uint64_t));
        if (rc) {
            hipFree(rec.d_lazy);
            hipFree(rec.d_ksk);
            return rc;
        }
        char* block = static_cast<char*>(rec.d_lazy);
        hipStream_t st = ctx->stream;
        const uint64_t* stage = static_cast<const uint64_t*>(ctx->d_stage);
        e = hipMemcpyAsync(block + lay.rows, rows, lay.rows_bytes, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(ctx->d_stage, packed.data(), packed.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = launch_bsk_expand(ctx->logn, expand_args(ctx, stage, static_cast<int>(s2Polys), reinterpret_cast<double*>(block + lay.s2e)), st);
        if (e == hipSuccess)
            e = launch_bsk_expand(ctx->logn, expand_args(ctx, stage + s2Polys * pw, static_cast<int>(row0Polys),
                                                         reinterpret_cast<double*>(block + lay.row0e)), st);
        if (e == hipSuccess) e = hipMemcpyAsync(rec.d_ksk, planes.data(), planes.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            hipStreamSynchronize(st);
            hipFree(rec.d_lazy);
            hipFree(rec.d_ksk);
            return fail(YGPU_ERR_HIP, std::string("lazy-pipe key upload: ") + hipGetErrorString(e));
        }
        return register_key(ctx, rec, key_id);
    });
}

int ygpu_set_lazy_window(ygpu_ctx* ctx, int32_t steps) {
    if (steps < 0) return fail(YGPU_ERR_INVALID_ARG, "negative window");
    return with_ctx(ctx, [&]() -> int {
        ctx->lazy_window = steps;
        return YGPU_OK;
    });
}

int ygpu_pbs_lazy_batch_device(ygpu_ctx* ctx, int32_t key_id, int32_t count, const int32_t* d_lwe_in, const void* d_v,
                               int32_t v_per_ct, int32_t* d_lwe_out, void* d_acc_out, int32_t* d_extract_out) {
    if (count < 0) return fail(YGPU_ERR_INVALID_ARG, "negative count");
    if (count == 0) return YGPU_OK;
    if (!d_lwe_in || !d_v || !d_lwe_out) return fail(YGPU_ERR_INVALID_ARG, "null buffer");
    return with_ctx(ctx, [&]() -> int {
        int rc = lazy_key_live(ctx, key_id);
        if (rc) return rc;
        return run_lazy(ctx, key_id, count, d_lwe_in, false, d_v, v_per_ct ? ctx->p.N : 0, d_lwe_out, d_acc_out, d_extract_out);
    });
}

// Host buffers, one ciphertext: the raw LWE (functionalBootstrapping's input)
// or, with `prescaled`, the rescaled ScaledTlwe of blindRotateLazyPipeNtt.
static int lazy_host(ygpu_ctx* ctx, int32_t key_id, const int32_t* lwe_in, bool prescaled, const void* v, int32_t* lwe_out,
                     void* acc_out, int32_t* extract_out) {
    if (!lwe_in || !v) return fail(YGPU_ERR_INVALID_ARG, "null buffer");
    return with_ctx(ctx, [&]() -> int {
        int rc = lazy_key_live(ctx, key_id);
        if (rc) return rc;
        rc = plain_buffers(ctx);
        if (rc) return rc;
        const ygpu_params& p = ctx->p;
        const size_t lweBytes = sizeof(int32_t) * (p.n + 1), vBytes = ctx->torus_bytes() * p.N;
        const size_t accBytes = ctx->torus_bytes() * (p.k + 1) * p.N, extBytes = sizeof(int32_t) * (p.k * p.N + 1);
        hipStream_t st = ctx->stream;
        YG_HIP(hipMemcpyAsync(ctx->d_lwe1, lwe_in, lweBytes, hipMemcpyHostToDevice, st));
        YG_HIP(hipMemcpyAsync(ctx->d_v1, v, vBytes, hipMemcpyHostToDevice, st));
        rc = run_lazy(ctx, key_id, 1, ctx->d_lwe1, prescaled, ctx->d_v1, 0, ctx->d_out1, ctx->d_acc1, extract_out ? ctx->d_ext1 : nullptr);
        if (rc) return rc;
        if (lwe_out) YG_HIP(hipMemcpyAsync(lwe_out, ctx->d_out1, lweBytes, hipMemcpyDeviceToHost, st));
        if (acc_out) YG_HIP(hipMemcpyAsync(acc_out, ctx->d_acc1, accBytes, hipMemcpyDeviceToHost, st));
        if (extract_out) YG_HIP(hipMemcpyAsync(extract_out, ctx->d_ext1, extBytes, hipMemcpyDeviceToHost, st));
        return sync_and_check(ctx);
    });
}

int ygpu_pbs_lazy(ygpu_ctx* ctx, int32_t key_id, const int32_t* lwe_in, const void* v, int32_t* lwe_out, void* acc_out,
                  int32_t* extract_out) {
    if (!lwe_out) return fail(YGPU_ERR_INVALID_ARG, "null buffer");
    return lazy_host(ctx, key_id, lwe_in, false, v, lwe_out, acc_out, extract_out);
}

int ygpu_blind_rotate_lazy(ygpu_ctx* ctx, int32_t key_id, const int32_t* abar, const void* v, void* acc_out) {
    if (!acc_out) return fail(YGPU_ERR_INVALID_ARG, "null buffer");
    return lazy_host(ctx, key_id, abar, true, v, nullptr, acc_out, nullptr);
}

}  // extern "C"
