// ygpu_runtime.hip -- host runtime behind the C ABI (include/yatfhe_gpu.h).
//
// Owns the device, its stream, the transform tables and the resident keys.
// Keys arrive in coefficient form; the host transforms each key polynomial
// exactly (128-bit integer NTT) into the engine's order and layout, folds 1/N
// in, and uploads.  Every entry point returns a status code and records a
// per-thread message; nothing throws across the boundary.
#include "ygpu_internal.h"

using namespace ygpu_rt;

// Scratch of ygpu_pbs_batch_device_noalloc: the accumulators and the keyswitch
// workspace of one call, in a static device array (each device holds its own
// copy from module load until the process ends).  In host code `scratch` only
// names the symbol: kernels get the device address hipGetSymbolAddress returns
// (ygpu_ctx::d_static_scratch), never the array's name.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t{96} << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

// The static scratch's device address on `device` (current), looked up once
// per device and kept.
void* static_scratch_address(int device) {
    static std::mutex mu;
    static std::map<int, void*> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(device);
    if (it != addr.end()) return it->second;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) {
        (void)hipGetLastError();
        ptr = nullptr;
    }
    addr[device] = ptr;
    return ptr;
}

// Scratch bytes of a no-allocation batch: the accumulators (unless the caller
// supplies them), then the keyswitch workspace, each 256-byte aligned.
size_t noalloc_bytes(const ygpu_ctx* ctx, int count, bool acc_given, size_t* acc_bytes) {
    const size_t acc = acc_given ? 0 : (ctx->torus_bytes() * (ctx->p.k + 1) * ctx->p.N * static_cast<size_t>(count) + 255) / 256 * 256;
    *acc_bytes = acc;
    return acc + ks_workspace_bytes(ctx, count);
}

}  // namespace

// ======================================================================= C ABI
extern "C" {

const char* ygpu_last_error(void) { return g_err.c_str(); }

int ygpu_abi_struct_sizes(int64_t* out, int32_t max) {
    const int64_t sizes[] = {sizeof(ygpu_params), sizeof(ygpu_job), sizeof(ygpu_server_config), sizeof(ygpu_request),
                             sizeof(ygpu_server_stats), sizeof(ygpu_loadgen_config), sizeof(ygpu_loadgen_result)};
    const int n = static_cast<int>(sizeof(sizes) / sizeof(sizes[0]));
    if (out)
        for (int i = 0; i < n && i < max; i++) out[i] = sizes[i];
    return n;
}

int ygpu_device_count(int32_t* count) {
    if (!count) return fail(YGPU_ERR_INVALID_ARG, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) { *count = 0; return fail(YGPU_ERR_NO_DEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e)); }
    *count = c;
    return YGPU_OK;
}

}  // extern "C"

namespace ygpu_rt {

int build_context(ygpu_ctx* ctx, const ygpu_params* params, int32_t device, hipStream_t stream);

int init_context(const ygpu_params* params, int32_t device, hipStream_t stream, ygpu_ctx** out) {
    if (!params || !out) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    int rc = check_params(*params);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(YGPU_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(YGPU_ERR_INVALID_ARG, "device index out of range");
    ygpu_ctx* ctx = new ygpu_ctx();
    rc = build_context(ctx, params, device, stream);
    if (rc) {
        const std::string msg = ygpu_last_error();
        ygpu_destroy(ctx);  // frees what was built
        return fail(rc, msg);
    }
    *out = ctx;
    return YGPU_OK;
}

int build_context(ygpu_ctx* ctx, const ygpu_params* params, int32_t device, hipStream_t stream) {
    ctx->p = *params;
    ctx->device = device;
    ctx->tbits = params->torus_bits;
    ctx->logn = params->N == 1024 ? 10 : 11;
    if (hipDeviceGetAttribute(&ctx->cu_count, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) ctx->cu_count = 0;
    const uint64_t q = params->q_ntt;
    if (q < (1ULL << 49)) {
        ctx->np = 1;
        ctx->primes[0] = q;
    } else {
        ctx->np = 2;
        ctx->primes[0] = P_AUX_0;
        ctx->primes[1] = P_AUX_1;
    }
    // default tile: small workgroups (a barrier couples few waves), twiddles in
    // LDS where they fit; TORUS32 N = 1024: two ciphertexts per workgroup and
    // the first two forward stages read from digit tables (45.9 ms against
    // 48.7 ms without tables for 4096 ciphertexts at n = 630, 2^7 x 3);
    // TORUS56 N = 2048, 2^10 x 3: stage 0 from a digit table (562.9 against
    // 569.4 ms for 8192 ciphertexts at n = 742)
    const bool t32 = params->torus_bits == 32 && params->N == 1024;
    const bool t56s = params->torus_bits == 56 && params->N == 2048 && params->level == 3 && params->radix_bits == 10;
    ctx->tile = BrTile{t32 ? 2 : 1, (t32 || params->torus_bits == 56) ? 1 : 0, 0, t32 ? 2 : t56s ? 1 : 0};
    if (const char* e = std::getenv("YGPU_CPW")) ctx->tile.cpw = std::max(1, std::atoi(e));
    YG_HIP(hipSetDevice(device));
    if (stream) {
        ctx->stream = stream;
        ctx->owns_stream = false;
    } else {
        YG_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    }
    const int N = params->N;
    std::vector<double> tw(static_cast<size_t>(ctx->np) * 2 * N);
    for (int k = 0; k < ctx->np; k++) {
        ctx->ntt.emplace_back(ctx->primes[k], N);
        const HostNtt& t = ctx->ntt.back();
        for (int i = 0; i < N; i++) {
            tw[static_cast<size_t>(k) * 2 * N + i] = centred_double(t.psi[i], t.p);
            tw[static_cast<size_t>(k) * 2 * N + N + i] = centred_double(t.ipsi[i], t.p);
        }
        ctx->prime[k].p = static_cast<double>(t.p);
        ctx->prime[k].pinv = 1.0 / static_cast<double>(t.p);
        ctx->prime[k].half = static_cast<double>((t.p - 1) / 2);
    }
    if (ctx->np == 2) {
        const uint64_t p0 = ctx->primes[0], p1 = ctx->primes[1];
        ctx->crt.p0 = static_cast<double>(p0);
        ctx->crt.inv01 = centred_double(invmod(p0 % p1, p1), p1);
        ctx->crt.p0i = static_cast<int64_t>(p0);
        ctx->crt.q = q;
        ctx->crt.qd = static_cast<double>(q);
        ctx->crt.qinv = 1.0 / static_cast<double>(q);
        ctx->crt.qhalf = static_cast<int64_t>((q - 1) / 2);
    }
    YG_HIP(hipMalloc(reinterpret_cast<void**>(&ctx->d_tw), sizeof(double) * tw.size()));
    // copies on the context's stream: no null-stream work, which would hold
    // one of the device's hardware queues for the rest of the process
    YG_HIP(hipMemcpyAsync(ctx->d_tw, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice, ctx->stream));
    // key tables of fixed capacity, all slots empty
    std::vector<const void*> empty(KEY_SLOTS, nullptr);
    YG_HIP(hipMalloc(reinterpret_cast<void**>(&ctx->d_bsk_table), sizeof(void*) * KEY_SLOTS));
    YG_HIP(hipMalloc(reinterpret_cast<void**>(&ctx->d_ksk_table), sizeof(void*) * KEY_SLOTS));
    YG_HIP(hipMemcpyAsync(ctx->d_bsk_table, empty.data(), sizeof(void*) * KEY_SLOTS, hipMemcpyHostToDevice, ctx->stream));
    YG_HIP(hipMemcpyAsync(ctx->d_ksk_table, empty.data(), sizeof(void*) * KEY_SLOTS, hipMemcpyHostToDevice, ctx->stream));
    YG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->d_static_scratch = static_scratch_address(device);
    YG_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->h_status), 64, hipHostMallocDefault));
    *ctx->h_status = 0u;
    {
        void* dp = nullptr;
        YG_HIP(hipHostGetDevicePointer(&dp, ctx->h_status, 0));
        ctx->hd_status = static_cast<uint32_t*>(dp);
    }
    return YGPU_OK;
}

}  // namespace ygpu_rt

extern "C" {

int ygpu_init(const ygpu_params* params, int32_t device, ygpu_ctx** out) {
    return init_context(params, device, nullptr, out);
}

int ygpu_destroy(ygpu_ctx* ctx) {
    if (!ctx) return YGPU_OK;
    {
        std::unique_lock<std::mutex> lk(ctx->mu);
        hipSetDevice(ctx->device);
        ctx->cv.wait(lk, [&] {
            for (auto& l : ctx->lanes)
                if (l->busy) return false;
            return true;
        });
        for (auto& l : ctx->lanes) free_lane(*l);
        ctx->lanes.clear();
        if (ctx->stream) hipStreamSynchronize(ctx->stream);
            for (auto& k : ctx->keys) {
            if (k.d_bsk) hipFree(k.d_bsk);
            if (k.d_ksk) hipFree(k.d_ksk);
            if (k.d_lazy) hipFree(k.d_lazy);
        }
        for (auto& g : ctx->graphs) if (g.live) free_graph(g);
        for (void* p : {static_cast<void*>(ctx->d_tw), static_cast<void*>(ctx->d_bsk_table), static_cast<void*>(ctx->d_ksk_table),
                        ctx->d_acc, static_cast<void*>(ctx->d_partial), static_cast<void*>(ctx->d_slots),
                        static_cast<void*>(ctx->d_lwe1), ctx->d_v1, static_cast<void*>(ctx->d_out1),
                        static_cast<void*>(ctx->d_ext1), ctx->d_acc1, ctx->d_accin1, ctx->d_stage, ctx->d_lazy_ws,
                        ctx->d_lazy_av, ctx->d_lazy_acc, ctx->d_lazy_tab, ctx->d_lazy_first})
            if (p) hipFree(p);
        if (ctx->h_status) hipHostFree(ctx->h_status);
        for (auto& t : ctx->pending) { hipEventDestroy(t.a); hipEventDestroy(t.b); }
        for (auto e : ctx->event_pool) hipEventDestroy(e);
        if (ctx->owns_stream && ctx->stream) hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return YGPU_OK;
}

int ygpu_get_params(const ygpu_ctx* ctx, ygpu_params* out) {
    if (!ctx || !out) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    *out = ctx->p;
    return YGPU_OK;
}

}  // extern "C"

namespace ygpu_rt {

// Gives an uploaded key a slot of the fixed-capacity key tables; frees its
// device memory on failure.
int register_key(ygpu_ctx* ctx, const KeyRec& rec, int32_t* key_id) {
    auto drop = [&](const KeyRec& k) {
        if (k.d_bsk) hipFree(k.d_bsk);
        if (k.d_ksk) hipFree(k.d_ksk);
        if (k.d_lazy) hipFree(k.d_lazy);
    };
    int id = -1;
    for (size_t i = 0; i < ctx->keys.size(); i++) if (!ctx->keys[i].live) { id = static_cast<int>(i); break; }
    if (id < 0) {
        if (static_cast<int>(ctx->keys.size()) >= KEY_SLOTS) {
            drop(rec);
            return fail(YGPU_ERR_UNSUPPORTED, "a context holds at most " + std::to_string(KEY_SLOTS) + " keys");
        }
        id = static_cast<int>(ctx->keys.size());
        ctx->keys.push_back(rec);
    } else {
        ctx->keys[id] = rec;
    }
    const int rc = set_table_slot(ctx, id);
    if (rc) {
        drop(ctx->keys[id]);
        ctx->keys[id] = KeyRec{};
        return rc;
    }
    *key_id = id;
    return YGPU_OK;
}

}  // namespace ygpu_rt

extern "C" {

namespace {

// Makes a host key form resident: device BSK + KSK, the packed residues staged
// and expanded on the context stream.  Called with the context lock held.
int upload_host_form(ygpu_ctx* ctx, const HostKeyForm& hk, int32_t* key_id) {
    KeyRec rec;
    rec.live = true;
    rec.bytes = hk.bytes();
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&rec.d_bsk), hk.bsk_bytes);
    if (e != hipSuccess) return fail(YGPU_ERR_OUT_OF_MEMORY, std::string("hipMalloc BSK: ") + hipGetErrorString(e));
    e = hipMalloc(reinterpret_cast<void**>(&rec.d_ksk), hk.planes.size() * sizeof(uint32_t));
    if (e != hipSuccess) { hipFree(rec.d_bsk); return fail(YGPU_ERR_OUT_OF_MEMORY, std::string("hipMalloc KSK: ") + hipGetErrorString(e)); }
    // packed residues -> staging -> engine form, on the context stream
    int rc = ensure(&ctx->d_stage, &ctx->stage_cap, stage_bytes(ctx->p, ctx->np, hk.s2_level));
    if (rc) { hipFree(rec.d_bsk); hipFree(rec.d_ksk); return rc; }
    e = hipMemcpyAsync(ctx->d_stage, hk.packed.data(), hk.packed.size() * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream);
    // timed as "k_bsk_expand" (full keys) or "succinct_expand" (the s^2 TRLEV's
    // k_bsk_expand + k_bsk_derive)
    if (e == hipSuccess)
        e = timed(ctx, hk.s2_level ? "succinct_expand" : "k_bsk_expand", ctx->stream,
                  [&] { return expand_key(ctx, hk.s2_level, ctx->d_stage, rec.d_bsk, ctx->stream); });
    if (e == hipSuccess)
        e = hipMemcpyAsync(rec.d_ksk, hk.planes.data(), hk.planes.size() * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        hipStreamSynchronize(ctx->stream);
        hipFree(rec.d_bsk);
        hipFree(rec.d_ksk);
        return fail(YGPU_ERR_HIP, std::string("key upload: ") + hipGetErrorString(e));
    }
    return register_key(ctx, rec, key_id);
}

}  // namespace

int ygpu_key_upload(ygpu_ctx* ctx, const uint64_t* bsk_coeff, const int32_t* ksk, int32_t* key_id) {
    if (!bsk_coeff || !ksk || !key_id) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    return with_ctx(ctx, [&]() -> int {
        HostKeyForm hk;
        int rc = prepare_key_host(ctx, bsk_coeff, ksk, hk);
        if (rc) return rc;
        return upload_host_form(ctx, hk, key_id);
    });
}

int ygpu_key_upload_succinct(ygpu_ctx* ctx, const uint64_t* bsk_half, const uint64_t* s2, int32_t s2_level, const int32_t* ksk,
                             int32_t* key_id) {
    if (!bsk_half || !s2 || !ksk || !key_id) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    return with_ctx(ctx, [&]() -> int {
        HostKeyForm hk;
        int rc = prepare_key_host_succinct(ctx, bsk_half, s2, s2_level, ksk, hk);
        if (rc) return rc;
        return upload_host_form(ctx, hk, key_id);
    });
}

int ygpu_key_release(ygpu_ctx* ctx, int32_t key_id) {
    if (!ctx) return fail(YGPU_ERR_INVALID_ARG, "null context");
    std::unique_lock<std::mutex> lk(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail(YGPU_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    int rc = key_live(ctx, key_id);
    if (rc) return rc;
    for (size_t g = 0; g < ctx->graphs.size(); g++)
        if (ctx->graphs[g].live && std::find(ctx->graphs[g].keys.begin(), ctx->graphs[g].keys.end(), key_id) != ctx->graphs[g].keys.end())
            return fail(YGPU_ERR_BUSY, "key " + std::to_string(key_id) + " is used by graph " + std::to_string(g) + ": destroy it first");
    // synchronous calls on plain lanes that use the key finish first
    ctx->cv.wait(lk, [&] { return ctx->keys[key_id].users == 0; });
    rc = key_live(ctx, key_id);  // another thread may have released it meanwhile
    if (rc) return rc;
    YG_HIP(hipStreamSynchronize(ctx->stream));
    KeyRec& k = ctx->keys[key_id];
    if (k.d_bsk) hipFree(k.d_bsk);
    if (k.d_ksk) hipFree(k.d_ksk);
    if (k.d_lazy) hipFree(k.d_lazy);
    k = KeyRec{};
    return set_table_slot(ctx, key_id);
}

int ygpu_key_bytes(const ygpu_ctx* ctx, int32_t key_id, int64_t* bytes) {
    if (!ctx || !bytes) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    int rc = key_live(ctx, key_id);
    if (rc) return rc;
    *bytes = static_cast<int64_t>(ctx->keys[key_id].bytes);
    return YGPU_OK;
}

int ygpu_pbs_batch_device(ygpu_ctx* ctx, int32_t key_id, int32_t count, const int32_t* d_lwe_in, const void* d_v,
                          int32_t v_per_ct, int32_t* d_lwe_out, void* d_acc_out, int32_t* d_extract_out) {
    if (count < 0) return fail(YGPU_ERR_INVALID_ARG, "negative count");
    if (count == 0) return YGPU_OK;
    if (!d_lwe_in || !d_v || !d_lwe_out) return fail(YGPU_ERR_INVALID_ARG, "null buffer");
    return with_ctx(ctx, [&]() -> int {
        int rc = engine_key_live(ctx, key_id);
        if (rc) return rc;
        void* acc = d_acc_out;
        if (!acc) {
            rc = ensure(&ctx->d_acc, &ctx->acc_cap, ctx->torus_bytes() * (ctx->p.k + 1) * ctx->p.N * static_cast<size_t>(count));
            if (rc) return rc;
            acc = ctx->d_acc;
        }
        BrArgs a = br_args(ctx, count, d_lwe_in, d_v, v_per_ct ? ctx->p.N : 0, acc, ctx->d_bsk_table + key_id, nullptr);
        rc = run_br(ctx, batch_tile(ctx, count), a);
        if (rc) return rc;
        return run_ks(ctx, count, acc, ctx->d_ksk_table + key_id, nullptr, d_lwe_out, d_extract_out);
    });
}

int ygpu_pbs_batch_device_noalloc(ygpu_ctx* ctx, int32_t key_id, int32_t count, const int32_t* d_lwe_in, const void* d_v,
                                  int32_t v_per_ct, int32_t* d_lwe_out, void* d_acc_out, int32_t* d_extract_out) {
    if (count < 0) return fail(YGPU_ERR_INVALID_ARG, "negative count");
    if (count == 0) return YGPU_OK;
    if (!d_lwe_in || !d_v || !d_lwe_out) return fail(YGPU_ERR_INVALID_ARG, "null buffer");
    return with_ctx(ctx, [&]() -> int {
        int rc = engine_key_live(ctx, key_id);
        if (rc) return rc;
        if (!ctx->d_static_scratch) return fail(YGPU_ERR_HIP, "the static scratch array has no device address on this device");
        size_t accBytes = 0;
        const size_t need = noalloc_bytes(ctx, count, d_acc_out != nullptr, &accBytes);
        if (need > NOALLOC_SCRATCH_BYTES) {
            size_t fit = 0;
            for (size_t lo = 1, hi = static_cast<size_t>(count); lo <= hi;) {
                const size_t mid = (lo + hi) / 2;
                size_t a = 0;
                if (noalloc_bytes(ctx, static_cast<int>(mid), d_acc_out != nullptr, &a) <= NOALLOC_SCRATCH_BYTES) { fit = mid; lo = mid + 1; }
                else hi = mid - 1;
            }
            return fail(YGPU_ERR_INVALID_ARG, "ygpu_pbs_batch_device_noalloc: " + std::to_string(count) + " ciphertexts need " +
                                                  std::to_string(need) + " bytes of scratch; the static scratch holds " +
                                                  std::to_string(NOALLOC_SCRATCH_BYTES) + " (at most " + std::to_string(fit) +
                                                  " ciphertexts of this parameter set per call)");
        }
        char* base = static_cast<char*>(ctx->d_static_scratch);
        void* acc = d_acc_out ? d_acc_out : base;
        int32_t* workspace = reinterpret_cast<int32_t*>(base + accBytes);
        BrArgs a = br_args(ctx, count, d_lwe_in, d_v, v_per_ct ? ctx->p.N : 0, acc, ctx->d_bsk_table + key_id, nullptr);
        // no timing events here: taking one may create it
        YG_HIP(launch_blind_rotate(ctx->tbits, ctx->logn, batch_tile(ctx, count), ctx->p.radix_bits, a, ctx->stream));
        KsArgs k = ks_args(ctx, count, acc, ctx->d_ksk_table + key_id, nullptr, d_lwe_out, workspace);
        YG_HIP(launch_extract_ks(ctx->tbits, k, ks_slices(ctx, count), ctx->stream));
        if (d_extract_out) YG_HIP(launch_extract(ctx->tbits, k, d_extract_out, ctx->stream));
        return YGPU_OK;
    });
}

int ygpu_pbs_jobs_device(ygpu_ctx* ctx, const ygpu_job* jobs, int32_t n_jobs, int64_t total, const int32_t* d_lwe_in,
                         const void* d_v, int32_t v_per_ct, int32_t* d_lwe_out, void* d_acc_out) {
    if (!jobs || n_jobs < 0 || total < 0 || total > (1LL << 30)) return fail(YGPU_ERR_INVALID_ARG, "bad job list");
    if (total == 0) return YGPU_OK;
    if (!d_lwe_in || !d_v || !d_lwe_out) return fail(YGPU_ERR_INVALID_ARG, "null buffer");
    return with_ctx(ctx, [&]() -> int {
        std::vector<int32_t> slots;
        int rc = job_slots(ctx, jobs, n_jobs, total, slots);
        if (rc) return rc;
        rc = ensure(reinterpret_cast<void**>(&ctx->d_slots), &ctx->slots_cap, sizeof(int32_t) * slots.size());
        if (rc) return rc;
        YG_HIP(hipMemcpyAsync(ctx->d_slots, slots.data(), sizeof(int32_t) * slots.size(), hipMemcpyHostToDevice, ctx->stream));
        void* acc = d_acc_out;
        if (!acc) {
            rc = ensure(&ctx->d_acc, &ctx->acc_cap, ctx->torus_bytes() * (ctx->p.k + 1) * ctx->p.N * static_cast<size_t>(total));
            if (rc) return rc;
            acc = ctx->d_acc;
        }
        rc = ensure(reinterpret_cast<void**>(&ctx->d_partial), &ctx->partial_cap, ks_workspace_bytes(ctx, static_cast<int>(total)));
        if (rc) return rc;
        rc = enqueue_jobs(ctx, static_cast<int>(total), d_lwe_in, d_v, v_per_ct, d_lwe_out, acc, ctx->d_slots, ctx->d_partial);
        if (rc) return rc;
        // the host slot vector dies with this call: finish the copy first
        return sync_and_check(ctx);
    });
}

int ygpu_pbs_plain_device(ygpu_ctx* ctx, int32_t key_id, const int32_t* d_lwe_in, const void* d_v, int32_t* d_lwe_out,
                          void* d_acc_out) {
    if (!d_lwe_in || !d_v || !d_lwe_out) return fail(YGPU_ERR_INVALID_ARG, "null buffer");
    return with_ctx(ctx, [&]() -> int {
        int rc = engine_key_live(ctx, key_id);
        if (rc) return rc;
        rc = plain_buffers(ctx);
        if (rc) return rc;
        void* acc = d_acc_out ? d_acc_out : ctx->d_acc1;
        BrArgs a = br_args(ctx, 1, d_lwe_in, d_v, 0, acc, ctx->d_bsk_table + key_id, nullptr);
        rc = run_br(ctx, plain_tile(ctx), a);
        if (rc) return rc;
        return run_ks(ctx, 1, acc, ctx->d_ksk_table + key_id, nullptr, d_lwe_out, nullptr);
    });
}

}  // extern "C"

namespace {

int lane_create(ygpu_ctx* ctx, std::unique_ptr<PlainLane>& out) {
    auto L = std::make_unique<PlainLane>();
    const ygpu_params& p = ctx->p;
    const size_t tb = ctx->torus_bytes();
    hipError_t e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&L->d_lwe), sizeof(int32_t) * (p.n + 1));
    if (e == hipSuccess) e = hipMalloc(&L->d_v, tb * p.N);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&L->d_out), sizeof(int32_t) * (p.n + 1));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&L->d_ext), sizeof(int32_t) * (p.k * p.N + 1));
    if (e == hipSuccess) e = hipMalloc(&L->d_acc, tb * (p.k + 1) * p.N);
    if (e == hipSuccess) e = hipMalloc(&L->d_accin, tb * (p.k + 1) * p.N);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&L->d_partial), ks_workspace_bytes(ctx, 1));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&L->h_status), 64, hipHostMallocDefault);
    void* dp = nullptr;
    if (e == hipSuccess) {
        *L->h_status = 0u;
        e = hipHostGetDevicePointer(&dp, L->h_status, 0);
    }
    if (e != hipSuccess) {
        free_lane(*L);
        return fail(YGPU_ERR_OUT_OF_MEMORY, std::string("plain lane: ") + hipGetErrorString(e));
    }
    L->hd_status = static_cast<uint32_t*>(dp);
    out = std::move(L);
    return YGPU_OK;
}

// Runs f(lane) on a plain lane of its own, holding the context lock only to
// take and return the lane (and to count the call as a user of key_id, so the
// key cannot be released under it).  Concurrent callers run side by side.
template <class F>
int with_lane(ygpu_ctx* ctx, int32_t key_id, F&& f) {
    if (!ctx) return fail(YGPU_ERR_INVALID_ARG, "null context");
    PlainLane* L = nullptr;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        hipError_t e = hipSetDevice(ctx->device);
        if (e != hipSuccess) return fail(YGPU_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
        if (key_id >= 0) {
            int rc = engine_key_live(ctx, key_id);
            if (rc) return rc;
        }
        for (auto& l : ctx->lanes)
            if (!l->busy) { L = l.get(); break; }
        if (!L) {
            std::unique_ptr<PlainLane> nl;
            int rc = lane_create(ctx, nl);
            if (rc) return rc;
            L = nl.get();
            ctx->lanes.push_back(std::move(nl));
        }
        L->busy = true;
        L->tile = plain_tile(ctx);
        L->extract_index = ctx->extract_index;
        L->barrier_polls = ctx->barrier_polls;
        if (key_id >= 0) ctx->keys[key_id].users++;
    }
    int rc;
    try {
        rc = f(*L);
    } catch (const std::exception& ex) {
        rc = fail(YGPU_ERR_INVALID_ARG, ex.what());
    }
    // nothing of this call may still run on the lane when another takes it
    hipStreamSynchronize(L->stream);
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        L->busy = false;
        if (key_id >= 0) ctx->keys[key_id].users--;
    }
    ctx->cv.notify_all();
    return rc;
}

// Synchronises the lane's stream and reports its barrier-failure word.
int lane_sync(PlainLane& L) {
    YG_HIP(hipStreamSynchronize(L.stream));
    volatile uint32_t* st = L.h_status;
    if (*st == 0u) return YGPU_OK;
    *st = 0u;
    return fail(YGPU_ERR_DEVICE, "a blind-rotation workgroup's team barrier timed out: the call's results are invalid "
                                 "(its accumulator reads the most negative torus value)");
}

BrArgs lane_br_args(ygpu_ctx* ctx, PlainLane& L, const int32_t* d_lwe, const void* d_v, void* d_acc, int32_t key_id) {
    BrArgs a = br_args(ctx, 1, d_lwe, d_v, 0, d_acc, ctx->d_bsk_table + key_id, nullptr);
    a.status = L.hd_status;
    a.barrier_polls = L.barrier_polls;
    return a;
}

// Host accumulator Torus[(k+1)N] -> the lane's device accumulator buffer `dst`.
hipError_t lane_upload(ygpu_ctx* ctx, PlainLane& L, void* dst, const void* src, size_t bytes) {
    (void)ctx;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, L.stream);
}

// Extract (at `index`) + keyswitch of the accumulator at L.d_acc into L.d_out
// (and the extracted LWE into L.d_ext when want_ext).
int lane_extract_keyswitch(ygpu_ctx* ctx, PlainLane& L, int32_t key_id, bool want_ext, int index) {
    return run_ks(ctx, 1, L.d_acc, ctx->d_ksk_table + key_id, nullptr, L.d_out, want_ext ? L.d_ext : nullptr, L.d_partial, L.stream,
                  index);
}

}  // namespace

extern "C" {

int ygpu_pbs(ygpu_ctx* ctx, int32_t key_id, const int32_t* lwe_in, const void* v, int32_t* lwe_out, void* acc_out,
             int32_t* extract_out) {
    if (!lwe_in || !v || !lwe_out) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    return with_lane(ctx, key_id, [&](PlainLane& L) -> int {
        const ygpu_params& p = ctx->p;
        const size_t tb = ctx->torus_bytes();
        YG_HIP(hipMemcpyAsync(L.d_lwe, lwe_in, sizeof(int32_t) * (p.n + 1), hipMemcpyHostToDevice, L.stream));
        YG_HIP(hipMemcpyAsync(L.d_v, v, tb * p.N, hipMemcpyHostToDevice, L.stream));
        int rc = run_br(ctx, L.tile, lane_br_args(ctx, L, L.d_lwe, L.d_v, L.d_acc, key_id), L.stream);
        if (rc) return rc;
        rc = lane_extract_keyswitch(ctx, L, key_id, extract_out != nullptr, L.extract_index);
        if (rc) return rc;
        YG_HIP(hipMemcpyAsync(lwe_out, L.d_out, sizeof(int32_t) * (p.n + 1), hipMemcpyDeviceToHost, L.stream));
        if (acc_out) YG_HIP(hipMemcpyAsync(acc_out, L.d_acc, tb * (p.k + 1) * p.N, hipMemcpyDeviceToHost, L.stream));
        if (extract_out) YG_HIP(hipMemcpyAsync(extract_out, L.d_ext, sizeof(int32_t) * (p.k * p.N + 1), hipMemcpyDeviceToHost, L.stream));
        return lane_sync(L);
    });
}

int ygpu_blind_rotate(ygpu_ctx* ctx, int32_t key_id, const int32_t* lwe_in, const void* v, void* acc_out) {
    if (!lwe_in || !v || !acc_out) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    return with_lane(ctx, key_id, [&](PlainLane& L) -> int {
        const ygpu_params& p = ctx->p;
        const size_t tb = ctx->torus_bytes();
        YG_HIP(hipMemcpyAsync(L.d_lwe, lwe_in, sizeof(int32_t) * (p.n + 1), hipMemcpyHostToDevice, L.stream));
        YG_HIP(hipMemcpyAsync(L.d_v, v, tb * p.N, hipMemcpyHostToDevice, L.stream));
        int rc = run_br(ctx, L.tile, lane_br_args(ctx, L, L.d_lwe, L.d_v, L.d_acc, key_id), L.stream);
        if (rc) return rc;
        YG_HIP(hipMemcpyAsync(acc_out, L.d_acc, tb * (p.k + 1) * p.N, hipMemcpyDeviceToHost, L.stream));
        return lane_sync(L);
    });
}

int ygpu_blind_rotate_scaled(ygpu_ctx* ctx, int32_t key_id, const int32_t* abar, void* acc) {
    if (!abar || !acc) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    if (!ctx) return fail(YGPU_ERR_INVALID_ARG, "null context");
    const int twoN = 2 * ctx->p.N;
    for (int i = 0; i < ctx->p.n; i++)
        if (abar[i] < -twoN || abar[i] >= twoN) return fail(YGPU_ERR_INVALID_ARG, "abar out of [-2N, 2N)");
    return with_lane(ctx, key_id, [&](PlainLane& L) -> int {
        const ygpu_params& p = ctx->p;
        const size_t tb = ctx->torus_bytes();
        YG_HIP(hipMemcpyAsync(L.d_lwe, abar, sizeof(int32_t) * p.n, hipMemcpyHostToDevice, L.stream));
        // the caller's initial accumulator, in a buffer of its own
        YG_HIP(lane_upload(ctx, L, L.d_accin, acc, tb * (p.k + 1) * p.N));
        BrArgs a = lane_br_args(ctx, L, L.d_lwe, L.d_v, L.d_acc, key_id);
        a.prescaled = 1;
        a.acc_in = L.d_accin;
        int rc = run_br(ctx, L.tile, a, L.stream);
        if (rc) return rc;
        YG_HIP(hipMemcpyAsync(acc, L.d_acc, tb * (p.k + 1) * p.N, hipMemcpyDeviceToHost, L.stream));
        return lane_sync(L);
    });
}

int ygpu_sample_extract_index(ygpu_ctx* ctx, const void* acc, int32_t index, int32_t* extract_out) {
    if (!acc || !extract_out) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    if (!ctx) return fail(YGPU_ERR_INVALID_ARG, "null context");
    if (index < 0 || index >= ctx->p.N) return fail(YGPU_ERR_INVALID_ARG, "extraction index must be in [0, N)");
    return with_lane(ctx, -1, [&](PlainLane& L) -> int {
        const ygpu_params& p = ctx->p;
        const size_t tb = ctx->torus_bytes();
        YG_HIP(lane_upload(ctx, L, L.d_acc, acc, tb * (p.k + 1) * p.N));
        KsArgs a = ks_args(ctx, 1, L.d_acc, nullptr, nullptr, nullptr, nullptr);
        a.extract_index = index;
        YG_HIP(timed(ctx, "k_extract", L.stream, [&] { return launch_extract(ctx->tbits, a, L.d_ext, L.stream); }));
        YG_HIP(hipMemcpyAsync(extract_out, L.d_ext, sizeof(int32_t) * (p.k * p.N + 1), hipMemcpyDeviceToHost, L.stream));
        return lane_sync(L);
    });
}

int ygpu_sample_extract(ygpu_ctx* ctx, const void* acc, int32_t* extract_out) {
    if (!ctx) return fail(YGPU_ERR_INVALID_ARG, "null context");
    int32_t index = 0;
    int rc = ygpu_get_extract_index(ctx, &index);
    if (rc) return rc;
    return ygpu_sample_extract_index(ctx, acc, index, extract_out);
}

int ygpu_extract_keyswitch(ygpu_ctx* ctx, int32_t key_id, const void* acc, int32_t* extract_out, int32_t* lwe_out) {
    if (!acc || !lwe_out) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    return with_lane(ctx, key_id, [&](PlainLane& L) -> int {
        const ygpu_params& p = ctx->p;
        const size_t tb = ctx->torus_bytes();
        YG_HIP(lane_upload(ctx, L, L.d_acc, acc, tb * (p.k + 1) * p.N));
        int rc = lane_extract_keyswitch(ctx, L, key_id, extract_out != nullptr, L.extract_index);
        if (rc) return rc;
        YG_HIP(hipMemcpyAsync(lwe_out, L.d_out, sizeof(int32_t) * (p.n + 1), hipMemcpyDeviceToHost, L.stream));
        if (extract_out) YG_HIP(hipMemcpyAsync(extract_out, L.d_ext, sizeof(int32_t) * (p.k * p.N + 1), hipMemcpyDeviceToHost, L.stream));
        return lane_sync(L);
    });
}

int ygpu_keyswitch(ygpu_ctx* ctx, int32_t key_id, const int32_t* lwe_in, int32_t* lwe_out) {
    if (!lwe_in || !lwe_out) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    if (!ctx) return fail(YGPU_ERR_INVALID_ARG, "null context");
    // Embed the LWE (dimension kN) as the accumulator whose extraction at
    // index 0 it is: acc_c[0] = a_{cN}, acc_c[N - j] = -a_{cN+j} (j > 0),
    // body[0] = b, scaled by 2^(w-32) (exact, and undone exactly by the
    // extraction's modswitch); the keyswitch then extracts at index 0.
    const ygpu_params& p = ctx->p;
    const int N = p.N, w = p.torus_bits;
    std::vector<int64_t> acc(static_cast<size_t>(p.k + 1) * N, 0);
    for (int c = 0; c < p.k; c++) {
        acc[static_cast<size_t>(c) * N] = static_cast<int64_t>(lwe_in[c * N]) << (w - 32);
        for (int j = 1; j < N; j++) acc[static_cast<size_t>(c) * N + N - j] = sext(-(static_cast<int64_t>(lwe_in[c * N + j]) << (w - 32)), w);
    }
    acc[static_cast<size_t>(p.k) * N] = static_cast<int64_t>(lwe_in[p.k * N]) << (w - 32);
    std::vector<int32_t> a32;
    const void* src = acc.data();
    if (w == 32) {
        a32.resize(acc.size());
        for (size_t i = 0; i < acc.size(); i++) a32[i] = static_cast<int32_t>(acc[i]);
        src = a32.data();
    }
    return with_lane(ctx, key_id, [&](PlainLane& L) -> int {
        YG_HIP(lane_upload(ctx, L, L.d_acc, src, ctx->torus_bytes() * (p.k + 1) * N));
        int rc = lane_extract_keyswitch(ctx, L, key_id, false, 0);
        if (rc) return rc;
        YG_HIP(hipMemcpyAsync(lwe_out, L.d_out, sizeof(int32_t) * (p.n + 1), hipMemcpyDeviceToHost, L.stream));
        return lane_sync(L);
    });
}

int ygpu_malloc(ygpu_ctx* ctx, size_t bytes, void** d_ptr) {
    if (!d_ptr) return fail(YGPU_ERR_INVALID_ARG, "null pointer");
    return with_ctx(ctx, [&]() -> int {
        hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
        if (e != hipSuccess) return fail(YGPU_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
        return YGPU_OK;
    });
}

int ygpu_free(ygpu_ctx* ctx, void* d_ptr) {
    return with_ctx(ctx, [&]() -> int {
        YG_HIP(hipStreamSynchronize(ctx->stream));
        if (d_ptr) YG_HIP(hipFree(d_ptr));
        return YGPU_OK;
    });
}

int ygpu_memcpy_htod(ygpu_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
    return with_ctx(ctx, [&]() -> int {
        YG_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        YG_HIP(hipStreamSynchronize(ctx->stream));
        return YGPU_OK;
    });
}

int ygpu_memcpy_dtoh(ygpu_ctx* ctx, void* dst, const void* d_src, size_t bytes) {
    return with_ctx(ctx, [&]() -> int {
        YG_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        return sync_and_check(ctx);
    });
}

int ygpu_sync(ygpu_ctx* ctx) {
    return with_ctx(ctx, [&]() -> int {
        YG_HIP(hipStreamSynchronize(ctx->stream));
        const int rc = ctx->timing ? resolve_timing(ctx) : YGPU_OK;
        const int st = check_status(ctx);
        return st ? st : rc;
    });
}

int ygpu_set_extract_index(ygpu_ctx* ctx, int32_t index) {
    return with_ctx(ctx, [&]() -> int {
        if (index < 0 || index >= ctx->p.N) return fail(YGPU_ERR_INVALID_ARG, "extraction index must be in [0, N)");
        ctx->extract_index = index;
        return YGPU_OK;
    });
}

int ygpu_get_extract_index(ygpu_ctx* ctx, int32_t* index) {
    if (!index) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    return with_ctx(ctx, [&]() -> int {
        *index = ctx->extract_index;
        return YGPU_OK;
    });
}

int ygpu_set_barrier_polls(ygpu_ctx* ctx, int32_t polls) {
    if (polls < 0) return fail(YGPU_ERR_INVALID_ARG, "polls must be >= 0");
    return with_ctx(ctx, [&]() -> int {
        ctx->barrier_polls = polls > 0 ? polls : 0;
        return YGPU_OK;
    });
}

int ygpu_modmul_peak(ygpu_ctx* ctx, int32_t kind, double* modmul_per_sec) {
    if (!modmul_per_sec || (kind != 0 && kind != 1)) return fail(YGPU_ERR_INVALID_ARG, "bad argument");
    return with_ctx(ctx, [&]() -> int {
        hipDeviceProp_t prop;
        YG_HIP(hipGetDeviceProperties(&prop, ctx->device));
        const int blocks = prop.multiProcessorCount * 16;
        const int iters = kind == 0 ? 4096 : 1024;
        void* sink = nullptr;
        YG_HIP(hipMalloc(&sink, sizeof(double) * blocks));
        const PrimeF64 pr = ctx->prime[0];
        const uint64_t q = ctx->primes[0];
        hipEvent_t a, b;
        YG_HIP(hipEventCreate(&a));
        YG_HIP(hipEventCreate(&b));
        YG_HIP(launch_modmul_peak(kind, pr, q, blocks, 16, sink, ctx->stream));  // warm-up
        YG_HIP(hipEventRecord(a, ctx->stream));
        YG_HIP(launch_modmul_peak(kind, pr, q, blocks, iters, sink, ctx->stream));
        YG_HIP(hipEventRecord(b, ctx->stream));
        YG_HIP(hipEventSynchronize(b));
        float ms = 0;
        YG_HIP(hipEventElapsedTime(&ms, a, b));
        hipEventDestroy(a);
        hipEventDestroy(b);
        hipFree(sink);
        const double ops = static_cast<double>(blocks) * 256.0 * MODMUL_CHAINS * iters;
        *modmul_per_sec = ops / (ms * 1e-3);
        return YGPU_OK;
    });
}

int ygpu_copy_device(ygpu_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
    if (!d_dst || !d_src || bytes % 16) return fail(YGPU_ERR_INVALID_ARG, "copy needs device buffers and a multiple of 16 bytes");
    return with_ctx(ctx, [&]() -> int {
        YG_HIP(timed(ctx, "k_copy_u32x4", ctx->stream, [&] { return launch_copy(d_src, d_dst, bytes, ctx->stream); }));
        return YGPU_OK;
    });
}

int ygpu_get_tile(ygpu_ctx* ctx, int32_t* cpw, int32_t* twl) {
    if (!cpw || !twl) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    return with_ctx(ctx, [&]() -> int {
        *cpw = ctx->tile.cpw;
        *twl = ctx->tile.twl;
        return YGPU_OK;
    });
}

int ygpu_set_tile_ex(ygpu_ctx* ctx, int32_t cpw, int32_t twl, int32_t tab) {
    return with_ctx(ctx, [&]() -> int {
        BrTile c[MAX_TILES];
        const int n = blind_rotate_tiles(ctx->tbits, ctx->logn, ctx->p.level, ctx->p.radix_bits, c, MAX_TILES);
        // tab < 0: the current tile's tab where that combination exists, else the first
        for (int pass = 0; pass < 2; pass++) {
            const int want = tab >= 0 ? tab : (pass == 0 ? ctx->tile.tab : -1);
            for (int i = 0; i < n; i++)
                if (c[i].cpw == cpw && c[i].twl == twl && (want < 0 || c[i].tab == want)) {
                    ctx->tile = c[i];
                    ctx->tile_fixed = true;
                    return YGPU_OK;
                }
            if (tab >= 0) break;
        }
        return fail(YGPU_ERR_UNSUPPORTED, "tile not instantiated for this parameter set");
    });
}

int ygpu_set_tile(ygpu_ctx* ctx, int32_t cpw, int32_t twl) { return ygpu_set_tile_ex(ctx, cpw, twl, -1); }

int ygpu_get_tile_ex(ygpu_ctx* ctx, int32_t* cpw, int32_t* twl, int32_t* tab) {
    if (!cpw || !twl || !tab) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    return with_ctx(ctx, [&]() -> int {
        *cpw = ctx->tile.cpw;
        *twl = ctx->tile.twl;
        *tab = ctx->tile.tab;
        return YGPU_OK;
    });
}

int ygpu_autotune_results(ygpu_ctx* ctx, int32_t max, int32_t* cpw, int32_t* twl, int32_t* tab, double* ms, int32_t* count) {
    if (!count || max < 0 || (max > 0 && (!cpw || !twl || !tab || !ms))) return fail(YGPU_ERR_INVALID_ARG, "bad argument");
    return with_ctx(ctx, [&]() -> int {
        const int n = static_cast<int>(ctx->tune.size());
        for (int i = 0; i < n && i < max; i++) {
            cpw[i] = ctx->tune[i].first.cpw;
            twl[i] = ctx->tune[i].first.twl;
            tab[i] = ctx->tune[i].first.tab;
            ms[i] = ctx->tune[i].second;
        }
        *count = n;
        return YGPU_OK;
    });
}

int ygpu_autotune(ygpu_ctx* ctx, int32_t key_id, int32_t count, int32_t* cpw, int32_t* twl, double* best_ms) {
    if (count < 1) return fail(YGPU_ERR_INVALID_ARG, "count must be positive");
    return with_ctx(ctx, [&]() -> int {
        int rc = engine_key_live(ctx, key_id);
        if (rc) return rc;
        const ygpu_params& p = ctx->p;
        BrTile cand[MAX_TILES];
        const int nc = blind_rotate_tiles(ctx->tbits, ctx->logn, p.level, p.radix_bits, cand, MAX_TILES);
        ctx->tune.clear();
        // synthetic batch: pseudo-random ciphertexts, a zero test polynomial
        std::vector<int32_t> lwe(static_cast<size_t>(count) * (p.n + 1));
        uint64_t st = 0x243F6A8885A308D3ULL;
        for (auto& x : lwe) x = static_cast<int32_t>(splitmix64(st));
        // released on every return path
        struct AutotuneBuffers {
            void *lwe = nullptr, *v = nullptr, *acc = nullptr;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            ~AutotuneBuffers() {
                for (void* b : {lwe, v, acc}) if (b) hipFree(b);
                if (e0) hipEventDestroy(e0);
                if (e1) hipEventDestroy(e1);
            }
        } sc;
        YG_HIP(hipMalloc(&sc.lwe, lwe.size() * sizeof(int32_t)));
        YG_HIP(hipMalloc(&sc.v, ctx->torus_bytes() * p.N));
        YG_HIP(hipMalloc(&sc.acc, ctx->torus_bytes() * (p.k + 1) * p.N * static_cast<size_t>(count)));
        YG_HIP(hipMemcpyAsync(sc.lwe, lwe.data(), lwe.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
        YG_HIP(hipMemsetAsync(sc.v, 0, ctx->torus_bytes() * p.N, ctx->stream));
        YG_HIP(hipStreamSynchronize(ctx->stream));
        YG_HIP(hipEventCreate(&sc.e0));
        YG_HIP(hipEventCreate(&sc.e1));
        void *d_lwe = sc.lwe, *d_v = sc.v, *d_acc = sc.acc;
        hipEvent_t e0 = sc.e0, e1 = sc.e1;
        double best = 1e30;
        BrTile bestT = ctx->tile;
        for (int c = 0; c < nc; c++) {
            BrArgs a = br_args(ctx, count, static_cast<int32_t*>(d_lwe), d_v, 0, d_acc, ctx->d_bsk_table + key_id, nullptr);
            YG_HIP(launch_blind_rotate(ctx->tbits, ctx->logn, cand[c], p.radix_bits, a, ctx->stream));  // warm-up
            double ms = 1e30;
            for (int rep = 0; rep < 2; rep++) {
                YG_HIP(hipEventRecord(e0, ctx->stream));
                YG_HIP(launch_blind_rotate(ctx->tbits, ctx->logn, cand[c], p.radix_bits, a, ctx->stream));
                YG_HIP(hipEventRecord(e1, ctx->stream));
                YG_HIP(hipEventSynchronize(e1));
                float t = 0;
                YG_HIP(hipEventElapsedTime(&t, e0, e1));
                ms = std::min(ms, static_cast<double>(t));
            }
            ctx->tune.emplace_back(cand[c], ms);
            if (ms < best) { best = ms; bestT = cand[c]; }
        }
        ctx->tile = bestT;
        if (cpw) *cpw = bestT.cpw;
        if (twl) *twl = bestT.twl;
        if (best_ms) *best_ms = best;
        return YGPU_OK;
    });
}

namespace {
int capture_graph(ygpu_ctx* ctx, GraphRec& g, const std::function<int()>& body, int32_t* graph_id) {
    YG_HIP(hipStreamSynchronize(ctx->stream));
    YG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    int rc = body();
    ctx->capturing = false;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc) { if (graph) hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return fail(YGPU_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    g.graph = graph;
    YG_HIP(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
    g.live = true;
    int id = -1;
    for (size_t i = 0; i < ctx->graphs.size(); i++) if (!ctx->graphs[i].live) { id = static_cast<int>(i); break; }
    if (id < 0) { id = static_cast<int>(ctx->graphs.size()); ctx->graphs.push_back(g); }
    else ctx->graphs[id] = g;
    *graph_id = id;
    return YGPU_OK;
}

}  // namespace

int ygpu_graph_create(ygpu_ctx* ctx, const ygpu_job* jobs, int32_t n_jobs, int64_t total, const int32_t* d_lwe_in,
                      const void* d_v, int32_t v_per_ct, int32_t* d_lwe_out, void* d_acc_out, int32_t* graph_id) {
    if (!jobs || n_jobs < 1 || total < 1 || total > (1LL << 30) || !d_lwe_in || !d_v || !d_lwe_out || !graph_id)
        return fail(YGPU_ERR_INVALID_ARG, "bad graph arguments");
    return with_ctx(ctx, [&]() -> int {
        std::vector<int32_t> slots;
        int rc = job_slots(ctx, jobs, n_jobs, total, slots);
        if (rc) return rc;
        GraphRec g;
        const ygpu_params& p = ctx->p;
        YG_HIP(hipMalloc(reinterpret_cast<void**>(&g.d_slots), sizeof(int32_t) * slots.size()));
        YG_HIP(hipMemcpyAsync(g.d_slots, slots.data(), sizeof(int32_t) * slots.size(), hipMemcpyHostToDevice, ctx->stream));
        YG_HIP(hipStreamSynchronize(ctx->stream));
        YG_HIP(hipMalloc(reinterpret_cast<void**>(&g.d_partial), ks_workspace_bytes(ctx, static_cast<int>(total))));
        void* acc = d_acc_out;
        if (!acc) {
            YG_HIP(hipMalloc(&g.d_acc, ctx->torus_bytes() * (p.k + 1) * p.N * static_cast<size_t>(total)));
            acc = g.d_acc;
        }
        for (int j = 0; j < n_jobs; j++)
            if (std::find(g.keys.begin(), g.keys.end(), jobs[j].key_id) == g.keys.end()) g.keys.push_back(jobs[j].key_id);
        rc = capture_graph(ctx, g, [&]() {
            return enqueue_jobs(ctx, static_cast<int>(total), d_lwe_in, d_v, v_per_ct, d_lwe_out, acc,
                                g.d_slots, g.d_partial);
        }, graph_id);
        if (rc) free_graph(g);
        return rc;
    });
}

int ygpu_graph_launch(ygpu_ctx* ctx, int32_t graph_id) {
    return with_ctx(ctx, [&]() -> int {
        if (graph_id < 0 || graph_id >= static_cast<int32_t>(ctx->graphs.size()) || !ctx->graphs[graph_id].live)
            return fail(YGPU_ERR_NOT_FOUND, "unknown graph id");
        YG_HIP(timed(ctx, "graph_pbs", ctx->stream, [&] { return hipGraphLaunch(ctx->graphs[graph_id].exec, ctx->stream); }));
        return YGPU_OK;
    });
}

int ygpu_graph_destroy(ygpu_ctx* ctx, int32_t graph_id) {
    return with_ctx(ctx, [&]() -> int {
        if (graph_id < 0 || graph_id >= static_cast<int32_t>(ctx->graphs.size()) || !ctx->graphs[graph_id].live)
            return fail(YGPU_ERR_NOT_FOUND, "unknown graph id");
        YG_HIP(hipStreamSynchronize(ctx->stream));
        free_graph(ctx->graphs[graph_id]);
        return YGPU_OK;
    });
}

int ygpu_timing_enable(ygpu_ctx* ctx, int32_t on) {
    return with_ctx(ctx, [&]() -> int {
        ctx->timing = on != 0;
        return YGPU_OK;
    });
}

int ygpu_timing_get(ygpu_ctx* ctx, const char* kernel, int64_t* launches, double* total_ms) {
    if (!kernel || !launches || !total_ms) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    return with_ctx(ctx, [&]() -> int {
        int rc = resolve_timing(ctx);
        if (rc) return rc;
        std::lock_guard<std::mutex> tl(ctx->timing_mu);
        auto it = ctx->stats.find(kernel);
        *launches = it == ctx->stats.end() ? 0 : it->second.first;
        *total_ms = it == ctx->stats.end() ? 0.0 : it->second.second;
        return YGPU_OK;
    });
}

int ygpu_timing_reset(ygpu_ctx* ctx) {
    return with_ctx(ctx, [&]() -> int {
        int rc = resolve_timing(ctx);
        std::lock_guard<std::mutex> tl(ctx->timing_mu);
        ctx->stats.clear();
        return rc;
    });
}

}  // extern "C"
