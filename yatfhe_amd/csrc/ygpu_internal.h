// ygpu_internal.h -- runtime state and helpers shared by the C ABI
// implementation (ygpu_runtime.hip) and the serving engine (ygpu_server.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstring>
#include <cstdio>
#include <string>
#include <vector>
#include <map>
#include <mutex>
#include <thread>
#include <atomic>
#include <memory>
#include <functional>
#include <condition_variable>
#include <cstdlib>
#include <algorithm>
#include "yatfhe_gpu.h"
#include "ygpu_math.h"
#include "ygpu_kernels.h"

using namespace ygpu;


namespace ygpu_rt {

inline thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define YG_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) return fail(YGPU_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct KeyRec {
    bool live = false;
    double* d_bsk = nullptr;   // engine-form BSK (null for a lazy-pipe key)
    int32_t* d_ksk = nullptr;
    void* d_lazy = nullptr;    // a lazy-pipe key's device block (lazy_layout)
    bool lazy = false;
    bool lazy_alt = false;     // the Alt form: torus rows instead of digits
    int lazy_digits = 0;       // its digit levels (param.l)
    size_t bytes = 0;
    int users = 0;  // synchronous calls running on a plain lane with this key
};

// One calling thread's resources for the synchronous host-buffer entry
// points (ygpu_pbs and its parts): a stream, one ciphertext's device buffers
// and a failure-report word of its own.  Threads that call at the same time
// each take a lane, so their bootstraps run side by side on the device
// instead of queueing behind one lock.
struct PlainLane {
    bool busy = false;
    // the context's settings as the current call found them
    BrTile tile{};
    int extract_index = 0;
    int barrier_polls = TEAM_BARRIER_POLLS;
    hipStream_t stream = nullptr;
    int32_t* d_lwe = nullptr;
    void* d_v = nullptr;
    int32_t* d_out = nullptr;
    int32_t* d_ext = nullptr;
    void* d_acc = nullptr;
    void* d_accin = nullptr;
    int32_t* d_partial = nullptr;   // keyswitch workspace of one ciphertext
    uint32_t* h_status = nullptr;   // pinned; the lane's barrier-failure report
    uint32_t* hd_status = nullptr;
};

inline void free_lane(PlainLane& L) {
    if (L.stream) hipStreamDestroy(L.stream);
    for (void* p : {static_cast<void*>(L.d_lwe), L.d_v, static_cast<void*>(L.d_out), static_cast<void*>(L.d_ext), L.d_acc,
                    L.d_accin, static_cast<void*>(L.d_partial)})
        if (p) hipFree(p);
    if (L.h_status) hipHostFree(L.h_status);
    L = PlainLane{};
}

struct Timing {
    std::string name;
    hipEvent_t a, b;
};

// A captured blind-rotate -> extract -> keyswitch pipeline with its own
// scratch; `keys` are the key slots its kernels read (they cannot be released
// while the graph exists).
struct GraphRec {
    bool live = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* d_acc = nullptr;
    int32_t* d_partial = nullptr;
    int32_t* d_slots = nullptr;
    std::vector<int32_t> keys;
};

inline void free_graph(GraphRec& g) {
    if (g.exec) hipGraphExecDestroy(g.exec);
    if (g.graph) hipGraphDestroy(g.graph);
    for (void* p : {g.d_acc, static_cast<void*>(g.d_partial), static_cast<void*>(g.d_slots)}) if (p) hipFree(p);
    g = GraphRec{};
}


}  // namespace ygpu_rt

using namespace ygpu_rt;

struct ygpu_ctx {
    ygpu_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = true;  // false: borrowed from the owner (a server's copy stream)
    int logn = 0, np = 1, tbits = 32;
    BrTile tile{};          // batched-form tile (autotuned or default)
    std::vector<std::pair<BrTile, double>> tune;  // the last autotune's candidates and their ms
    bool tile_fixed = false; // set by ygpu_set_tile: used for every batch size
    int cu_count = 0;       // compute units of the device
    bool capturing = false;
    std::vector<GraphRec> graphs;
    uint64_t primes[2] = {0, 0};
    std::vector<HostNtt> ntt;
    double* d_tw = nullptr;
    PrimeF64 prime[2]{};
    CrtConst crt{};
    std::vector<KeyRec> keys;
    // device key tables (BSK / KSK base per key slot), allocated once at init
    // with KEY_SLOTS entries: captured graphs hold their addresses
    const double** d_bsk_table = nullptr;
    const int32_t** d_ksk_table = nullptr;
    // failure report of the team barriers (pinned host word and its device
    // address): set by a kernel whose bounded barrier wait ran out, read and
    // cleared after each synchronisation
    uint32_t* h_status = nullptr;
    uint32_t* hd_status = nullptr;
    int barrier_polls = TEAM_BARRIER_POLLS;
    // this device's copy of the library's static scratch array (the device
    // address hipGetSymbolAddress returned at init; null if that failed)
    void* d_static_scratch = nullptr;
    int extract_index = 0;  // extraction coefficient of the bootstraps (param.driftPhase)
    void* d_acc = nullptr;
    size_t acc_cap = 0;
    int32_t* d_partial = nullptr;
    size_t partial_cap = 0;
    int32_t* d_slots = nullptr;
    size_t slots_cap = 0;
    int32_t* d_lwe1 = nullptr;
    void* d_v1 = nullptr;
    int32_t* d_out1 = nullptr;
    int32_t* d_ext1 = nullptr;
    void* d_acc1 = nullptr;
    void* d_accin1 = nullptr;  // caller's initial accumulator (ygpu_blind_rotate_scaled)
    void* d_stage = nullptr;   // packed key staging for uploads
    size_t stage_cap = 0;
    // lazy-pipe runs: derived-row workspace, rotation amounts, accumulators,
    // per-ciphertext row tables; and the window (steps per derive launch, 0: auto)
    void* d_lazy_ws = nullptr;
    size_t lazy_ws_cap = 0;
    void* d_lazy_av = nullptr;
    size_t lazy_av_cap = 0;
    void* d_lazy_acc = nullptr;
    size_t lazy_acc_cap = 0;
    void* d_lazy_tab = nullptr;
    size_t lazy_tab_cap = 0;
    void* d_lazy_first = nullptr;  // the Alt form's derived first components
    size_t lazy_first_cap = 0;
    int lazy_window = 0;
    std::mutex mu;
    std::condition_variable cv;   // a plain lane or a key's last user finished
    std::vector<std::unique_ptr<PlainLane>> lanes;
    std::atomic<bool> timing{false};
    std::mutex timing_mu;         // pending, event_pool, stats
    std::vector<Timing> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<int64_t, double>> stats;

    size_t torus_bytes() const { return tbits == 32 ? 4 : 8; }
};

namespace ygpu_rt {

inline int check_params(const ygpu_params& p) {
    if (p.torus_bits != 32 && p.torus_bits != 56) return fail(YGPU_ERR_UNSUPPORTED, "torus_bits must be 32 or 56");
    if (p.k != 1) return fail(YGPU_ERR_UNSUPPORTED, "only GLWE dimension k = 1 is supported");
    if (p.N != 1024 && p.N != 2048) return fail(YGPU_ERR_UNSUPPORTED, "ring degree N must be 1024 or 2048");
    if (p.n < 1 || p.n > KS_MAX_N) return fail(YGPU_ERR_UNSUPPORTED, "LWE dimension n must be in [1, 1023]");
    // digits lie in [-B/2, B/2] and travel in 16-bit fields: B/2 <= 2^14
    if (p.radix_bits < 1 || p.radix_bits > 15) return fail(YGPU_ERR_UNSUPPORTED, "radix_bits must be in [1, 15]");
    // two 32-bit words of digits per position: 8-bit fields (B <= 2^7, up to 8
    // levels) or 16-bit fields (up to 4 levels)
    if (p.level < 1 || p.level > (p.radix_bits <= 7 ? 8 : 4))
        return fail(YGPU_ERR_UNSUPPORTED, "gadget level must be in [1, 8] for radix_bits <= 7 and in [1, 4] above");
    // decomposeRow's own checks (gadget_decomposition.h:194-206)
    if (p.level * p.radix_bits > p.torus_bits)
        return fail(YGPU_ERR_INVALID_ARG, "decomposeRow: " + std::to_string(p.level) + " digits of " + std::to_string(p.radix_bits) +
                                              " bits do not fit the " + std::to_string(p.torus_bits) + "-bit torus");
    if (p.lwe_q_bits != 32) return fail(YGPU_ERR_UNSUPPORTED, "qLwe must be 2^32");
    if (p.ks_width_bits < 1 || p.ks_width_bits > 32) return fail(YGPU_ERR_UNSUPPORTED, "ks_width_bits must be in [1, 32]");
    // the keyswitch GEMM takes its digits as signed bytes: a balanced digit of
    // base 2^b lies in [-2^(b-1), 2^(b-1)], and +128 (b = 8) has no int8 form
    if (p.ks_radix_bits < 1 || p.ks_radix_bits > 7)
        return fail(YGPU_ERR_UNSUPPORTED, "ks_radix_bits must be in [1, 7]: the keyswitch digits travel as signed bytes, "
                                          "and base 2^8 yields the digit +128");
    if (p.ks_level < 1 || p.ks_level > KS_MAX_LEVEL) return fail(YGPU_ERR_UNSUPPORTED, "ks_level must be in [1, 16]");
    if (p.ks_level * p.ks_radix_bits > p.ks_width_bits) return fail(YGPU_ERR_INVALID_ARG, "keyswitch digits exceed ks_width_bits");
    if (!is_prime_u64(p.q_ntt)) return fail(YGPU_ERR_INVALID_ARG, "q_ntt is not prime");
    if ((p.q_ntt - 1) % (2 * static_cast<uint64_t>(p.N)) != 0) return fail(YGPU_ERR_INVALID_ARG, "q_ntt is not 1 mod 2N");
    if (p.q_ntt >= (1ULL << 63)) return fail(YGPU_ERR_UNSUPPORTED, "q_ntt must be below 2^63");
    if (p.torus_bits == 56 && p.q_ntt <= (1ULL << 56)) return fail(YGPU_ERR_INVALID_ARG, "TORUS56 needs q_ntt > 2^56");
    return YGPU_OK;
}

template <class F>
inline int with_ctx(ygpu_ctx* ctx, F&& f) {
    if (!ctx) return fail(YGPU_ERR_INVALID_ARG, "null context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail(YGPU_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    try {
        return f();
    } catch (const std::exception& ex) {
        return fail(YGPU_ERR_INVALID_ARG, ex.what());
    }
}

// timing_mu held.
inline hipEvent_t take_event(ygpu_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Launch `f` bracketed by events on `stream` (the stream `f` launches on)
// when timing is on.
template <class F>
hipError_t timed(ygpu_ctx* ctx, const char* name, hipStream_t stream, F&& f) {
    if (!ctx->timing || ctx->capturing) return f();
    Timing t;
    {
        std::lock_guard<std::mutex> lk(ctx->timing_mu);
        t = Timing{name, take_event(ctx), take_event(ctx)};
    }
    hipEventRecord(t.a, stream);
    hipError_t e = f();
    hipEventRecord(t.b, stream);
    std::lock_guard<std::mutex> lk(ctx->timing_mu);
    ctx->pending.push_back(t);
    return e;
}

inline int resolve_timing(ygpu_ctx* ctx) {
    std::lock_guard<std::mutex> lk(ctx->timing_mu);
    for (auto& t : ctx->pending) {
        YG_HIP(hipEventSynchronize(t.b));
        float ms = 0;
        YG_HIP(hipEventElapsedTime(&ms, t.a, t.b));
        auto& s = ctx->stats[t.name];
        s.first += 1;
        s.second += ms;
        ctx->event_pool.push_back(t.a);
        ctx->event_pool.push_back(t.b);
    }
    ctx->pending.clear();
    return YGPU_OK;
}

inline int ensure(void** buf, size_t* cap, size_t bytes) {
    if (*cap >= bytes) return YGPU_OK;
    if (*buf) hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc(buf, bytes);
    if (e != hipSuccess) return fail(YGPU_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    *cap = bytes;
    return YGPU_OK;
}

// Key slots per context: the device key tables never move, so graphs that
// captured their addresses stay valid whatever keys are uploaded later.
constexpr int KEY_SLOTS = 4096;

// Writes key slot `id`'s table entries (the key's device bases, or null for a
// released slot) on the context stream.  A kernel in flight reads only the
// slots of its own keys, which are live.
inline int set_table_slot(ygpu_ctx* ctx, int id) {
    if (id < 0 || id >= KEY_SLOTS) return fail(YGPU_ERR_INVALID_ARG, "key slot out of range");
    const KeyRec& k = ctx->keys[id];
    const double* b = k.live ? k.d_bsk : nullptr;
    const int32_t* q = k.live ? k.d_ksk : nullptr;
    YG_HIP(hipMemcpyAsync(ctx->d_bsk_table + id, &b, sizeof(b), hipMemcpyHostToDevice, ctx->stream));
    YG_HIP(hipMemcpyAsync(ctx->d_ksk_table + id, &q, sizeof(q), hipMemcpyHostToDevice, ctx->stream));
    YG_HIP(hipStreamSynchronize(ctx->stream));
    return YGPU_OK;
}

// Reads and clears the context's barrier-failure report (after the stream it
// covers has been synchronised).
inline int check_status(ygpu_ctx* ctx) {
    if (!ctx->h_status) return YGPU_OK;
    volatile uint32_t* st = ctx->h_status;
    if (*st == 0u) return YGPU_OK;
    *st = 0u;
    return fail(YGPU_ERR_DEVICE, "a blind-rotation workgroup's team barrier timed out: the results of the calls since the "
                                 "last synchronisation are invalid (their accumulators read the most negative torus value)");
}

// Synchronises the context stream and reports a device-side failure.
inline int sync_and_check(ygpu_ctx* ctx) {
    YG_HIP(hipStreamSynchronize(ctx->stream));
    return check_status(ctx);
}

inline int key_live(const ygpu_ctx* ctx, int32_t key_id) {
    if (key_id < 0 || key_id >= static_cast<int32_t>(ctx->keys.size()) || !ctx->keys[key_id].live)
        return fail(YGPU_ERR_NOT_FOUND, "unknown key id " + std::to_string(key_id));
    return YGPU_OK;
}
// A live key in the engine's BSK form (every entry point but the lazy-pipe ones).
inline int engine_key_live(const ygpu_ctx* ctx, int32_t key_id) {
    const int rc = key_live(ctx, key_id);
    if (rc) return rc;
    if (ctx->keys[key_id].lazy)
        return fail(YGPU_ERR_UNSUPPORTED, "key " + std::to_string(key_id) + " is a lazy-pipe key: use the ygpu_*_lazy entry points");
    return YGPU_OK;
}
// A live lazy-pipe key.
inline int lazy_key_live(const ygpu_ctx* ctx, int32_t key_id) {
    const int rc = key_live(ctx, key_id);
    if (rc) return rc;
    if (!ctx->keys[key_id].lazy) return fail(YGPU_ERR_UNSUPPORTED, "key " + std::to_string(key_id) + " is not a lazy-pipe key");
    return YGPU_OK;
}

// Device position of transform output index k for lane tl, register r.
inline int last_pass_index(int logn, int tl, int r) {
    return logn == 10 ? Pass<10, 2>::idx(tl, r) : Pass<11, 2>::idx(tl, r);
}

inline BrArgs br_args(ygpu_ctx* ctx, int count, const int32_t* d_lwe, const void* d_v, int v_stride, void* d_acc,
               const double* const* table, const int32_t* slots) {
    BrArgs a{};
    a.n = ctx->p.n;
    a.level = ctx->p.level;
    a.radix_bits = ctx->p.radix_bits;
    a.count = count;
    a.prescaled = 0;
    a.acc_in = nullptr;
    a.lwe_in = d_lwe;
    a.v = d_v;
    a.v_stride = v_stride;
    a.acc_out = d_acc;
    a.bsk_table = table;
    a.ct_slot = slots;
    a.tw = ctx->d_tw;
    a.prime[0] = ctx->prime[0];
    a.prime[1] = ctx->prime[1];
    a.crt = ctx->crt;
    a.status = ctx->hd_status;
    a.barrier_polls = ctx->barrier_polls;
    return a;
}

// Bytes of the keyswitch workspace (sliced partial sums + packed digits) of
// `count` ciphertexts.
inline size_t ks_workspace_bytes(const ygpu_ctx* ctx, int count) {
    return sizeof(uint32_t) * ks_workspace_words(ctx->p.n, ctx->p.k, ctx->p.N, ctx->p.ks_level, count);
}

// `workspace` holds ks_workspace_bytes(count) (null: extraction only).
inline KsArgs ks_args(ygpu_ctx* ctx, int count, const void* d_acc, const int32_t* const* table, const int32_t* slots, int32_t* d_out,
               int32_t* workspace) {
    KsArgs a{};
    a.n = ctx->p.n;
    a.N = ctx->p.N;
    a.k = ctx->p.k;
    a.count = count;
    a.ks_level = ctx->p.ks_level;
    a.ks_radix_bits = ctx->p.ks_radix_bits;
    a.ks_width_bits = ctx->p.ks_width_bits;
    a.acc = d_acc;
    a.ksk_table = table;
    a.ct_slot = slots;
    a.lwe_out = d_out;
    a.partial = workspace;
    a.digits = workspace ? reinterpret_cast<uint32_t*>(workspace) + (static_cast<size_t>(count) * (ctx->p.n + 1) + 63) / 64 * 64
                         : nullptr;
    a.extract_index = ctx->extract_index;
    return a;
}

// Keyswitch row slices for a batch of `count`: enough workgroups to cover the device.
inline int ks_slices(const ygpu_ctx* ctx, int count) {
    const int tiles = (count + KS_TC - 1) / KS_TC * ((ctx->p.n + 1 + KS_TK - 1) / KS_TK);
    int slices = 1;
    while (tiles * slices < 256 && slices < 32) slices *= 2;
    return slices;
}

// Extract + keyswitch of `count` accumulators.  `workspace` holds
// ks_workspace_bytes(count); null selects the context's buffer.
// extract_index < 0 takes the context's.
inline int run_ks(ygpu_ctx* ctx, int count, const void* d_acc, const int32_t* const* table, const int32_t* slots, int32_t* d_out,
           int32_t* d_ext, int32_t* workspace = nullptr, hipStream_t stream = nullptr, int extract_index = -1) {
    hipStream_t st = stream ? stream : ctx->stream;
    const int slices = ks_slices(ctx, count);
    if (!workspace) {
        int rc = ensure(reinterpret_cast<void**>(&ctx->d_partial), &ctx->partial_cap, ks_workspace_bytes(ctx, count));
        if (rc) return rc;
        workspace = ctx->d_partial;
    }
    KsArgs a = ks_args(ctx, count, d_acc, table, slots, d_out, workspace);
    if (extract_index >= 0) a.extract_index = extract_index;
    YG_HIP(timed(ctx, "k_extract_ks", st, [&] { return launch_extract_ks(ctx->tbits, a, slices, st); }));
    if (d_ext) YG_HIP(timed(ctx, "k_extract", st, [&] { return launch_extract(ctx->tbits, a, d_ext, st); }));
    return YGPU_OK;
}

inline int run_br(ygpu_ctx* ctx, const BrTile& tile, const BrArgs& a, hipStream_t stream = nullptr) {
    hipStream_t st = stream ? stream : ctx->stream;
    YG_HIP(timed(ctx, "k_blind_rotate", st, [&] { return launch_blind_rotate(ctx->tbits, ctx->logn, tile, ctx->p.radix_bits, a, st); }));
    return YGPU_OK;
}

// One ciphertext per workgroup, on the latency form where it is instantiated
// (elsewhere with the current tile's twiddle and digit-table choice).
inline BrTile plain_tile(const ygpu_ctx* ctx) {
    const int lat = blind_rotate_lat_supported(ctx->tbits, ctx->logn, ctx->p.level, ctx->p.radix_bits) ? 1 : 0;
    return BrTile{1, ctx->tile.twl, lat, lat ? 0 : ctx->tile.tab};
}
// Tile of a batch of `count` ciphertexts: a batch no larger than the number of
// compute units gives every ciphertext a workgroup of its own on the plain
// tile (the latency form where instantiated), which finishes sooner than pairs
// in lockstep; larger batches take the throughput tile.  A tile set explicitly
// is used for every size.
inline BrTile batch_tile(const ygpu_ctx* ctx, int count) {
    if (!ctx->tile_fixed && count <= ctx->cu_count) return plain_tile(ctx);
    return ctx->tile;
}

// Validated job list -> per-ciphertext key slots.
inline int job_slots(const ygpu_ctx* ctx, const ygpu_job* jobs, int n_jobs, int64_t total, std::vector<int32_t>& slots) {
    slots.assign(static_cast<size_t>(total), -1);
    for (int j = 0; j < n_jobs; j++) {
        int rc = engine_key_live(ctx, jobs[j].key_id);
        if (rc) return rc;
        if (jobs[j].count < 0 || jobs[j].first < 0 || jobs[j].first + jobs[j].count > total)
            return fail(YGPU_ERR_INVALID_ARG, "job " + std::to_string(j) + " out of range");
        for (int64_t c = jobs[j].first; c < jobs[j].first + jobs[j].count; c++) {
            if (slots[c] != -1) return fail(YGPU_ERR_INVALID_ARG, "ciphertext " + std::to_string(c) + " in two jobs");
            slots[c] = jobs[j].key_id;
        }
    }
    for (int64_t c = 0; c < total; c++)
        if (slots[c] < 0) return fail(YGPU_ERR_INVALID_ARG, "ciphertext " + std::to_string(c) + " in no job");
    return YGPU_OK;
}

// The key-grouped pipeline on device buffers (eager or under capture);
// d_slots holds each ciphertext's key slot.
inline int enqueue_jobs(ygpu_ctx* ctx, int count, const int32_t* d_lwe_in, const void* d_v,
                 int32_t v_per_ct, int32_t* d_lwe_out, void* acc, const int32_t* d_slots, int32_t* partial,
                 hipStream_t stream = nullptr) {
    BrArgs a = br_args(ctx, count, d_lwe_in, d_v, v_per_ct ? ctx->p.N : 0, acc, ctx->d_bsk_table, d_slots);
    int rc = run_br(ctx, batch_tile(ctx, count), a, stream);
    if (rc) return rc;
    // one keyswitch launch for all jobs: the kernel walks each tile's runs of
    // one key through the per-ciphertext slots
    return run_ks(ctx, count, acc, ctx->d_ksk_table, d_slots, d_lwe_out, nullptr, partial, stream);
}

inline int plain_buffers(ygpu_ctx* ctx) {
    if (ctx->d_lwe1) return YGPU_OK;
    const size_t tb = ctx->torus_bytes();
    YG_HIP(hipMalloc(reinterpret_cast<void**>(&ctx->d_lwe1), sizeof(int32_t) * (ctx->p.n + 1)));
    YG_HIP(hipMalloc(&ctx->d_v1, tb * ctx->p.N));
    YG_HIP(hipMalloc(reinterpret_cast<void**>(&ctx->d_out1), sizeof(int32_t) * (ctx->p.n + 1)));
    YG_HIP(hipMalloc(reinterpret_cast<void**>(&ctx->d_ext1), sizeof(int32_t) * (ctx->p.k * ctx->p.N + 1)));
    YG_HIP(hipMalloc(&ctx->d_acc1, tb * (ctx->p.k + 1) * ctx->p.N));
    YG_HIP(hipMalloc(&ctx->d_accin1, tb * (ctx->p.k + 1) * ctx->p.N));
    return YGPU_OK;
}

inline void parallel_for(int count, int threads, const std::function<void(int)>& fn) {
    std::atomic<int> next{0};
    auto work = [&]() {
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= count) break;
            fn(i);
        }
    };
    std::vector<std::thread> ts;
    const int nt = std::max(1, std::min(threads, count));
    for (int t = 1; t < nt; t++) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
}

inline int host_threads() {
    const char* e = std::getenv("YGPU_HOST_THREADS");
    if (e) return std::max(1, std::atoi(e));
    const unsigned hc = std::thread::hardware_concurrency();
    return static_cast<int>(std::max(1u, std::min(hc, 16u)));
}

// Key material as it crosses to the device: the BSK residues bit-packed
// (bits = bit length of q - 1 per value, polynomials word-aligned because N is a
// multiple of 64), expanded into the engine's form on the device, and the KSK as
// signed-byte planes.  Two BSK forms: the full key (every RGSW row; expanded by
// k_bsk_expand) and the WWL+24 succinct key (only the body-digit rows cPrime
// plus a TRLEV of s^2 with s2_level levels, appended after them; the mask-digit
// rows are derived on the device by k_bsk_derive).  Either way the device holds
// the same engine-form BSK.
struct HostKeyForm {
    std::vector<uint64_t> packed;
    std::vector<uint32_t> planes;
    size_t bsk_bytes = 0;  // engine-form BSK bytes on the device
    int s2_level = 0;      // 0: full key; > 0: succinct key
    size_t bytes() const { return bsk_bytes + planes.size() * sizeof(uint32_t); }
    size_t transfer_bytes() const { return packed.size() * sizeof(uint64_t) + planes.size() * sizeof(uint32_t); }
};

inline int key_bits(uint64_t q) { return 64 - __builtin_clzll(q - 1); }

inline size_t bsk_polys(const ygpu_params& p) { return static_cast<size_t>(p.n) * p.level * (p.k + 1) * (p.k + 1); }
// Body-digit rows of a succinct key, and the polynomials of its s^2 TRLEV.
inline size_t half_polys(const ygpu_params& p) { return static_cast<size_t>(p.n) * p.level * (p.k + 1); }
inline size_t s2_polys(const ygpu_params& p, int s2_level) { return static_cast<size_t>(s2_level) * (p.k + 1); }
inline size_t poly_words(const ygpu_params& p) { return static_cast<size_t>(p.N) * key_bits(p.q_ntt) / 64; }

// Words of the packed BSK (+1 so that a two-word read of the last value stays in bounds).
inline size_t packed_words(const ygpu_params& p) { return bsk_polys(p) * poly_words(p) + 1; }
inline size_t packed_words_succinct(const ygpu_params& p, int s2_level) {
    return (half_polys(p) + s2_polys(p, s2_level)) * poly_words(p) + 1;
}
constexpr int S2_MAX_LEVEL = 8;
// Device staging for one key load: the packed words, then (succinct keys) the
// engine-form s^2 TRLEV, 256-byte aligned.
inline size_t s2_stage_offset(const ygpu_params& p, int s2_level) {
    return (packed_words_succinct(p, s2_level) * sizeof(uint64_t) + 255) / 256 * 256;
}
inline size_t stage_bytes(const ygpu_params& p, int np, int s2_level) {
    if (s2_level <= 0) return packed_words(p) * sizeof(uint64_t);
    return s2_stage_offset(p, s2_level) + s2_polys(p, s2_level) * np * p.N * sizeof(double);
}
// Enough staging for a load of either form.
inline size_t stage_bytes_any(const ygpu_params& p, int np) {
    size_t b = stage_bytes(p, np, 0);
    for (int l = 1; l <= S2_MAX_LEVEL; l++) b = std::max(b, stage_bytes(p, np, l));
    return b;
}

inline int check_s2_level(const ygpu_params& p, int s2_level) {
    if (s2_level < 1 || s2_level > S2_MAX_LEVEL) return fail(YGPU_ERR_UNSUPPORTED, "s2_level must be in [1, 8]");
    // decomposeRow's own check (gadget_decomposition.h:194-206) on the s^2 digits
    if (s2_level * p.radix_bits > p.torus_bits)
        return fail(YGPU_ERR_INVALID_ARG, "decomposeRow: " + std::to_string(s2_level) + " digits of " + std::to_string(p.radix_bits) +
                                              " bits do not fit the " + std::to_string(p.torus_bits) + "-bit torus");
    return YGPU_OK;
}

// Packs `polys` polynomials of N residues (each < q) at `bits` per value.
// Returns the index of the first residue not below q, or -1.
inline int64_t pack_residues(const uint64_t* src, size_t polys, const ygpu_params& p, uint64_t* dst) {
    const int N = p.N;
    const uint64_t q = p.q_ntt;
    const int bits = key_bits(q);
    const size_t polyWords = poly_words(p);  // exact: N is a multiple of 64
    std::atomic<int64_t> bad{-1};
    parallel_for(static_cast<int>(polys), host_threads(), [&](int pi) {
        const uint64_t* in = src + static_cast<size_t>(pi) * N;
        uint64_t* out = dst + static_cast<size_t>(pi) * polyWords;
        size_t bit = 0;
        for (int j = 0; j < N; j++, bit += bits) {
            const uint64_t v = in[j];
            if (v >= q) {
                int64_t expect = -1;
                bad.compare_exchange_strong(expect, static_cast<int64_t>(pi) * N + j);
                return;
            }
            const size_t w = bit >> 6;
            const int off = static_cast<int>(bit & 63);
            out[w] |= v << off;
            if (off + bits > 64) out[w + 1] |= v >> (64 - off);
        }
    });
    return bad.load();
}

// KSK as signed-byte planes: rows grouped by four, word = the four rows' byte b
// of the same column; x = sum_b s_b 2^(8b) (mod 2^32), s_b signed.
// KSK [row][col] (int32) -> four signed-byte planes, x = sum_b s_b 2^(8b)
// mod 2^32 (s_b = byte b of x plus the carry out of byte b - 1, taken in
// [-128, 127]), laid out [plane][col][group] with four consecutive rows of a
// group packed one byte each (the keyswitch's MFMA B fragments are 16
// contiguous bytes).  The transpose runs on a few host threads, each taking
// a range of columns; blocks of 64 groups keep reads and writes in cache.
inline void ksk_planes(const ygpu_params& p, const int32_t* ksk, std::vector<uint32_t>& planes) {
    const size_t cols = static_cast<size_t>(p.n) + 1;
    const size_t kskRows = static_cast<size_t>(p.k) * p.N * p.ks_level;
    const size_t groups = (kskRows + 3) / 4;
    planes.assign(groups * 4 * cols, 0u);
    uint32_t* out = planes.data();
    auto work = [&](size_t cLo, size_t cHi) {
        constexpr size_t GB = 64;
        for (size_t g0 = 0; g0 < groups; g0 += GB) {
            const size_t g1 = std::min(groups, g0 + GB);
            for (size_t c = cLo; c < cHi; c++) {
                uint32_t* o[4];
                for (int b = 0; b < 4; b++) o[b] = out + (b * cols + c) * groups;
                for (size_t g = g0; g < g1; g++) {
                    uint32_t w[4] = {0, 0, 0, 0};
                    for (size_t r = 0; r < 4 && g * 4 + r < kskRows; r++) {
                        uint32_t x = static_cast<uint32_t>(ksk[(g * 4 + r) * cols + c]);
                        for (int b = 0; b < 4; b++) {
                            const uint32_t byte = x & 0xFFu;
                            w[b] |= byte << (8 * r);
                            x = (x >> 8) + (byte >> 7);  // (x - s_b) / 256 with s_b = byte - 256 [byte >= 128]
                        }
                    }
                    for (int b = 0; b < 4; b++) o[b][g] = w[b];
                }
            }
        }
    };
    const size_t nt = std::min<size_t>(8, std::max<size_t>(1, std::min<size_t>(std::thread::hardware_concurrency(), cols / 64)));
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; t++) pool.emplace_back(work, cols * t / nt, cols * (t + 1) / nt);
    work(0, cols / nt);
    for (auto& th : pool) th.join();
}

inline int prepare_key_host(ygpu_ctx* ctx, const uint64_t* bsk_coeff, const int32_t* ksk, HostKeyForm& out) {
    const ygpu_params& p = ctx->p;
    out.s2_level = 0;
    out.bsk_bytes = bsk_polys(p) * ctx->np * p.N * sizeof(double);
    out.packed.assign(packed_words(p), 0);
    const int64_t bad = pack_residues(bsk_coeff, bsk_polys(p), p, out.packed.data());
    if (bad >= 0) return fail(YGPU_ERR_INVALID_ARG, "BSK coefficient " + std::to_string(bad) + " is not below q_ntt");
    ksk_planes(p, ksk, out.planes);
    return YGPU_OK;
}

// Succinct (WWL+24) key: bsk_half [i][lvl][col][N] (the body-digit rows
// cPrime[lvl] of every step) and s2 [l2][col][N] (TRLEV of s^2), residues mod q.
inline int prepare_key_host_succinct(ygpu_ctx* ctx, const uint64_t* bsk_half, const uint64_t* s2, int s2_level, const int32_t* ksk,
                                     HostKeyForm& out) {
    const ygpu_params& p = ctx->p;
    int rc = check_s2_level(p, s2_level);
    if (rc) return rc;
    out.s2_level = s2_level;
    out.bsk_bytes = bsk_polys(p) * ctx->np * p.N * sizeof(double);
    out.packed.assign(packed_words_succinct(p, s2_level), 0);
    int64_t bad = pack_residues(bsk_half, half_polys(p), p, out.packed.data());
    if (bad >= 0) return fail(YGPU_ERR_INVALID_ARG, "succinct BSK coefficient " + std::to_string(bad) + " is not below q_ntt");
    bad = pack_residues(s2, s2_polys(p, s2_level), p, out.packed.data() + half_polys(p) * poly_words(p));
    if (bad >= 0) return fail(YGPU_ERR_INVALID_ARG, "s2 coefficient " + std::to_string(bad) + " is not below q_ntt");
    ksk_planes(p, ksk, out.planes);
    return YGPU_OK;
}

inline BskExpandArgs expand_args(const ygpu_ctx* ctx, const uint64_t* d_packed, int polys, double* d_out) {
    BskExpandArgs a{};
    a.packed = d_packed;
    a.bits = key_bits(ctx->p.q_ntt);
    a.q = ctx->p.q_ntt;
    a.polys = polys;
    a.np = ctx->np;
    a.out = d_out;
    a.tw = ctx->d_tw;
    for (int k = 0; k < ctx->np; k++) {
        a.prime[k] = ctx->prime[k];
        a.p[k] = ctx->primes[k];
        a.ninv[k] = centred_double(ctx->ntt[k].ninv, ctx->ntt[k].p);
    }
    return a;
}

// Enqueues the expansion of a packed BSK (device memory) into the engine form.
inline hipError_t expand_bsk(const ygpu_ctx* ctx, const uint64_t* d_packed, double* d_out, hipStream_t s) {
    return launch_bsk_expand(ctx->logn, expand_args(ctx, d_packed, static_cast<int>(bsk_polys(ctx->p)), d_out), s);
}

// Enqueues the expansion of a key staged at d_stage (stage_bytes of its form):
// full keys as above; succinct keys by expanding the s^2 TRLEV into the staging
// tail and deriving every row with k_bsk_derive.
inline hipError_t expand_key(const ygpu_ctx* ctx, int s2_level, void* d_stage, double* d_out, hipStream_t s) {
    const uint64_t* packed = static_cast<const uint64_t*>(d_stage);
    if (s2_level <= 0) return expand_bsk(ctx, packed, d_out, s);
    const ygpu_params& p = ctx->p;
    double* s2x = reinterpret_cast<double*>(static_cast<char*>(d_stage) + s2_stage_offset(p, s2_level));
    hipError_t e = launch_bsk_expand(ctx->logn, expand_args(ctx, packed + half_polys(p) * poly_words(p),
                                                            static_cast<int>(s2_polys(p, s2_level)), s2x), s);
    if (e != hipSuccess) return e;
    const BskExpandArgs x = expand_args(ctx, packed, 0, d_out);
    BskDeriveArgs a{};
    a.half = packed;
    a.s2 = s2x;
    a.bits = x.bits;
    a.q = x.q;
    a.n = p.n;
    a.level = p.level;
    a.s2_level = s2_level;
    a.radix_bits = p.radix_bits;
    a.np = ctx->np;
    a.out = d_out;
    a.tw = ctx->d_tw;
    for (int k = 0; k < 2; k++) { a.prime[k] = x.prime[k]; a.p[k] = x.p[k]; a.ninv[k] = x.ninv[k]; }
    a.crt = ctx->crt;
    return launch_bsk_derive(ctx->logn, ctx->tbits, a, s);
}

int register_key(ygpu_ctx* ctx, const KeyRec& rec, int32_t* key_id);

// A context issuing its work on `stream`, which the caller owns and keeps
// alive until ygpu_destroy; nullptr creates a stream of the context's own
// (ygpu_init).  Initialisation issues no work on the null stream.
int init_context(const ygpu_params* params, int32_t device, hipStream_t stream, ygpu_ctx** out);

inline bool params_equal(const ygpu_params& a, const ygpu_params& b) {
    return a.torus_bits == b.torus_bits && a.n == b.n && a.k == b.k && a.N == b.N && a.radix_bits == b.radix_bits &&
           a.level == b.level && a.ks_radix_bits == b.ks_radix_bits && a.ks_level == b.ks_level &&
           a.ks_width_bits == b.ks_width_bits && a.lwe_q_bits == b.lwe_q_bits && a.q_ntt == b.q_ntt;
}

// Words of the KSK byte planes (ksk_planes) of a parameter set.
inline size_t ksk_plane_words(const ygpu_params& p) {
    const size_t rows = static_cast<size_t>(p.k) * p.N * p.ks_level;
    return (rows + 3) / 4 * 4 * (static_cast<size_t>(p.n) + 1);
}

// Bytes a key occupies on the device, without building it.
inline size_t key_device_bytes(const ygpu_ctx* ctx) {
    const ygpu_params& p = ctx->p;
    const size_t bsk = static_cast<size_t>(p.n) * p.level * (p.k + 1) * (p.k + 1) * ctx->np * p.N * sizeof(double);
    const size_t rows = static_cast<size_t>(p.k) * p.N * p.ks_level;
    return bsk + (rows + 3) / 4 * 4 * (static_cast<size_t>(p.n) + 1) * sizeof(uint32_t);
}


}  // namespace ygpu_rt
