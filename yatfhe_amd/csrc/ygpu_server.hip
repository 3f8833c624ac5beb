// Continuous-batching bootstrapping server on one GPU.
//
// The reference serves bootstraps from a CPU worker pool: a shared key cache
// with single-flight loads (ya_benchmark/include/blindrotate_serving.h:25-111),
// a request scheduler (132-273) and one blind rotation per request and worker
// (ya_benchmark/blindrotate_serving.cpp:121-353).  On the GPU one bootstrap per
// launch would leave the device idle, so this server batches instead:
//
//  * Requests queue per tenant within their parameter set.  A dispatcher
//    thread per parameter set forms a batch whenever one of its lanes is free:
//    it walks the tenants with pending requests, oldest first, takes the
//    requests of every tenant whose key is resident and loaded (up to
//    max_batch) and launches one key-grouped blind rotation over them, then
//    extract + keyswitch per tenant, on the lane's stream.  Requests arriving
//    while batches run form later batches, so batches of different sizes (and
//    of both torus widths) overlap on the device.
//  * Keys live in one device arena of a fixed byte budget (allocated once: no
//    hipMalloc/hipFree while serving).  A tenant's key is loaded on demand from
//    its device-ready form in pinned host memory (the spill tier) by an
//    asynchronous copy on the parameter set's copy stream; a load watcher
//    thread marks it ready when the copy completes.  Requests whose key is not
//    ready stay queued, so batches never wait for other tenants' loads, and a
//    key is loaded once however many requests ask for it (single flight).
//    When the arena is full, a contiguous window of idle keys is evicted:
//    the one whose keys imply the least reload traffic (decayed request rate
//    times transfer bytes), or, under the LRU policy, whose newest use is
//    oldest; keys used by batches in flight or being loaded are never evicted.
//  * Admission by device memory: a tenant whose key cannot be placed (every
//    window holds a key in use) stays queued and is retried when a batch or a
//    load completes.
//  * A completion thread per parameter set retires lanes in launch order,
//    copies results to the callers' buffers and wakes waiters.
//  * Which lane batches which requests follows the reference Scheduler's
//    modes (blindrotate_serving.h:132-273), lanes standing for its workers:
//    shared (any lane, any tenant), affinity (tenant t on lane t % lanes) and
//    soft (affinity with spill on admission and stealing by idle lanes).
//
// The PBS itself is a deterministic function of (key, ciphertext, test
// polynomial), and the batched kernels compute every ciphertext independently,
// so a request's result does not depend on what it was batched with.
#include "ygpu_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <random>
#include <thread>
#include <unordered_map>

namespace {

using Clock = std::chrono::steady_clock;

inline int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now().time_since_epoch()).count();
}

// A key's transfer form in pinned host memory: bit-packed BSK residues (full
// key, or succinct key with s2_level > 0) and KSK planes.  bsk_bytes is the
// engine-form BSK size on the device.  A file-backed form (the lowest tier)
// holds only its key file's path until a load needs it: it is then read into
// one of its parameter set's pinned slots, which it keeps until the slot is
// reused for another file.
struct PinnedKey {
    uint64_t* packed = nullptr;
    uint32_t* planes = nullptr;
    size_t packed_bytes = 0, planes_bytes = 0, bsk_bytes = 0;
    int s2_level = 0;
    std::string path;      // key file (file-backed forms only)
    int pset = -1;
    int slot = -1;         // pinned slot holding it (file-backed, read or being read)
    bool present = true;   // packed and planes hold the form
    bool reading = false;  // queued for the reader or being read
    int copies = 0;        // device loads in flight that read packed and planes
    bool unused = false;   // read, and no load has used it yet (its slot is kept)
    int read_rc = 0;       // the last read's status, and its message
    std::string read_err;
};

// A pinned buffer of a parameter set for file-backed forms (a full key's
// transfer form; a succinct form is smaller).
struct FileSlot {
    uint64_t* packed = nullptr;
    uint32_t* planes = nullptr;
    int form = -1;  // the form it holds or is being read into (-1: free)
    uint64_t last_use = 0;
};

struct Tenant {
    int pset = 0;
    int form = 0;
    size_t bytes = 0;      // arena bytes (BSK + KSK planes, each 256-byte aligned)
    bool resident = false; // holds arena memory (loaded or loading)
    bool ready = false;    // the load has completed
    size_t off = 0;        // arena offset when resident
    hipEvent_t loaded = nullptr;
    int inuse = 0;         // batches in flight holding the key
    bool fresh = false;    // loaded but not yet used by a batch
    uint64_t last_use = 0;
    double freq = 0;       // request rate, decayed with time constant FREQ_TAU_NS
    int64_t freq_ns = 0;   // when freq was last brought up to date
    size_t load_bytes = 0; // host-to-device bytes of one load
    std::deque<ygpu_request*> pending;
    std::deque<int> pending_lane;  // per pending request: the lane it is queued on (-1: any)
    bool listed = false;   // in its parameter set's `order`
};

struct Region {
    size_t off, size;
    int tenant;  // -1: free
};

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int32_t* h_in = nullptr;
    void* h_v = nullptr;
    int32_t* h_out = nullptr;
    int32_t* h_slots = nullptr;
    uint32_t* h_status = nullptr;  // the lane's barrier-failure report (pinned)
    // device addresses of the pinned staging buffers (mapped host memory)
    const int32_t* hd_in = nullptr;
    const void* hd_v = nullptr;
    int32_t* hd_out = nullptr;
    const int32_t* hd_slots = nullptr;
    uint32_t* hd_status = nullptr;
    int32_t* d_in = nullptr;
    void* d_v = nullptr;
    int32_t* d_out = nullptr;
    void* d_acc = nullptr;
    int32_t* d_slots = nullptr;
    int32_t* d_partial = nullptr;
    bool busy = false;
    int rc = 0;
    int64_t launch_ns = 0;
    std::vector<ygpu_request*> reqs;
    std::vector<int> tenants;
};

struct PSet {
    ygpu_ctx* ctx = nullptr;
    ygpu_params p{};
    size_t tbytes = 4;
    std::vector<Lane> lanes;
    std::deque<int> order;            // tenants with pending requests, oldest first
    int64_t queued = 0;               // pending requests
    std::vector<int64_t> lane_depth;  // pending requests queued on each lane (affinity, soft)
    hipStream_t copy_stream = nullptr;  // the server's (shared by the parameter sets)
    void* d_stage = nullptr;          // packed BSK of the load being expanded (either form)
    std::deque<int> loads;            // tenants being loaded, in copy order
    std::deque<int> inflight;         // lanes in launch order
    std::condition_variable cv_work;  // requests, a free lane, a ready key or free memory
    std::condition_variable cv_done;  // a batch was launched
    std::condition_variable cv_load;  // a load was started
    std::thread dispatcher, completer, watcher;
    bool dispatcher_done = false;
    std::vector<FileSlot> slots;  // pinned slots for file-backed forms (made with the first)
    int slot_count = 2;
};

size_t align256(size_t x) { return (x + 255) & ~size_t{255}; }

}  // namespace

struct ygpu_server {
    ygpu_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_req;
    std::vector<std::unique_ptr<PSet>> psets;
    std::vector<Tenant> tenants;
    std::vector<PinnedKey> forms;
    char* d_arena = nullptr;
    size_t arena_bytes = 0;
    std::vector<Region> regions;
    const double** d_bsk_tab = nullptr;
    const int32_t** d_ksk_tab = nullptr;
    const double** h_bsk_tab = nullptr;   // pinned mirrors (copy sources)
    const int32_t** h_ksk_tab = nullptr;
    hipStream_t copy_stream = nullptr;  // key loads of every parameter set; the contexts' stream
    std::deque<int> read_queue;         // file-backed forms to read, in request order
    std::condition_variable cv_read;
    std::thread reader;
    bool started = false;
    bool stop = false;
    uint64_t tick = 0;
    int64_t load_change_ns = 0;  // when the count of loads under way last changed
    ygpu_server_stats st{};
};

namespace {

using namespace ygpu_rt;

int server_fail_hip(hipError_t e, const char* what) {
    return fail(YGPU_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void notify_all_psets(ygpu_server* s) {
    for (auto& q : s->psets) q->cv_work.notify_all();
}

// ---- arena ------------------------------------------------------------------

void merge_free(ygpu_server* s) {
    std::vector<Region> out;
    for (const Region& r : s->regions) {
        if (!out.empty() && out.back().tenant < 0 && r.tenant < 0) out.back().size += r.size;
        else out.push_back(r);
    }
    s->regions.swap(out);
}

void evict(ygpu_server* s, int t) {
    Tenant& tn = s->tenants[t];
    tn.resident = false;
    tn.ready = false;
    s->st.key_evictions++;
    s->st.resident_bytes -= static_cast<int64_t>(tn.bytes);
    s->st.resident_keys--;
    for (Region& r : s->regions)
        if (r.tenant == t) r.tenant = -1;
}

// A key may be evicted once loaded and used: a freshly loaded key serves its
// tenant's waiting requests at least once first, so two parameter sets short
// of memory cannot keep evicting each other's new keys (every load makes
// progress).
bool evictable(const Tenant& tn) { return tn.inuse == 0 && tn.ready && !tn.fresh; }

constexpr double FREQ_TAU_NS = 1e9;

double decayed_freq(const Tenant& tn, int64_t now) {
    return tn.freq * std::exp(-static_cast<double>(now - tn.freq_ns) / FREQ_TAU_NS);
}

// Cost of evicting a key.  Keys with waiting requests come last under both
// policies.  LRU: the newest use (a window's cost is its newest key);
// FREQ_BYTES: the reload traffic the key's request rate implies (a window's
// cost is the sum over its keys).
double evict_cost(const ygpu_server* s, const Tenant& tn, int64_t now) {
    const double waiting = tn.pending.empty() ? 0.0 : 1e30;
    if (s->cfg.eviction == YGPU_EVICT_LRU) return static_cast<double>(tn.last_use) + waiting;
    return decayed_freq(tn, now) * static_cast<double>(tn.load_bytes) + waiting;
}

// Finds `bytes` of arena: best-fitting free region, else the contiguous window
// of free regions and idle keys of least eviction cost.  Returns false when
// every window holds a key in use or being loaded.
bool arena_alloc(ygpu_server* s, size_t bytes, int tenant, size_t* off) {
    auto best_free = [&]() {
        int best = -1;
        for (size_t i = 0; i < s->regions.size(); i++) {
            const Region& r = s->regions[i];
            if (r.tenant < 0 && r.size >= bytes && (best < 0 || r.size < s->regions[best].size)) best = static_cast<int>(i);
        }
        return best;
    };
    int best = best_free();
    if (best < 0) {
        const int64_t now = now_ns();
        const bool lru = s->cfg.eviction == YGPU_EVICT_LRU;
        size_t bi = 0, bj = 0;
        bool found = false;
        double bcost = 0;
        size_t bevict = SIZE_MAX;
        for (size_t i = 0; i < s->regions.size(); i++) {
            size_t sum = 0, evicted = 0;
            double cost = 0;
            size_t j = i;
            for (; j < s->regions.size() && sum < bytes; j++) {
                const Region& r = s->regions[j];
                if (r.tenant >= 0) {
                    const Tenant& tn = s->tenants[r.tenant];
                    if (!evictable(tn)) break;
                    const double c = evict_cost(s, tn, now);
                    cost = lru ? std::max(cost, c) : cost + c;
                    evicted += r.size;
                }
                sum += r.size;
            }
            if (sum >= bytes && (!found || cost < bcost || (cost == bcost && evicted < bevict))) {
                found = true;
                bcost = cost;
                bevict = evicted;
                bi = i;
                bj = j;
            }
        }
        if (!found) return false;
        for (size_t k = bi; k < bj; k++)
            if (s->regions[k].tenant >= 0) evict(s, s->regions[k].tenant);
        merge_free(s);
        best = best_free();
        if (best < 0) return false;  // not reached: the window merged into one free region
    }
    Region r = s->regions[best];
    *off = r.off;
    Region used{r.off, bytes, tenant};
    if (r.size == bytes) {
        s->regions[best] = used;
    } else {
        s->regions[best] = Region{r.off + bytes, r.size - bytes, -1};
        s->regions.insert(s->regions.begin() + best, used);
    }
    return true;
}

// Starts loading `t` into the arena.  Returns 1 when no memory can be freed
// now, 0 when the load is enqueued, <0 on error.  Server lock held.
// Keeps the count of key loads under way, its peak and its integral over
// time (server lock held).
void load_count(ygpu_server* s, int delta) {
    const int64_t now = now_ns();
    if (s->load_change_ns) s->st.load_inflight_ns += s->st.loads_in_flight * (now - s->load_change_ns);
    s->load_change_ns = now;
    s->st.loads_in_flight += delta;
    s->st.peak_loads_in_flight = std::max(s->st.peak_loads_in_flight, s->st.loads_in_flight);
}

void fail_pending(ygpu_server* s, PSet& ps, int t, int rc);

// Queues the read of file-backed form `fi` into a slot of `ps`: a free slot,
// else the least recently used one whose form no load is reading.  Returns 2
// when the read is queued or under way, 3 when every slot is busy now, <0 on
// error.  Server lock held.
int request_read(ygpu_server* s, PSet& ps, int fi) {
    PinnedKey& f = s->forms[fi];
    if (f.reading) return 2;
    int best = -1;
    for (int i = 0; i < static_cast<int>(ps.slots.size()); i++) {
        const FileSlot& sl = ps.slots[i];
        if (sl.form < 0) {
            best = i;
            break;
        }
        const PinnedKey& o = s->forms[sl.form];
        if (o.reading || o.copies > 0 || o.unused) continue;
        if (best < 0 || sl.last_use < ps.slots[best].last_use) best = i;
    }
    if (best < 0) return ps.slots.empty() ? fail(YGPU_ERR_INVALID_ARG, "no pinned slots for file-backed keys") : 3;
    FileSlot& sl = ps.slots[best];
    if (sl.form >= 0) {
        PinnedKey& o = s->forms[sl.form];
        o.present = false;
        o.packed = nullptr;
        o.planes = nullptr;
        o.slot = -1;
    }
    sl.form = fi;
    sl.last_use = ++s->tick;
    f.slot = best;
    f.reading = true;
    s->read_queue.push_back(fi);
    s->cv_read.notify_one();
    return 2;
}

// Reads a key file into a slot: the file's residues are packed and its KSK
// cut into byte planes exactly as for a key handed over in memory.
int read_form(const ygpu_ctx* ctx, const std::string& path, int s2_level, size_t packed_bytes, size_t planes_bytes,
              uint64_t* packed, uint32_t* planes) {
    const ygpu_params& p = ctx->p;
    int32_t form = 0;
    int rc = ygpu_key_file_form(path.c_str(), &form);
    if (rc) return rc;
    if (form != s2_level) return fail(YGPU_ERR_CORRUPT, path + ": key form changed since the tenant was registered");
    const size_t half = half_polys(p) * p.N;
    const size_t bskWords = s2_level ? half + s2_polys(p, s2_level) * p.N : bsk_polys(p) * p.N;
    std::vector<uint64_t> bsk(bskWords);
    std::vector<int32_t> ksk(static_cast<size_t>(p.k) * p.N * p.ks_level * (p.n + 1));
    HostKeyForm hk;
    ygpu_ctx* c = const_cast<ygpu_ctx*>(ctx);  // the preparers read only its parameters
    if (s2_level == 0) {
        rc = ygpu_key_file_read(path.c_str(), &p, bsk.data(), ksk.data());
        if (!rc) rc = prepare_key_host(c, bsk.data(), ksk.data(), hk);
    } else {
        rc = ygpu_key_file_read_succinct(path.c_str(), &p, bsk.data(), bsk.data() + half, ksk.data());
        if (!rc) rc = prepare_key_host_succinct(c, bsk.data(), bsk.data() + half, s2_level, ksk.data(), hk);
    }
    if (rc) return rc;
    if (hk.packed.size() * sizeof(uint64_t) != packed_bytes || hk.planes.size() * sizeof(uint32_t) != planes_bytes)
        return fail(YGPU_ERR_CORRUPT, path + ": key size differs from the registered one");
    std::memcpy(packed, hk.packed.data(), packed_bytes);
    std::memcpy(planes, hk.planes.data(), planes_bytes);
    return YGPU_OK;
}

// Reads file-backed forms as loads ask for them, one at a time, without the
// server lock (the reference's DiskReader on a cache miss,
// cache_manager.h:13-51, blindrotate_serving.cpp:210-219).  A failed read
// fails the pending requests of every tenant of that form.
void reader_loop(ygpu_server* s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_read.wait(lk, [&] { return s->stop || !s->read_queue.empty(); });
        if (s->read_queue.empty()) break;
        const int fi = s->read_queue.front();
        s->read_queue.pop_front();
        const PinnedKey f = s->forms[fi];  // a copy: forms may grow meanwhile
        PSet& ps = *s->psets[f.pset];
        const FileSlot sl = ps.slots[f.slot];
        lk.unlock();
        const int64_t t0 = now_ns();
        const int rc = read_form(ps.ctx, f.path, f.s2_level, f.packed_bytes, f.planes_bytes, sl.packed, sl.planes);
        const std::string err = rc ? ygpu_last_error() : std::string();
        const int64_t t1 = now_ns();
        lk.lock();
        PinnedKey& g = s->forms[fi];
        g.reading = false;
        g.read_rc = rc;
        g.read_err = err;
        s->st.file_read_ns += t1 - t0;
        if (rc == 0) {
            g.present = true;
            g.unused = true;
            g.packed = sl.packed;
            g.planes = sl.planes;
            s->st.file_reads++;
            s->st.file_read_bytes += static_cast<int64_t>(f.packed_bytes + f.planes_bytes);
        } else {
            ps.slots[g.slot].form = -1;
            g.slot = -1;
            for (size_t t = 0; t < s->tenants.size(); t++)
                if (s->tenants[t].form == fi && !s->tenants[t].resident) fail_pending(s, ps, static_cast<int>(t), fail(rc, err));
        }
        notify_all_psets(s);
        s->cv_req.notify_all();
    }
}

int start_load(ygpu_server* s, PSet& ps, int t) {
    Tenant& tn = s->tenants[t];
    if (!s->forms[tn.form].present) return request_read(s, ps, tn.form);
    size_t off = 0;
    if (!arena_alloc(s, tn.bytes, t, &off)) return 1;
    PinnedKey& f = s->forms[tn.form];
    char* base = s->d_arena + off;
    s->h_bsk_tab[t] = reinterpret_cast<const double*>(base);
    s->h_ksk_tab[t] = reinterpret_cast<const int32_t*>(base + align256(f.bsk_bytes));
    hipStream_t cs = ps.copy_stream;
    // packed residues -> staging -> engine form; loads run in order on the
    // server's copy stream, so one staging buffer per parameter set serves them all
    hipError_t e = hipMemcpyAsync(ps.d_stage, f.packed, f.packed_bytes, hipMemcpyHostToDevice, cs);
    if (e == hipSuccess) e = expand_key(ps.ctx, f.s2_level, ps.d_stage, reinterpret_cast<double*>(base), cs);
    if (e == hipSuccess) e = hipMemcpyAsync(base + align256(f.bsk_bytes), f.planes, f.planes_bytes, hipMemcpyHostToDevice, cs);
    if (e == hipSuccess) e = hipMemcpyAsync(s->d_bsk_tab + t, s->h_bsk_tab + t, sizeof(void*), hipMemcpyHostToDevice, cs);
    if (e == hipSuccess) e = hipMemcpyAsync(s->d_ksk_tab + t, s->h_ksk_tab + t, sizeof(void*), hipMemcpyHostToDevice, cs);
    if (e == hipSuccess) e = hipEventRecord(tn.loaded, cs);
    if (e != hipSuccess) {
        for (Region& r : s->regions)
            if (r.tenant == t) r.tenant = -1;
        merge_free(s);
        return server_fail_hip(e, "key load");
    }
    tn.resident = true;
    tn.ready = false;
    tn.fresh = true;
    tn.off = off;
    f.copies++;
    f.unused = false;
    if (f.slot >= 0) ps.slots[f.slot].last_use = ++s->tick;
    tn.last_use = ++s->tick;
    ps.loads.push_back(t);
    load_count(s, +1);
    ps.cv_load.notify_one();
    s->st.key_loads++;
    s->st.bytes_loaded += static_cast<int64_t>(f.packed_bytes + f.planes_bytes);
    s->st.resident_bytes += static_cast<int64_t>(tn.bytes);
    s->st.resident_keys++;
    s->st.peak_resident_bytes = std::max(s->st.peak_resident_bytes, s->st.resident_bytes);
    return 0;
}

// ---- lanes --------------------------------------------------------------------

int lane_alloc(const ygpu_server* s, PSet& ps, Lane& L) {
    const ygpu_params& p = ps.p;
    const size_t B = static_cast<size_t>(s->cfg.max_batch);
    const size_t lwe = sizeof(int32_t) * (p.n + 1);
    YG_HIP(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    YG_HIP(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
    YG_HIP(hipHostMalloc(reinterpret_cast<void**>(&L.h_in), B * lwe, hipHostMallocDefault));
    YG_HIP(hipHostMalloc(&L.h_v, B * ps.tbytes * p.N, hipHostMallocDefault));
    YG_HIP(hipHostMalloc(reinterpret_cast<void**>(&L.h_out), B * lwe, hipHostMallocDefault));
    YG_HIP(hipHostMalloc(reinterpret_cast<void**>(&L.h_slots), B * sizeof(int32_t), hipHostMallocDefault));
    YG_HIP(hipHostMalloc(reinterpret_cast<void**>(&L.h_status), 64, hipHostMallocDefault));
    *L.h_status = 0u;
    void* dp = nullptr;
    YG_HIP(hipHostGetDevicePointer(&dp, L.h_in, 0));
    L.hd_in = static_cast<const int32_t*>(dp);
    YG_HIP(hipHostGetDevicePointer(&dp, L.h_v, 0));
    L.hd_v = dp;
    YG_HIP(hipHostGetDevicePointer(&dp, L.h_out, 0));
    L.hd_out = static_cast<int32_t*>(dp);
    YG_HIP(hipHostGetDevicePointer(&dp, L.h_slots, 0));
    L.hd_slots = static_cast<const int32_t*>(dp);
    YG_HIP(hipHostGetDevicePointer(&dp, L.h_status, 0));
    L.hd_status = static_cast<uint32_t*>(dp);
    YG_HIP(hipMalloc(reinterpret_cast<void**>(&L.d_in), B * lwe));
    YG_HIP(hipMalloc(&L.d_v, B * ps.tbytes * p.N));
    YG_HIP(hipMalloc(reinterpret_cast<void**>(&L.d_out), B * lwe));
    YG_HIP(hipMalloc(&L.d_acc, B * ps.tbytes * (p.k + 1) * p.N));
    YG_HIP(hipMalloc(reinterpret_cast<void**>(&L.d_slots), B * sizeof(int32_t)));
    YG_HIP(hipMalloc(reinterpret_cast<void**>(&L.d_partial), ks_workspace_bytes(ps.ctx, static_cast<int>(B))));
    return YGPU_OK;
}

void lane_free(Lane& L) {
    if (L.stream) hipStreamDestroy(L.stream);
    if (L.done) hipEventDestroy(L.done);
    for (void* h : {static_cast<void*>(L.h_in), L.h_v, static_cast<void*>(L.h_out), static_cast<void*>(L.h_slots),
                    static_cast<void*>(L.h_status)})
        if (h) hipHostFree(h);
    for (void* d : {static_cast<void*>(L.d_in), L.d_v, static_cast<void*>(L.d_out), L.d_acc, static_cast<void*>(L.d_slots),
                    static_cast<void*>(L.d_partial)})
        if (d) hipFree(d);
    L = Lane{};
}

// Stages the lane's requests (grouped by tenant, keys loaded) and enqueues
// the pipeline on the lane's stream.  Runs without the server lock.
int launch_lane(ygpu_server* s, PSet& ps, Lane& L) {
    const ygpu_params& p = ps.p;
    const int count = static_cast<int>(L.reqs.size());
    const size_t lwe = sizeof(int32_t) * (p.n + 1);
    const size_t vbytes = ps.tbytes * p.N;
    for (int c = 0; c < count; c++) {
        const ygpu_request* r = L.reqs[c];
        std::memcpy(L.h_in + static_cast<size_t>(c) * (p.n + 1), r->lwe_in, lwe);
        std::memcpy(static_cast<char*>(L.h_v) + c * vbytes, r->v, vbytes);
        L.h_slots[c] = r->tenant;
    }
    // staging copies by kernel over mapped pinned memory (not behind key loads)
    YG_HIP(launch_copy_words(L.hd_in, L.d_in, count * lwe / 4, L.stream));
    YG_HIP(launch_copy_words(L.hd_v, L.d_v, count * vbytes / 4, L.stream));
    YG_HIP(launch_copy_words(L.hd_slots, L.d_slots, count, L.stream));
    BrArgs a = br_args(ps.ctx, count, L.d_in, L.d_v, p.N, L.d_acc, s->d_bsk_tab, L.d_slots);
    a.status = L.hd_status;  // this batch's own failure report
    int rc = run_br(ps.ctx, batch_tile(ps.ctx, count), a, L.stream);
    if (rc) return rc;
    // one keyswitch launch; the kernel follows the per-ciphertext tenant slots
    rc = run_ks(ps.ctx, count, L.d_acc, s->d_ksk_tab, L.d_slots, L.d_out, nullptr, L.d_partial, L.stream);
    if (rc) return rc;
    YG_HIP(launch_copy_words(L.d_out, L.hd_out, count * lwe / 4, L.stream));
    YG_HIP(hipEventRecord(L.done, L.stream));
    return YGPU_OK;
}

void finish_request(ygpu_request* r, int status, int64_t launch, int64_t done, int batch) {
    r->launch_ns = launch;
    r->done_ns = done;
    r->batch_size = batch;
    r->status = status;
}

// Whether anything in progress will free arena memory: batches in flight,
// loads in flight, or loaded keys waiting for their first batch.
bool any_busy(const ygpu_server* s) {
    for (const auto& q : s->psets) {
        if (!q->inflight.empty() || !q->loads.empty()) return true;
        for (const auto& l : q->lanes)
            if (l.busy) return true;
    }
    for (const Tenant& t : s->tenants)
        if (t.fresh) return true;
    return false;
}

// Fails every pending request of `t` (server lock held).
void fail_pending(ygpu_server* s, PSet& ps, int t, int rc) {
    const int64_t now = now_ns();
    Tenant& tn = s->tenants[t];
    tn.fresh = false;
    for (ygpu_request* r : tn.pending) finish_request(r, rc, now, now, 0);
    for (int l : tn.pending_lane)
        if (l >= 0) ps.lane_depth[l]--;
    ps.queued -= static_cast<int64_t>(tn.pending.size());
    s->st.failed += static_cast<int64_t>(tn.pending.size());
    tn.pending.clear();
    tn.pending_lane.clear();
}

// Queue lane of a new request of tenant t (server lock held): -1 under SHARED,
// else the owner lane t % lanes, or under SOFT the shortest lane when the
// owner already holds spill_threshold requests and the shortest is shorter
// by more than one (Scheduler::push, blindrotate_serving.h:154-175).
int queue_lane(ygpu_server* s, PSet& ps, int t) {
    if (s->cfg.sched == YGPU_SCHED_SHARED) return -1;
    const int nl = static_cast<int>(ps.lanes.size());
    int lane = t % nl;
    if (s->cfg.sched == YGPU_SCHED_SOFT && ps.lane_depth[lane] >= s->cfg.spill_threshold) {
        int alt = 0;
        for (int l = 1; l < nl; l++)
            if (ps.lane_depth[l] < ps.lane_depth[alt]) alt = l;
        if (ps.lane_depth[alt] + 1 < ps.lane_depth[lane]) {
            lane = alt;
            s->st.spilled++;
        }
    }
    return lane;
}

// Fills lane L's batch from the ready tenants, oldest first (server lock
// held): every pending request when `want` < 0, else the requests queued on
// lane `want`; at most `cap`.  Tenants left with requests stay listed.
void form_batch(ygpu_server* s, PSet& ps, Lane& L, int want, int cap) {
    L.reqs.clear();
    L.tenants.clear();
    std::deque<int> keep;
    for (int t : ps.order) {
        Tenant& tn = s->tenants[t];
        if (tn.pending.empty()) {
            tn.listed = false;
            continue;
        }
        const int room = cap - static_cast<int>(L.reqs.size());
        if (!tn.ready || room == 0) {
            keep.push_back(t);
            continue;
        }
        int took = 0;
        if (want < 0) {
            took = std::min<int>(room, static_cast<int>(tn.pending.size()));
            for (int c = 0; c < took; c++) {
                L.reqs.push_back(tn.pending.front());
                if (tn.pending_lane.front() >= 0) ps.lane_depth[tn.pending_lane.front()]--;
                tn.pending.pop_front();
                tn.pending_lane.pop_front();
            }
        } else {
            std::deque<ygpu_request*> restR;
            std::deque<int> restL;
            while (!tn.pending.empty()) {
                ygpu_request* r = tn.pending.front();
                const int l = tn.pending_lane.front();
                tn.pending.pop_front();
                tn.pending_lane.pop_front();
                if (l == want && took < room) {
                    L.reqs.push_back(r);
                    ps.lane_depth[l]--;
                    took++;
                } else {
                    restR.push_back(r);
                    restL.push_back(l);
                }
            }
            tn.pending.swap(restR);
            tn.pending_lane.swap(restL);
        }
        if (took) {
            tn.inuse++;
            tn.fresh = false;
            tn.last_use = ++s->tick;
            L.tenants.push_back(t);
        }
        if (tn.pending.empty()) tn.listed = false;
        else keep.push_back(t);
    }
    ps.order.swap(keep);
}

void dispatcher_loop(ygpu_server* s, PSet* ps) {
    hipSetDevice(s->cfg.device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        if (ps->order.empty()) {
            ps->cv_work.wait(lk);
            continue;
        }
        // 1. start loads for the oldest tenants whose keys are absent
        bool mem_blocked = false;
        for (size_t k = 0; k < ps->order.size() && static_cast<int>(ps->loads.size()) < s->cfg.max_loads; k++) {
            const int t = ps->order[k];
            if (s->tenants[t].resident) continue;
            const int rc = start_load(s, *ps, t);
            if (rc == 1) {
                mem_blocked = true;
                break;
            }
            if (rc < 0) fail_pending(s, *ps, t, rc);
        }
        // 2. a batch for a free lane: under SHARED the requests of every
        // ready tenant, oldest tenants first; under AFFINITY and SOFT those
        // queued on that lane, and under SOFT a lane with none of its own
        // takes the longest lane's once it holds steal_threshold requests
        // (Scheduler::pop, blindrotate_serving.h:177-215)
        // Under SHARED one lane takes at most a lane's share of the queue (and
        // never less than a ciphertext per compute unit), so a backlog is
        // spread over the lanes instead of one batch holding it all while the
        // other lanes wait for new arrivals.
        bool launched = false;
        const int nl = static_cast<int>(ps->lanes.size());
        const int share = static_cast<int>(std::min<int64_t>(
            s->cfg.max_batch, std::max<int64_t>(ps->ctx->cu_count, (ps->queued + nl - 1) / nl)));
        for (int li = 0; li < nl && !launched; li++) {
            Lane* L = &ps->lanes[li];
            if (L->busy) continue;
            const int own = s->cfg.sched == YGPU_SCHED_SHARED ? -1 : li;
            form_batch(s, *ps, *L, own, own < 0 ? share : s->cfg.max_batch);
            if (L->reqs.empty() && s->cfg.sched == YGPU_SCHED_SOFT) {
                int lm = -1;
                for (int l = 0; l < nl; l++)
                    if (ps->lane_depth[l] > 0 && (lm < 0 || ps->lane_depth[l] > ps->lane_depth[lm])) lm = l;
                if (lm >= 0 && lm != li && ps->lane_depth[lm] >= s->cfg.steal_threshold) {
                    form_batch(s, *ps, *L, lm, s->cfg.max_batch);
                    s->st.stolen += static_cast<int64_t>(L->reqs.size());
                }
            }
            if (L->reqs.empty()) continue;
            for (ygpu_request* r : L->reqs) r->lane = li;
            ps->queued -= static_cast<int64_t>(L->reqs.size());
            L->busy = true;
            L->launch_ns = now_ns();
            s->st.batches++;
            s->st.batched_cts += static_cast<int64_t>(L->reqs.size());
            lk.unlock();
            L->rc = launch_lane(s, *ps, *L);
            lk.lock();
            ps->inflight.push_back(li);
            ps->cv_done.notify_all();
            launched = true;
        }
        if (launched) continue;  // look for more work at once
        if (mem_blocked) {
            s->st.memory_waits++;
            if (!any_busy(s)) {
                // nothing in flight will free memory: keys that fit the budget never get here
                for (int t : ps->order) fail_pending(s, *ps, t, fail(YGPU_ERR_OUT_OF_MEMORY, "key does not fit the key budget"));
                for (int t : ps->order) s->tenants[t].listed = false;
                ps->order.clear();
                s->cv_req.notify_all();
                continue;
            }
        }
        // nothing to launch now: wait for a request, a lane, a key or memory
        ps->cv_work.wait(lk);
    }
    ps->dispatcher_done = true;
    ps->cv_done.notify_all();
    ps->cv_load.notify_all();
}

// Marks keys ready as their copies complete (copies run in order on the
// parameter set's copy stream).
void watcher_loop(ygpu_server* s, PSet* ps) {
    hipSetDevice(s->cfg.device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        ps->cv_load.wait(lk, [&] { return !ps->loads.empty() || ps->dispatcher_done; });
        if (ps->loads.empty()) break;
        const int t = ps->loads.front();
        hipEvent_t ev = s->tenants[t].loaded;
        lk.unlock();
        const hipError_t e = hipEventSynchronize(ev);
        lk.lock();
        ps->loads.pop_front();
        load_count(s, -1);
        Tenant& tn = s->tenants[t];
        s->forms[tn.form].copies--;  // its pinned form may now be replaced
        if (e != hipSuccess) {
            // the key never arrived: release its memory and fail its requests
            tn.ready = true;
            tn.fresh = false;
            evict(s, t);
            merge_free(s);
            fail_pending(s, *ps, t, server_fail_hip(e, "key load"));
            s->cv_req.notify_all();
        } else {
            tn.ready = true;
        }
        notify_all_psets(s);
        s->cv_req.notify_all();
    }
}

void completer_loop(ygpu_server* s, PSet* ps) {
    hipSetDevice(s->cfg.device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        ps->cv_done.wait(lk, [&] { return !ps->inflight.empty() || ps->dispatcher_done; });
        if (ps->inflight.empty()) break;
        const int li = ps->inflight.front();
        Lane& L = ps->lanes[li];
        lk.unlock();
        int rc = L.rc;
        if (rc == 0) {
            const hipError_t e = hipEventSynchronize(L.done);
            if (e != hipSuccess) {
                rc = server_fail_hip(e, "batch");
            } else if (*reinterpret_cast<volatile uint32_t*>(L.h_status) != 0u) {
                *reinterpret_cast<volatile uint32_t*>(L.h_status) = 0u;
                rc = fail(YGPU_ERR_DEVICE, "a team barrier of the batch's blind rotation timed out: its results are invalid");
            }
        } else {
            hipStreamSynchronize(L.stream);
        }
        const int64_t t = now_ns();
        const int count = static_cast<int>(L.reqs.size());
        if (rc == 0) {
            for (int c = 0; c < count; c++)
                std::memcpy(L.reqs[c]->lwe_out, L.h_out + static_cast<size_t>(c) * (ps->p.n + 1), sizeof(int32_t) * (ps->p.n + 1));
        }
        lk.lock();
        ps->inflight.pop_front();
        for (int tn : L.tenants) s->tenants[tn].inuse--;
        for (int c = 0; c < count; c++) finish_request(L.reqs[c], rc, L.launch_ns, t, count);
        if (rc == 0) s->st.completed += count;
        else s->st.failed += count;
        L.reqs.clear();
        L.tenants.clear();
        L.busy = false;
        s->cv_req.notify_all();
        notify_all_psets(s);
    }
}

// Hardware queues HIP multiplexes the process's streams over
// (GPU_MAX_HW_QUEUES, 4 by default).
int hardware_queues() {
    const char* e = std::getenv("GPU_MAX_HW_QUEUES");
    const int q = e ? std::atoi(e) : 0;
    return q > 0 ? q : 4;
}

// Lanes of each parameter set.  A lane count given in the configuration is
// used for every set.  Otherwise the server's streams are fitted to the
// hardware queues: one queue for the copy stream and one per lane, so that no
// two lanes share a queue (a stream sharing a queue waits behind the other
// stream's kernels, and a short TORUS32 batch behind a long TORUS56 batch
// loses its latency).  Every set has one lane; the remaining queues go to the
// sets of the cheapest blind rotation first (N times the number of primes),
// where overlap of short batches pays most, at most 3 lanes a set.
std::vector<int> lane_plan(const ygpu_server* s) {
    const int P = static_cast<int>(s->psets.size());
    std::vector<int> n(P, s->cfg.lanes > 0 ? s->cfg.lanes : 1);
    if (s->cfg.lanes > 0 || P == 0) return n;
    std::vector<int> order(P);
    for (int i = 0; i < P; i++) order[i] = i;
    auto cost = [&](int i) { return static_cast<int64_t>(s->psets[i]->p.N) * s->psets[i]->ctx->np; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) < cost(b); });
    int spare = hardware_queues() - 1 - P;
    for (int round = 0; round < 2 && spare > 0; round++)
        for (int i : order)
            if (spare > 0) { n[i]++; spare--; }
    return n;
}

int start_threads(ygpu_server* s) {
    if (s->started) return YGPU_OK;
    const std::vector<int> plan = lane_plan(s);
    for (size_t i = 0; i < s->psets.size(); i++) {
        PSet& ps = *s->psets[i];
        if (!ps.lanes.empty()) continue;
        ps.lanes.resize(plan[i]);
        ps.lane_depth.assign(plan[i], 0);
        for (Lane& L : ps.lanes) {
            const int rc = lane_alloc(s, ps, L);
            if (rc) {
                for (Lane& l2 : ps.lanes) lane_free(l2);
                ps.lanes.clear();
                ps.lane_depth.clear();
                return rc;
            }
        }
    }
    s->reader = std::thread(reader_loop, s);
    for (auto& ps : s->psets) {
        ps->dispatcher = std::thread(dispatcher_loop, s, ps.get());
        ps->completer = std::thread(completer_loop, s, ps.get());
        ps->watcher = std::thread(watcher_loop, s, ps.get());
    }
    s->started = true;
    return YGPU_OK;
}

template <class F>
int with_server(ygpu_server* s, F&& f) {
    if (!s) return fail(YGPU_ERR_INVALID_ARG, "null server");
    try {
        return f();
    } catch (const std::exception& e) {
        return fail(YGPU_ERR_INVALID_ARG, e.what());
    }
}

}  // namespace

// ======================================================================= C ABI

extern "C" {

int ygpu_server_create(const ygpu_server_config* cfg, ygpu_server** out) {
    if (!cfg || !out) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(YGPU_ERR_NO_DEVICE, "no HIP device");
    if (cfg->device < 0 || cfg->device >= ndev) return fail(YGPU_ERR_INVALID_ARG, "device index out of range");
    auto s = std::make_unique<ygpu_server>();
    s->cfg = *cfg;
    if (s->cfg.lanes < 0) s->cfg.lanes = 0;
    if (s->cfg.max_batch <= 0) s->cfg.max_batch = 4096;
    if (s->cfg.queue_max <= 0) s->cfg.queue_max = 1 << 20;
    if (s->cfg.max_tenants <= 0) s->cfg.max_tenants = 4096;
    if (s->cfg.max_loads <= 0) s->cfg.max_loads = 4;
    if (s->cfg.eviction != YGPU_EVICT_FREQ_BYTES && s->cfg.eviction != YGPU_EVICT_LRU)
        return fail(YGPU_ERR_INVALID_ARG, "unknown eviction policy");
    if (s->cfg.sched != YGPU_SCHED_SHARED && s->cfg.sched != YGPU_SCHED_AFFINITY && s->cfg.sched != YGPU_SCHED_SOFT)
        return fail(YGPU_ERR_INVALID_ARG, "unknown scheduling mode");
    if (s->cfg.spill_threshold <= 0) s->cfg.spill_threshold = 2;
    if (s->cfg.steal_threshold <= 0) s->cfg.steal_threshold = 2;
    if (s->cfg.lanes > 16 || s->cfg.max_batch > (1 << 16) || s->cfg.max_tenants > (1 << 20))
        return fail(YGPU_ERR_INVALID_ARG, "lanes <= 16, max_batch <= 65536 and max_tenants <= 2^20 required");
    YG_HIP(hipSetDevice(cfg->device));
    size_t freeb = 0, total = 0;
    YG_HIP(hipMemGetInfo(&freeb, &total));
    size_t budget = cfg->key_budget_bytes > 0 ? static_cast<size_t>(cfg->key_budget_bytes) : freeb / 4 * 3;
    budget &= ~size_t{255};
    if (budget == 0 || budget > freeb) return fail(YGPU_ERR_OUT_OF_MEMORY, "key budget exceeds free device memory");
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&s->d_arena), budget);
    if (e != hipSuccess) return fail(YGPU_ERR_OUT_OF_MEMORY, std::string("key arena: ") + hipGetErrorString(e));
    s->arena_bytes = budget;
    s->regions.push_back(Region{0, budget, -1});
    s->cfg.key_budget_bytes = static_cast<int64_t>(budget);
    s->st.key_budget_bytes = static_cast<int64_t>(budget);
    const size_t tab = sizeof(void*) * s->cfg.max_tenants;
    YG_HIP(hipMalloc(reinterpret_cast<void**>(&s->d_bsk_tab), tab));
    YG_HIP(hipMalloc(reinterpret_cast<void**>(&s->d_ksk_tab), tab));
    YG_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->h_bsk_tab), tab, hipHostMallocDefault));
    YG_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->h_ksk_tab), tab, hipHostMallocDefault));
    std::memset(s->h_bsk_tab, 0, tab);
    std::memset(s->h_ksk_tab, 0, tab);
    YG_HIP(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
    YG_HIP(hipMemcpyAsync(s->d_bsk_tab, s->h_bsk_tab, tab, hipMemcpyHostToDevice, s->copy_stream));
    YG_HIP(hipMemcpyAsync(s->d_ksk_tab, s->h_ksk_tab, tab, hipMemcpyHostToDevice, s->copy_stream));
    YG_HIP(hipStreamSynchronize(s->copy_stream));
    *out = s.release();
    return YGPU_OK;
}

int ygpu_server_add_params(ygpu_server* s, const ygpu_params* params, int32_t* pset) {
    return with_server(s, [&] {
        if (!params || !pset) return fail(YGPU_ERR_INVALID_ARG, "null argument");
        std::lock_guard<std::mutex> g(s->mu);
        if (s->started) return fail(YGPU_ERR_INVALID_ARG, "parameter sets must be added before the first request");
        auto ps = std::make_unique<PSet>();
        // the context's own work (set-up, autotuning) runs on the copy stream:
        // the server's streams are the copy stream and the lanes, made at the
        // first request
        int rc = init_context(params, s->cfg.device, s->copy_stream, &ps->ctx);
        if (rc) return rc;
        ps->p = ps->ctx->p;
        ps->tbytes = ps->ctx->torus_bytes();
        ps->copy_stream = s->copy_stream;
        if (hipMalloc(&ps->d_stage, stage_bytes_any(ps->p, ps->ctx->np)) != hipSuccess) {
            ygpu_destroy(ps->ctx);
            return fail(YGPU_ERR_OUT_OF_MEMORY, "key staging buffer");
        }
        *pset = static_cast<int32_t>(s->psets.size());
        s->psets.push_back(std::move(ps));
        return YGPU_OK;
    });
}

int ygpu_server_lanes(ygpu_server* s, int32_t pset, int32_t* lanes) {
    return with_server(s, [&] {
        if (!lanes) return fail(YGPU_ERR_INVALID_ARG, "null argument");
        std::lock_guard<std::mutex> g(s->mu);
        if (pset < 0 || pset >= static_cast<int32_t>(s->psets.size())) return fail(YGPU_ERR_NOT_FOUND, "unknown parameter set");
        *lanes = s->started ? static_cast<int32_t>(s->psets[pset]->lanes.size()) : lane_plan(s)[pset];
        return YGPU_OK;
    });
}

int ygpu_server_set_tile(ygpu_server* s, int32_t pset, int32_t cpw, int32_t twl) {
    return with_server(s, [&] {
        std::lock_guard<std::mutex> g(s->mu);
        if (pset < 0 || pset >= static_cast<int32_t>(s->psets.size())) return fail(YGPU_ERR_NOT_FOUND, "unknown parameter set");
        if (s->started) return fail(YGPU_ERR_INVALID_ARG, "tiles must be set before the first request");
        return ygpu_set_tile(s->psets[pset]->ctx, cpw, twl);
    });
}

int ygpu_server_set_extract_index(ygpu_server* s, int32_t pset, int32_t index) {
    return with_server(s, [&] {
        std::lock_guard<std::mutex> g(s->mu);
        if (pset < 0 || pset >= static_cast<int32_t>(s->psets.size())) return fail(YGPU_ERR_NOT_FOUND, "unknown parameter set");
        if (s->started) return fail(YGPU_ERR_INVALID_ARG, "the extraction index must be set before the first request");
        return ygpu_set_extract_index(s->psets[pset]->ctx, index);
    });
}

}  // extern "C"

namespace {

// Registers a tenant whose host form `make` builds with the parameter set's
// context: pinned copy of the transfer form, arena size, load bytes.
template <class MakeForm>
int add_tenant_form(ygpu_server* s, int32_t pset, int32_t* tenant, MakeForm&& make) {
    ygpu_ctx* ctx = nullptr;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (pset < 0 || pset >= static_cast<int32_t>(s->psets.size())) return fail(YGPU_ERR_NOT_FOUND, "unknown parameter set");
        if (static_cast<int>(s->tenants.size()) >= s->cfg.max_tenants) return fail(YGPU_ERR_INVALID_ARG, "max_tenants reached");
        ctx = s->psets[pset]->ctx;
    }
    hipSetDevice(s->cfg.device);
    HostKeyForm hk;
    int rc = make(ctx, hk);
    if (rc) return rc;
    PinnedKey f;
    f.bsk_bytes = hk.bsk_bytes;
    f.s2_level = hk.s2_level;
    f.packed_bytes = hk.packed.size() * sizeof(uint64_t);
    f.planes_bytes = hk.planes.size() * sizeof(uint32_t);
    if (align256(f.bsk_bytes) + align256(f.planes_bytes) > s->arena_bytes)
        return fail(YGPU_ERR_OUT_OF_MEMORY, "key larger than the key budget");
    YG_HIP(hipHostMalloc(reinterpret_cast<void**>(&f.packed), f.packed_bytes, hipHostMallocDefault));
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&f.planes), f.planes_bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        hipHostFree(f.packed);
        return server_fail_hip(e, "pinned key");
    }
    std::memcpy(f.packed, hk.packed.data(), f.packed_bytes);
    std::memcpy(f.planes, hk.planes.data(), f.planes_bytes);
    Tenant t;
    t.pset = pset;
    t.bytes = align256(f.bsk_bytes) + align256(f.planes_bytes);
    t.load_bytes = f.packed_bytes + f.planes_bytes;
    e = hipEventCreateWithFlags(&t.loaded, hipEventDisableTiming);
    if (e != hipSuccess) {
        hipHostFree(f.packed);
        hipHostFree(f.planes);
        return server_fail_hip(e, "event");
    }
    std::lock_guard<std::mutex> g(s->mu);
    // another registration may have taken the last slot meanwhile
    if (static_cast<int>(s->tenants.size()) >= s->cfg.max_tenants) {
        hipHostFree(f.packed);
        hipHostFree(f.planes);
        hipEventDestroy(t.loaded);
        return fail(YGPU_ERR_INVALID_ARG, "max_tenants reached");
    }
    t.form = static_cast<int>(s->forms.size());
    s->forms.push_back(f);
    *tenant = static_cast<int32_t>(s->tenants.size());
    s->tenants.push_back(t);
    s->st.host_key_bytes += static_cast<int64_t>(f.packed_bytes + f.planes_bytes);
    return YGPU_OK;
}

}  // namespace

extern "C" {

int ygpu_server_add_tenant(ygpu_server* s, int32_t pset, const uint64_t* bsk_coeff, const int32_t* ksk, int32_t* tenant) {
    return with_server(s, [&] {
        if (!bsk_coeff || !ksk || !tenant) return fail(YGPU_ERR_INVALID_ARG, "null argument");
        return add_tenant_form(s, pset, tenant, [&](ygpu_ctx* ctx, HostKeyForm& hk) { return prepare_key_host(ctx, bsk_coeff, ksk, hk); });
    });
}

int ygpu_server_add_tenant_succinct(ygpu_server* s, int32_t pset, const uint64_t* bsk_half, const uint64_t* s2, int32_t s2_level,
                                    const int32_t* ksk, int32_t* tenant) {
    return with_server(s, [&] {
        if (!bsk_half || !s2 || !ksk || !tenant) return fail(YGPU_ERR_INVALID_ARG, "null argument");
        return add_tenant_form(s, pset, tenant, [&](ygpu_ctx* ctx, HostKeyForm& hk) {
            return prepare_key_host_succinct(ctx, bsk_half, s2, s2_level, ksk, hk);
        });
    });
}

int ygpu_server_set_file_slots(ygpu_server* s, int32_t pset, int32_t slots) {
    return with_server(s, [&] {
        std::lock_guard<std::mutex> g(s->mu);
        if (pset < 0 || pset >= static_cast<int32_t>(s->psets.size())) return fail(YGPU_ERR_NOT_FOUND, "unknown parameter set");
        if (slots < 1 || slots > 64) return fail(YGPU_ERR_INVALID_ARG, "slots must be in [1, 64]");
        if (!s->psets[pset]->slots.empty())
            return fail(YGPU_ERR_INVALID_ARG, "slots must be set before the parameter set's first file-backed tenant");
        s->psets[pset]->slot_count = slots;
        return YGPU_OK;
    });
}

int ygpu_server_add_tenant_file(ygpu_server* s, int32_t pset, const char* path, int32_t* tenant) {
    return with_server(s, [&] {
        if (!path || !tenant) return fail(YGPU_ERR_INVALID_ARG, "null argument");
        ygpu_params fp{};
        int rc = ygpu_key_file_info(path, &fp);
        if (rc) return rc;
        int32_t s2_level = 0;
        rc = ygpu_key_file_form(path, &s2_level);
        if (rc) return rc;
        hipSetDevice(s->cfg.device);
        std::lock_guard<std::mutex> g(s->mu);
        if (pset < 0 || pset >= static_cast<int32_t>(s->psets.size())) return fail(YGPU_ERR_NOT_FOUND, "unknown parameter set");
        if (static_cast<int>(s->tenants.size()) >= s->cfg.max_tenants) return fail(YGPU_ERR_INVALID_ARG, "max_tenants reached");
        PSet& ps = *s->psets[pset];
        const ygpu_params& p = ps.p;
        if (!params_equal(fp, p)) return fail(YGPU_ERR_INVALID_ARG, std::string(path) + ": parameters differ from the parameter set's");
        if (s2_level > 0) {
            rc = check_s2_level(p, s2_level);
            if (rc) return rc;
        }
        PinnedKey f;
        f.path = path;
        f.pset = pset;
        f.present = false;
        f.s2_level = s2_level;
        f.bsk_bytes = bsk_polys(p) * ps.ctx->np * p.N * sizeof(double);
        f.packed_bytes = (s2_level ? packed_words_succinct(p, s2_level) : packed_words(p)) * sizeof(uint64_t);
        f.planes_bytes = ksk_plane_words(p) * sizeof(uint32_t);
        if (align256(f.bsk_bytes) + align256(f.planes_bytes) > s->arena_bytes)
            return fail(YGPU_ERR_OUT_OF_MEMORY, "key larger than the key budget");
        if (ps.slots.empty()) {
            // the set's pinned slots, each sized for a full key's transfer form
            std::vector<FileSlot> made(ps.slot_count);
            hipError_t e = hipSuccess;
            for (FileSlot& sl : made) {
                if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.packed), packed_words(p) * sizeof(uint64_t), hipHostMallocDefault);
                if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.planes), f.planes_bytes, hipHostMallocDefault);
            }
            if (e != hipSuccess) {
                for (FileSlot& sl : made) {
                    if (sl.packed) hipHostFree(sl.packed);
                    if (sl.planes) hipHostFree(sl.planes);
                }
                return server_fail_hip(e, "pinned slots for file-backed keys");
            }
            ps.slots = std::move(made);
            s->st.host_key_bytes += static_cast<int64_t>(ps.slots.size() * (packed_words(p) * sizeof(uint64_t) + f.planes_bytes));
        }
        Tenant t;
        t.pset = pset;
        t.bytes = align256(f.bsk_bytes) + align256(f.planes_bytes);
        t.load_bytes = f.packed_bytes + f.planes_bytes;
        YG_HIP(hipEventCreateWithFlags(&t.loaded, hipEventDisableTiming));
        t.form = static_cast<int>(s->forms.size());
        s->forms.push_back(f);
        *tenant = static_cast<int32_t>(s->tenants.size());
        s->tenants.push_back(t);
        return YGPU_OK;
    });
}

int ygpu_server_add_tenant_shared(ygpu_server* s, int32_t src_tenant, int32_t* tenant) {
    return with_server(s, [&] {
        if (!tenant) return fail(YGPU_ERR_INVALID_ARG, "null argument");
        std::lock_guard<std::mutex> g(s->mu);
        if (src_tenant < 0 || src_tenant >= static_cast<int32_t>(s->tenants.size())) return fail(YGPU_ERR_NOT_FOUND, "unknown tenant");
        if (static_cast<int>(s->tenants.size()) >= s->cfg.max_tenants) return fail(YGPU_ERR_INVALID_ARG, "max_tenants reached");
        Tenant t;
        t.pset = s->tenants[src_tenant].pset;
        t.form = s->tenants[src_tenant].form;
        t.bytes = s->tenants[src_tenant].bytes;
        t.load_bytes = s->tenants[src_tenant].load_bytes;
        YG_HIP(hipSetDevice(s->cfg.device));
        YG_HIP(hipEventCreateWithFlags(&t.loaded, hipEventDisableTiming));
        *tenant = static_cast<int32_t>(s->tenants.size());
        s->tenants.push_back(t);
        return YGPU_OK;
    });
}

int ygpu_server_prefetch(ygpu_server* s, int32_t tenant) {
    return with_server(s, [&] {
        std::unique_lock<std::mutex> lk(s->mu);
        if (tenant < 0 || tenant >= static_cast<int32_t>(s->tenants.size())) return fail(YGPU_ERR_NOT_FOUND, "unknown tenant");
        if (!s->started) {
            const int rc = start_threads(s);
            if (rc) return rc;
        }
        Tenant& tn = s->tenants[tenant];
        while (!tn.resident) {
            const int rc = start_load(s, *s->psets[tn.pset], tenant);
            if (rc == 1) return fail(YGPU_ERR_BUSY, "no evictable key memory for a prefetch");
            if (rc == 3) return fail(YGPU_ERR_BUSY, "every pinned slot for file-backed keys is in use");
            if (rc < 0) return rc;
            if (rc == 2) {
                // a file-backed form: wait for the reader, then load
                const int fi = tn.form;
                s->cv_req.wait(lk, [&] { return !s->forms[fi].reading || s->stop; });
                if (s->stop) return fail(YGPU_ERR_INVALID_ARG, "server is shutting down");
                if (!s->forms[fi].present) return fail(s->forms[fi].read_rc ? s->forms[fi].read_rc : YGPU_ERR_CORRUPT, s->forms[fi].read_err);
                continue;
            }
            tn.fresh = false;  // no request is waiting for it: an ordinary LRU entry
        }
        s->cv_req.wait(lk, [&] { return tn.ready || !tn.resident; });
        return tn.resident ? YGPU_OK : fail(YGPU_ERR_HIP, "prefetch load failed");
    });
}

int ygpu_server_tenant_params(ygpu_server* s, int32_t tenant, int32_t* pset) {
    return with_server(s, [&] {
        if (!pset) return fail(YGPU_ERR_INVALID_ARG, "null argument");
        std::lock_guard<std::mutex> g(s->mu);
        if (tenant < 0 || tenant >= static_cast<int32_t>(s->tenants.size())) return fail(YGPU_ERR_NOT_FOUND, "unknown tenant");
        *pset = s->tenants[tenant].pset;
        return YGPU_OK;
    });
}

int ygpu_server_submit(ygpu_server* s, ygpu_request* req) {
    return with_server(s, [&] {
        if (!req || !req->lwe_in || !req->v || !req->lwe_out) return fail(YGPU_ERR_INVALID_ARG, "null request buffer");
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->stop) return fail(YGPU_ERR_INVALID_ARG, "server is shutting down");
        if (req->tenant < 0 || req->tenant >= static_cast<int32_t>(s->tenants.size())) return fail(YGPU_ERR_NOT_FOUND, "unknown tenant");
        Tenant& tn = s->tenants[req->tenant];
        PSet& ps = *s->psets[tn.pset];
        s->st.submitted++;
        if (ps.queued >= s->cfg.queue_max) {
            s->st.shed++;
            req->status = YGPU_ERR_BUSY;
            return fail(YGPU_ERR_BUSY, "queue full: request shed");
        }
        if (!s->started) {
            const int rc = start_threads(s);
            if (rc) return rc;
        }
        req->status = 1;
        if (req->arrival_ns == 0) req->arrival_ns = now_ns();
        req->seq = ++s->tick;
        req->key_hit = tn.ready ? 1 : 0;
        // absent key already awaited (loading, or an earlier request queued for it)
        req->coalesced = (!tn.ready && (tn.resident || !tn.pending.empty())) ? 1 : 0;
        req->lane = -1;
        s->st.coalesced += req->coalesced;
        {
            const int64_t now = now_ns();
            tn.freq = decayed_freq(tn, now) + 1.0;
            tn.freq_ns = now;
        }
        s->st.request_hits += req->key_hit;
        const int ql = queue_lane(s, ps, req->tenant);
        tn.pending.push_back(req);
        tn.pending_lane.push_back(ql);
        ps.queued++;
        if (ql >= 0) ps.lane_depth[ql]++;
        s->st.peak_queue_depth = std::max<int64_t>(s->st.peak_queue_depth, ql >= 0 ? ps.lane_depth[ql] : ps.queued);
        if (!tn.listed) {
            tn.listed = true;
            ps.order.push_back(req->tenant);
        }
        ps.cv_work.notify_one();
        return YGPU_OK;
    });
}

int ygpu_server_wait(ygpu_server* s, ygpu_request* req) {
    return with_server(s, [&] {
        if (!req) return fail(YGPU_ERR_INVALID_ARG, "null request");
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_req.wait(lk, [&] { return req->status != 1; });
        return static_cast<int>(req->status);
    });
}

int ygpu_server_wait_for(ygpu_server* s, ygpu_request* req, int32_t timeout_ms) {
    return with_server(s, [&] {
        if (!req) return fail(YGPU_ERR_INVALID_ARG, "null request");
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_req.wait_for(lk, std::chrono::milliseconds(std::max(0, timeout_ms)), [&] { return req->status != 1; });
        return static_cast<int>(req->status);
    });
}

int ygpu_server_stats_get(ygpu_server* s, ygpu_server_stats* out) {
    return with_server(s, [&] {
        if (!out) return fail(YGPU_ERR_INVALID_ARG, "null argument");
        std::lock_guard<std::mutex> g(s->mu);
        *out = s->st;
        return YGPU_OK;
    });
}

int ygpu_server_destroy(ygpu_server* s) {
    if (!s) return YGPU_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        for (auto& ps : s->psets) ps->cv_work.notify_all();
        s->cv_read.notify_all();
        s->cv_req.notify_all();
    }
    if (s->reader.joinable()) s->reader.join();
    for (auto& ps : s->psets) {
        if (ps->dispatcher.joinable()) ps->dispatcher.join();
        if (ps->completer.joinable()) ps->completer.join();
        if (ps->watcher.joinable()) ps->watcher.join();
    }
    hipSetDevice(s->cfg.device);
    {
        // requests still queued never ran
        std::lock_guard<std::mutex> g(s->mu);
        const int64_t t = now_ns();
        for (auto& ps : s->psets) {
            for (int tn : ps->order) {
                for (ygpu_request* r : s->tenants[tn].pending) finish_request(r, YGPU_ERR_INVALID_ARG, t, t, 0);
                s->tenants[tn].pending.clear();
                s->tenants[tn].pending_lane.clear();
            }
            ps->order.clear();
        }
        s->cv_req.notify_all();
    }
    if (s->copy_stream) hipStreamSynchronize(s->copy_stream);
    for (auto& ps : s->psets) {
        if (ps->d_stage) hipFree(ps->d_stage);
        for (Lane& L : ps->lanes) lane_free(L);
        ygpu_destroy(ps->ctx);
    }
    if (s->copy_stream) hipStreamDestroy(s->copy_stream);
    for (Tenant& t : s->tenants)
        if (t.loaded) hipEventDestroy(t.loaded);
    for (PinnedKey& f : s->forms) {
        if (!f.path.empty()) continue;  // file-backed: its buffers are its slot's
        if (f.packed) hipHostFree(f.packed);
        if (f.planes) hipHostFree(f.planes);
    }
    for (auto& ps : s->psets)
        for (FileSlot& sl : ps->slots) {
            if (sl.packed) hipHostFree(sl.packed);
            if (sl.planes) hipHostFree(sl.planes);
        }
    if (s->d_arena) hipFree(s->d_arena);
    if (s->d_bsk_tab) hipFree(s->d_bsk_tab);
    if (s->d_ksk_tab) hipFree(s->d_ksk_tab);
    if (s->h_bsk_tab) hipHostFree(s->h_bsk_tab);
    if (s->h_ksk_tab) hipHostFree(s->h_ksk_tab);
    delete s;
    return YGPU_OK;
}

// Open-loop load: Poisson arrivals (exponential gaps, std::mt19937_64 seeded
// with seed ^ 0x9e3779b97f4a7c15, as ArrivalProcess,
// ya_benchmark/include/blindrotate_serving.h:300-350) over tenants drawn from a
// Zipf(s) popularity (cumulative table + lower_bound, yatfhe/include/yautil/zipf.h:47-87;
// popularity rank r is tenant r), latency from each request's scheduled
// arrival to its completion, percentiles by the nearest-rank rule of
// ya_benchmark/blindrotate_serving.cpp:70-73.
int ygpu_server_loadgen(ygpu_server* s, const ygpu_loadgen_config* c, ygpu_loadgen_result* out) {
    return with_server(s, [&] {
        if (!c || !out || !c->lwe_in || !c->v) return fail(YGPU_ERR_INVALID_ARG, "null argument");
        if (c->rate <= 0 || c->requests < 0 || c->warm < 0 || c->n_tenants <= 0) return fail(YGPU_ERR_INVALID_ARG, "bad load parameters");
        std::vector<int> cols(c->n_tenants), psetOf(c->n_tenants);
        {
            std::lock_guard<std::mutex> g(s->mu);
            if (c->n_tenants > static_cast<int32_t>(s->tenants.size())) return fail(YGPU_ERR_INVALID_ARG, "more tenants than registered");
            for (int t = 0; t < c->n_tenants; t++) {
                psetOf[t] = s->tenants[t].pset;
                cols[t] = s->psets[psetOf[t]]->p.n + 1;
            }
        }
        const int maxCols = *std::max_element(cols.begin(), cols.end());
        // popularity
        std::vector<double> cdf(c->n_tenants);
        double h = 0;
        for (int k = 1; k <= c->n_tenants; k++) h += 1.0 / std::pow(static_cast<double>(k), c->zipf_s);
        double acc = 0;
        for (int r = 0; r < c->n_tenants; r++) {
            acc += 1.0 / (std::pow(static_cast<double>(r + 1), c->zipf_s) * h);
            cdf[r] = acc;
        }
        cdf.back() = 1.0;
        std::mt19937_64 zgen(c->seed);
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        const int64_t total = c->warm + c->requests;
        std::vector<int> pattern(total);
        for (auto& t : pattern) {
            const double u = uni(zgen);
            t = std::min<int>(static_cast<int>(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin()), c->n_tenants - 1);
        }
        // request records and outputs are owned by shared state that outlives
        // this call if requests are still pending at the deadline
        auto reqsHold = std::make_shared<std::vector<ygpu_request>>(total);
        auto outsHold = std::make_shared<std::vector<int32_t>>(static_cast<size_t>(total) * maxCols);
        std::vector<ygpu_request>& reqs = *reqsHold;
        std::vector<int32_t>& outs = *outsHold;
        auto make = [&](int64_t i) {
            ygpu_request& r = reqs[i];
            std::memset(&r, 0, sizeof(r));
            r.tenant = pattern[i];
            r.lwe_in = c->lwe_in[r.tenant];
            r.v = c->v[r.tenant];
            r.lwe_out = outs.data() + static_cast<size_t>(i) * maxCols;
        };
        // warm-up: one burst, not measured
        for (int64_t i = 0; i < c->warm; i++) {
            make(i);
            int rc = ygpu_server_submit(s, &reqs[i]);
            if (rc && rc != YGPU_ERR_BUSY) return rc;
        }
        for (int64_t i = 0; i < c->warm; i++)
            if (reqs[i].status != YGPU_ERR_BUSY && ygpu_server_wait_for(s, &reqs[i], 120000) == 1) {
                static std::mutex graveMu;
                static std::vector<std::shared_ptr<void>> grave;
                std::lock_guard<std::mutex> g(graveMu);
                grave.push_back(reqsHold);
                grave.push_back(outsHold);
                return fail(YGPU_ERR_BUSY, "load generator: warm-up requests still pending after 120 s");
            }
        int64_t loadsFrom = 0, loadsTo = 0;  // the window of the loads-in-flight integral
        ygpu_server_stats s0{};
        {
            // the queue-depth and load peaks are per measured run
            std::lock_guard<std::mutex> g(s->mu);
            s->st.peak_queue_depth = 0;
            s->st.peak_loads_in_flight = s->st.loads_in_flight;
            load_count(s, 0);
            loadsFrom = s->load_change_ns;
            s0 = s->st;
        }
        // measured open loop
        std::mt19937_64 agen(c->seed ^ 0x9e3779b97f4a7c15ULL);
        std::exponential_distribution<double> gap(1.0);
        const int64_t t0 = now_ns() + 1000000;  // start 1 ms from now
        double cursor_ms = 0;
        int64_t shed = 0;
        std::vector<char> submitted(total, 0);
        for (int64_t i = c->warm; i < total; i++) {
            cursor_ms += gap(agen) * 1000.0 / c->rate;
            const int64_t due = t0 + static_cast<int64_t>(cursor_ms * 1e6);
            for (;;) {
                const int64_t now = now_ns();
                if (now >= due) break;
                if (due - now > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(due - now - 100000));
                else std::this_thread::yield();
            }
            make(i);
            reqs[i].arrival_ns = due;
            const int rc = ygpu_server_submit(s, &reqs[i]);
            if (rc == YGPU_ERR_BUSY) shed++;
            else if (rc) return rc;
            else submitted[i] = 1;
        }
        int64_t errors = 0, mismatches = 0, done = 0, coalesced = 0, owned = 0;
        int64_t t_end = t0;
        std::vector<double> lat, queue, service, latHit, latMiss, latPset[4];
        lat.reserve(c->requests);
        queue.reserve(c->requests);
        service.reserve(c->requests);
        std::unordered_map<int, std::pair<double, int64_t>> perTenant;
        int64_t laneCount[16] = {};
        std::vector<int> laneBase(s->psets.size() + 1, 0);  // first global lane of each set
        for (size_t k = 0; k < s->psets.size(); k++)
            laneBase[k + 1] = laneBase[k] + static_cast<int>(s->psets[k]->lanes.size());
        const bool affine = s->cfg.sched != YGPU_SCHED_SHARED;
        // a run that has not drained 120 s after its last arrival is stuck
        const int64_t deadline = now_ns() + int64_t{120} * 1000000000;
        for (int64_t i = c->warm; i < total; i++) {
            if (!submitted[i]) continue;
            const int64_t left_ms = (deadline - now_ns()) / 1000000;
            const int rc = ygpu_server_wait_for(s, &reqs[i], static_cast<int32_t>(std::max<int64_t>(0, left_ms)));
            if (rc == 1) {
                int64_t pending = 0;
                for (int64_t j = i; j < total; j++) pending += submitted[j] && reqs[j].status == 1;
                // the server still references these buffers: keep them alive
                static std::mutex graveMu;
                static std::vector<std::shared_ptr<void>> grave;
                {
                    std::lock_guard<std::mutex> g(graveMu);
                    grave.push_back(reqsHold);
                    grave.push_back(outsHold);
                }
                return fail(YGPU_ERR_BUSY, "load generator: " + std::to_string(pending) + " requests still pending 120 s after the last arrival");
            }
            if (rc) {
                errors++;
                continue;
            }
            done++;
            t_end = std::max<int64_t>(t_end, reqs[i].done_ns);
            lat.push_back((reqs[i].done_ns - reqs[i].arrival_ns) * 1e-6);
            (reqs[i].key_hit ? latHit : latMiss).push_back(lat.back());
            if (psetOf[reqs[i].tenant] < 4) latPset[psetOf[reqs[i].tenant]].push_back(lat.back());
            queue.push_back(std::max<int64_t>(0, reqs[i].launch_ns - reqs[i].arrival_ns) * 1e-6);
            service.push_back(std::max<int64_t>(0, reqs[i].done_ns - reqs[i].launch_ns) * 1e-6);
            auto& pt = perTenant[reqs[i].tenant];
            pt.first += lat.back();
            pt.second++;
            coalesced += reqs[i].coalesced;
            const int gl = laneBase[psetOf[reqs[i].tenant]] + reqs[i].lane;
            if (reqs[i].lane >= 0 && gl < 16) laneCount[gl]++;
            const int nl = laneBase[psetOf[reqs[i].tenant] + 1] - laneBase[psetOf[reqs[i].tenant]];
            owned += affine && nl > 0 && reqs[i].lane == reqs[i].tenant % nl;
            if (c->expected && c->expected[reqs[i].tenant] &&
                std::memcmp(reqs[i].lwe_out, c->expected[reqs[i].tenant], sizeof(int32_t) * cols[reqs[i].tenant]) != 0)
                mismatches++;
        }
        ygpu_server_stats s1{};
        {
            std::lock_guard<std::mutex> g(s->mu);
            load_count(s, 0);
            loadsTo = s->load_change_ns;
            s1 = s->st;
        }
        auto pct = [](std::vector<double>& v, double p) {
            if (v.empty()) return 0.0;
            const long i = static_cast<long>(std::ceil(p / 100.0 * v.size())) - 1;
            return v[std::min<long>(std::max<long>(i, 0), static_cast<long>(v.size()) - 1)];
        };
        std::sort(lat.begin(), lat.end());
        std::sort(queue.begin(), queue.end());
        std::sort(latHit.begin(), latHit.end());
        std::sort(latMiss.begin(), latMiss.end());
        std::memset(out, 0, sizeof(*out));
        const double wall = (t_end - t0) * 1e-9;
        out->offered_rps = c->rate;
        out->wall_sec = wall;
        out->throughput_rps = wall > 0 ? done / wall : 0;
        out->completed = done;
        out->shed = shed;
        out->errors = errors;
        out->mismatches = mismatches;
        double sum = 0;
        for (double x : lat) sum += x;
        out->mean_ms = lat.empty() ? 0 : sum / lat.size();
        out->p50_ms = pct(lat, 50);
        out->p90_ms = pct(lat, 90);
        out->p99_ms = pct(lat, 99);
        out->p99_9_ms = pct(lat, 99.9);
        out->max_ms = lat.empty() ? 0 : lat.back();
        out->queue_p50_ms = pct(queue, 50);
        out->queue_p99_ms = pct(queue, 99);
        int64_t over = 0;
        for (double x : lat) over += x > c->slo_ms;
        out->slo_violation_rate = lat.empty() ? 0 : static_cast<double>(over) / lat.size();
        const int64_t nb = s1.batches - s0.batches;
        out->batches = nb;
        out->mean_batch = nb ? static_cast<double>(s1.batched_cts - s0.batched_cts) / nb : 0;
        const int64_t ns = s1.submitted - s0.submitted;
        out->hit_rate = ns ? static_cast<double>(s1.request_hits - s0.request_hits) / ns : 0;
        out->key_loads = s1.key_loads - s0.key_loads;
        out->key_evictions = s1.key_evictions - s0.key_evictions;
        out->bytes_loaded = s1.bytes_loaded - s0.bytes_loaded;
        out->memory_waits = s1.memory_waits - s0.memory_waits;
        out->hit_p50_ms = pct(latHit, 50);
        out->hit_p99_ms = pct(latHit, 99);
        out->miss_p50_ms = pct(latMiss, 50);
        out->miss_p99_ms = pct(latMiss, 99);
        for (int k = 0; k < 4; k++) {
            std::sort(latPset[k].begin(), latPset[k].end());
            out->pset_p50_ms[k] = pct(latPset[k], 50);
            out->pset_p99_ms[k] = pct(latPset[k], 99);
        }
        // the rest of the reference's record (blindrotate_serving.cpp:60-80, 253-353)
        double sq = 0;
        for (double x : lat) sq += (x - out->mean_ms) * (x - out->mean_ms);
        out->sd_ms = lat.empty() ? 0 : std::sqrt(sq / lat.size());
        double qsum = 0;
        for (double x : queue) qsum += x;
        out->queue_mean_ms = queue.empty() ? 0 : qsum / queue.size();
        out->queue_p90_ms = pct(queue, 90);
        out->queue_max_ms = queue.empty() ? 0 : queue.back();
        std::sort(service.begin(), service.end());
        double ssum = 0;
        for (double x : service) ssum += x;
        out->service_mean_ms = service.empty() ? 0 : ssum / service.size();
        out->service_p50_ms = pct(service, 50);
        out->service_p90_ms = pct(service, 90);
        out->service_p99_ms = pct(service, 99);
        out->service_max_ms = service.empty() ? 0 : service.back();
        int64_t o100 = 0, o300 = 0, o500 = 0;
        for (double x : lat) {
            o100 += x > 100.0;
            o300 += x > 300.0;
            o500 += x > 500.0;
        }
        out->slo_violation_rate_100 = lat.empty() ? 0 : static_cast<double>(o100) / lat.size();
        out->slo_violation_rate_300 = lat.empty() ? 0 : static_cast<double>(o300) / lat.size();
        out->slo_violation_rate_500 = lat.empty() ? 0 : static_cast<double>(o500) / lat.size();
        out->coalesced_loads = coalesced;
        out->peak_queue_depth = s1.peak_queue_depth;
        out->spilled = s1.spilled - s0.spilled;
        out->stolen = s1.stolen - s0.stolen;
        double fair = 1.0;
        if (perTenant.size() > 1) {
            double sm = 0, sm2 = 0;
            for (const auto& kv : perTenant) {
                const double m = kv.second.first / kv.second.second;
                sm += m;
                sm2 += m * m;
            }
            fair = sm * sm / (perTenant.size() * sm2);
        }
        out->tenant_fairness_jain = fair;
        const int64_t span = loadsTo - loadsFrom;
        out->load_concurrency_mean = span > 0 ? static_cast<double>(s1.load_inflight_ns - s0.load_inflight_ns) / span : 0;
        out->load_concurrency_peak = s1.peak_loads_in_flight;
        out->locality = done ? static_cast<double>(owned) / done : 0;
        for (int k = 0; k < 16; k++) out->lane_requests[k] = laneCount[k];
        return YGPU_OK;
    });
}

}  // extern "C"
