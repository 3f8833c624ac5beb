/* yatfhe_gpu.h -- C ABI of the MI355X TFHE programmable-bootstrapping engine.
 *
 * This is the boundary a caller of the reference library's bootstrapping path
 * binds to.  Each entry point names the reference interface it replaces
 * (paths relative to the reference repository root, xsong93/yatfhe).
 *
 * Conventions
 *  - Every function returns an int status (YGPU_OK == 0, negative on error);
 *    nothing throws across this boundary.  ygpu_last_error() returns the
 *    calling thread's last message.
 *  - Torus values (test polynomial v, accumulator) use the reference's Torus
 *    type: int32_t when params.torus_bits == 32 (TORUS32), int64_t holding the
 *    sign-extended 56-bit value when torus_bits == 56 (TORUS56).  They are passed
 *    as void* and sized by torus_bits.
 *  - LWE ciphertexts are int32_t[n+1] (a[0..n-1], b): every supported
 *    configuration has qLwe = 2^32, so LweTorus values fit 32 bits; callers
 *    built with LweTorus = int64_t narrow them losslessly.
 *  - Buffers whose name starts with d_ are device pointers (from ygpu_malloc);
 *    all others are host pointers owned by the caller.  The library owns every
 *    device buffer it allocates itself (keys, scratch).
 *  - Bootstrapping keys cross this boundary in coefficient form: for every
 *    NTT-form key polynomial of the reference, its coefficient-domain residues
 *    mod qNtt in [0, qNtt) (the inverse transform without centring).  The
 *    engine's own transform order never crosses the interface, so keys produced
 *    by any correct NTT (the reference uses Intel HEXL) are interchangeable.
 *    Layout: bsk[i][lvl][row][col][N], i < n, lvl < level, row < k+1 (row r < k
 *    is the RGSW row of the mask digit a_r, row k the row of the body digit b),
 *    col < k+1 (col c < k is the mask polynomial a_c, col k the body b).
 *  - KSK layout: ksk[i][j][n+1] int32, i < k*N, j < ks_level, entries
 *    a[0..n-1] then b (TlweKeySwitchingKey::decomposedKsk, keyswitching.h:14-31).
 *  - One context per GPU.  The synchronous host-buffer calls (ygpu_pbs,
 *    ygpu_blind_rotate[_scaled], ygpu_sample_extract[_index], ygpu_keyswitch,
 *    ygpu_extract_keyswitch) may be made from several threads at once: each
 *    call runs on a stream and buffers of its own, so bootstraps of different
 *    threads run side by side (the reference runs blindRotateJP22Ntt on many
 *    worker threads, blindrotate_serving.cpp:197-228).  Every other call on a
 *    context is serialised by the library and runs on the context's stream.
 *  - Device-side failures of asynchronous calls (YGPU_ERR_DEVICE) are reported
 *    by the next ygpu_sync, ygpu_memcpy_dtoh or synchronous bootstrapping call
 *    on the same context.
 */
#ifndef YATFHE_GPU_H
#define YATFHE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YGPU_OK 0
#define YGPU_ERR_INVALID_ARG (-1)
#define YGPU_ERR_UNSUPPORTED (-2)
#define YGPU_ERR_HIP (-3)
#define YGPU_ERR_NO_DEVICE (-4)
#define YGPU_ERR_OUT_OF_MEMORY (-5)
#define YGPU_ERR_NOT_FOUND (-6)
#define YGPU_ERR_BUSY (-7)      /* serving queue full: the request was shed */
#define YGPU_ERR_IO (-8)        /* a file could not be opened, read or written */
#define YGPU_ERR_CORRUPT (-9)   /* a key file failed its format or digest checks */
#define YGPU_ERR_DEVICE (-10)   /* a kernel reported a failed device-side check (a blind-rotation team
                                   barrier timed out): the results of the calls since the last
                                   synchronisation of that context are invalid */

/* Parameter set.  Mirrors YatfheParameters (yatfhe_parameters.h:9-110) for the
 * fields the bootstrapping path reads; `level` is the gadget length used by the
 * external product (lApprox == l for the key formats supported). */
typedef struct ygpu_params {
    int32_t torus_bits;     /* 32 or 56 (torus.h:11-21) */
    int32_t n;              /* LWE dimension */
    int32_t k;              /* GLWE dimension (1) */
    int32_t N;              /* ring degree (1024 or 2048) */
    int32_t radix_bits;     /* gadget base log2, 1..15 */
    int32_t level;          /* gadget levels, 1..4; level * radix_bits <= torus_bits */
    int32_t ks_radix_bits;  /* keyswitch base log2, 1..7 (digits travel as signed bytes) */
    int32_t ks_level;       /* keyswitch levels */
    int32_t ks_width_bits;  /* keyswitch precision (32) */
    int32_t lwe_q_bits;     /* log2 qLwe (32) */
    uint64_t q_ntt;         /* NTT prime: Q_49P_T32 or Q_61P_T56 (torus.h:151,157) */
} ygpu_params;

typedef struct ygpu_ctx ygpu_ctx;

/* Multi-key job: `count` ciphertexts starting at row `first` of the batch
 * buffers, all bootstrapped under key `key_id`. */
typedef struct ygpu_job {
    int32_t key_id;
    int32_t count;
    int64_t first;
} ygpu_job;

const char* ygpu_last_error(void);
int ygpu_device_count(int32_t* count);
/* Sizes in bytes of the ABI structs, in the order ygpu_params, ygpu_job,
 * ygpu_server_config, ygpu_request, ygpu_server_stats, ygpu_loadgen_config,
 * ygpu_loadgen_result; writes min(max, 7) of them and returns how many exist.
 * No counterpart in the reference (its boundary is C++): it lets a binding
 * check its struct mirrors against the library it loaded. */
int ygpu_abi_struct_sizes(int64_t* out, int32_t max);

/* Replaces initYatfhe (initializer.cpp:50-85): validates the parameter set,
 * builds the transform tables and binds the context to `device`.  Unlike the
 * reference's process-global state, contexts of different parameter sets
 * coexist in one process. */
int ygpu_init(const ygpu_params* params, int32_t device, ygpu_ctx** out);
int ygpu_destroy(ygpu_ctx* ctx);
int ygpu_get_params(const ygpu_ctx* ctx, ygpu_params* out);
/* The coefficient the bootstrap's sample extraction takes (YatfheParameters::
 * driftPhase, yatfhe_parameters.h:65, passed by functionalBootstrappingMPNtt
 * to extractTlweFromTrlwe, bootstrapping.cpp:78): 0 by default, in [0, N).
 * Applies to every later bootstrap of the context, batched and plain, and to
 * graphs captured afterwards. */
int ygpu_set_extract_index(ygpu_ctx* ctx, int32_t index);
int ygpu_get_extract_index(ygpu_ctx* ctx, int32_t* index);

/* Makes a bootstrapping key (BootstrappingKeyMP / BootstrappingKey,
 * bootstrapping.h:16-114) and its keyswitching key (keyswitching.h:14-31)
 * resident on the device, the BSK transformed into the engine's NTT form. */
int ygpu_key_upload(ygpu_ctx* ctx, const uint64_t* bsk_coeff, const int32_t* ksk, int32_t* key_id);
/* WWL+24 succinct bootstrapping key (BootstrappingKeyWWL24,
 * bootstrapping.h:116-136): per step only the body-digit rows cPrime[lvl] of
 * the RGSW, plus one TRLEV of s^2 shared by all steps (BootstrappingKeyWWL24::
 * s2Dft; symEncTrlevWithKeyNtt, trlev.cpp:75-94), both in coefficient form:
 * bsk_half[i][lvl][col][N] (i < n, lvl < level, col < k+1) and s2[l2][col][N]
 * (l2 < s2_level, the reference's param.l; s2_level * radix_bits must fit the
 * torus).  The device derives the mask-digit rows c[lvl][k] once, exactly as
 * switchTrlweToSecretEmbeddingNtt (trgsw.cpp:861-909) forms them at every step
 * of blindRotateWWL24Ntt (blind_rotate.cpp:1383-1412), so the key id bootstraps
 * with the results of blindRotateWWL24Ntt through every entry point below.  The
 * transfer is about half the full key's. */
int ygpu_key_upload_succinct(ygpu_ctx* ctx, const uint64_t* bsk_half, const uint64_t* s2, int32_t s2_level, const int32_t* ksk,
                             int32_t* key_id);
int ygpu_key_release(ygpu_ctx* ctx, int32_t key_id);
/* Device bytes held by a key (BSK in NTT form + KSK). */
int ygpu_key_bytes(const ygpu_ctx* ctx, int32_t key_id, int64_t* bytes);

/* Plain drop-in form, one ciphertext per call, host buffers.
 * Replaces functionalBootstrappingMPNtt (bootstrapping.cpp:71-80) and, for
 * keys with level == lApprox, functionalBootstrappingNtt (bootstrapping.cpp:60-69).
 * lwe_in, lwe_out: int32[n+1].  v: Torus[N].  Optional (may be NULL):
 * acc_out Torus[(k+1)N] (the blind-rotated accumulator, blindRotateMP21Ntt,
 * blind_rotate.cpp:1538-1565), extract_out int32[kN+1] (extractTlweFromTrlwe,
 * trlwe.cpp:241-250). */
int ygpu_pbs(ygpu_ctx* ctx, int32_t key_id, const int32_t* lwe_in, const void* v,
             int32_t* lwe_out, void* acc_out, int32_t* extract_out);

/* Blind rotation alone, host buffers: rescaleTlweToNewMod (tlwe.cpp:75-81) +
 * genNoiselessTrlweSample (trlwe.cpp:175-182) + blindRotateMP21Ntt
 * (blind_rotate.cpp:1538-1565).  acc_out: Torus[(k+1)N]. */
int ygpu_blind_rotate(ygpu_ctx* ctx, int32_t key_id, const int32_t* lwe_in, const void* v, void* acc_out);

/* Blind rotation of an already rescaled input from a caller-supplied
 * accumulator, host buffers: exactly blindRotateMP21Ntt / blindRotateJP22Ntt
 * (blind_rotate.cpp:1448-1478, 1538-1565) with ScaledTlwe::a (abar, n values in
 * [-2N, 2N)) and the in/out Trlwe accum (Torus[(k+1)N]). */
int ygpu_blind_rotate_scaled(ygpu_ctx* ctx, int32_t key_id, const int32_t* abar, void* acc);

/* Sample extraction alone, host buffers: extractTlweFromTrlwe(out, in, index)
 * (trlwe.cpp:241-250): a'_{iN+j} = a_i[index - j] for j <= index, else
 * -a_i[N + index - j]; b' = b[index]; each modswitched to 2^32.
 * acc Torus[(k+1)N] -> extract_out int32[kN+1].  ygpu_sample_extract takes the
 * context's extraction index. */
int ygpu_sample_extract_index(ygpu_ctx* ctx, const void* acc, int32_t index, int32_t* extract_out);
int ygpu_sample_extract(ygpu_ctx* ctx, const void* acc, int32_t* extract_out);

/* LWE keyswitch of an arbitrary ciphertext of dimension kN, host buffers:
 * switchKeyForTlwe (keyswitching.cpp:48-72).  lwe_in int32[kN+1] -> lwe_out int32[n+1]. */
int ygpu_keyswitch(ygpu_ctx* ctx, int32_t key_id, const int32_t* lwe_in, int32_t* lwe_out);

/* Sample extract + keyswitch of an accumulator, host buffers:
 * extractTlweFromTrlwe (trlwe.cpp:241-250, the context's index) then switchKeyForTlwe
 * (keyswitching.cpp:48-72).  acc: Torus[(k+1)N]; extract_out optional. */
int ygpu_extract_keyswitch(ygpu_ctx* ctx, int32_t key_id, const void* acc, int32_t* extract_out, int32_t* lwe_out);

/* Batched form on device buffers: `count` ciphertexts under one key in one
 * persistent blind-rotation launch, then extract + keyswitch.
 * d_lwe_in, d_lwe_out: int32[count][n+1].  d_v: Torus[N] shared by all, or
 * Torus[count][N] when v_per_ct != 0.  Optional: d_acc_out Torus[count][(k+1)N],
 * d_extract_out int32[count][kN+1].  Asynchronous on the context's stream. */
int ygpu_pbs_batch_device(ygpu_ctx* ctx, int32_t key_id, int32_t count, const int32_t* d_lwe_in,
                          const void* d_v, int32_t v_per_ct, int32_t* d_lwe_out,
                          void* d_acc_out, int32_t* d_extract_out);

/* ygpu_pbs_batch_device without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * accumulators (when d_acc_out is NULL) and the keyswitch workspace live in a
 * static device array of 96 MiB that the library holds on every device it
 * runs on until the process ends, so the call makes no hipMalloc, hipFree or
 * other allocating call.  A batch whose scratch exceeds the array (about
 * 5,300 TORUS32 N=1024 or 1,900 TORUS56 N=2048 ciphertexts) is rejected with
 * YGPU_ERR_INVALID_ARG naming the limit.  Every call on one device shares that
 * device's array: calls that overlap on one GPU (other contexts, other
 * streams) race and must be ordered by the caller.  Arguments, results and
 * asynchrony as ygpu_pbs_batch_device; no kernel timing is recorded. */
int ygpu_pbs_batch_device_noalloc(ygpu_ctx* ctx, int32_t key_id, int32_t count, const int32_t* d_lwe_in,
                                  const void* d_v, int32_t v_per_ct, int32_t* d_lwe_out,
                                  void* d_acc_out, int32_t* d_extract_out);

/* Multi-key batched form: jobs grouped by key in one persistent launch;
 * buffers as in ygpu_pbs_batch_device with `total` rows. */
int ygpu_pbs_jobs_device(ygpu_ctx* ctx, const ygpu_job* jobs, int32_t n_jobs, int64_t total,
                         const int32_t* d_lwe_in, const void* d_v, int32_t v_per_ct,
                         int32_t* d_lwe_out, void* d_acc_out);

/* Plain form on device buffers (one ciphertext, the per-call persistent
 * kernel), for timing the drop-in path without host copies. */
int ygpu_pbs_plain_device(ygpu_ctx* ctx, int32_t key_id, const int32_t* d_lwe_in, const void* d_v,
                          int32_t* d_lwe_out, void* d_acc_out);

/* Batched-form tile: ciphertexts per workgroup (cpw) and whether the transform
 * twiddles live in LDS (twl = 1) or are read through L1 (twl = 0); twl = 2
 * (TORUS32 N=1024 2^7 x 3, with tab = 2) keeps lane-major LDS copies of the
 * twiddles of transform passes 1 and 2, read without bank conflicts.
 * ygpu_autotune times every tile instantiated for the context's parameter set
 * on a synthetic batch of `count` ciphertexts under `key_id` and keeps the
 * fastest (no counterpart in the reference, which has no tiles). */
int ygpu_get_tile(ygpu_ctx* ctx, int32_t* cpw, int32_t* twl);
int ygpu_set_tile(ygpu_ctx* ctx, int32_t cpw, int32_t twl);
int ygpu_autotune(ygpu_ctx* ctx, int32_t key_id, int32_t count, int32_t* cpw, int32_t* twl, double* best_ms);
/* The tile's third dimension where the gadget digits are signed bytes
 * (TORUS32 N=1024 2^7 x 3): tab forward-transform stages (0, 1 or 2) taken
 * from per-workgroup LDS tables of w * d mod p for every digit d instead of
 * computed products, trading VALU work for LDS reads.  tab = -1 in
 * ygpu_set_tile_ex (and ygpu_set_tile) keeps the current tab where that
 * combination is instantiated, else takes the first. */
int ygpu_get_tile_ex(ygpu_ctx* ctx, int32_t* cpw, int32_t* twl, int32_t* tab);
int ygpu_set_tile_ex(ygpu_ctx* ctx, int32_t cpw, int32_t twl, int32_t tab);
/* The candidates of the last ygpu_autotune and their best launch times (ms):
 * writes min(max, count) of them and sets *count. */
int ygpu_autotune_results(ygpu_ctx* ctx, int32_t max, int32_t* cpw, int32_t* twl, int32_t* tab, double* ms, int32_t* count);

/* hipGraph of the key-grouped pipeline (blind rotate -> extract -> keyswitch)
 * over fixed device buffers, captured once and replayed with one launch.
 * Arguments as ygpu_pbs_jobs_device; the graph owns its scratch (key slots,
 * keyswitch partial sums, and the accumulators when d_acc_out is NULL).  Keys
 * used by a graph must stay resident while it exists. */
int ygpu_graph_create(ygpu_ctx* ctx, const ygpu_job* jobs, int32_t n_jobs, int64_t total, const int32_t* d_lwe_in,
                      const void* d_v, int32_t v_per_ct, int32_t* d_lwe_out, void* d_acc_out, int32_t* graph_id);
int ygpu_graph_launch(ygpu_ctx* ctx, int32_t graph_id);
int ygpu_graph_destroy(ygpu_ctx* ctx, int32_t graph_id);

/* Device memory and stream helpers. */
int ygpu_malloc(ygpu_ctx* ctx, size_t bytes, void** d_ptr);
int ygpu_free(ygpu_ctx* ctx, void* d_ptr);
int ygpu_memcpy_htod(ygpu_ctx* ctx, void* d_dst, const void* src, size_t bytes);
int ygpu_memcpy_dtoh(ygpu_ctx* ctx, void* dst, const void* d_src, size_t bytes);
int ygpu_sync(ygpu_ctx* ctx);

/* Kernel timing: when enabled, the library brackets each kernel it launches
 * with HIP events on the stream the kernel runs on and accumulates the
 * durations per kernel name. */
int ygpu_timing_enable(ygpu_ctx* ctx, int32_t on);
int ygpu_timing_get(ygpu_ctx* ctx, const char* kernel, int64_t* launches, double* total_ms);
int ygpu_timing_reset(ygpu_ctx* ctx);

/* Measurement helpers (no counterpart in the reference).
 * ygpu_modmul_peak: throughput of a microbenchmark running independent chains
 * of modular products mod the context's first transform prime; kind 0 is the
 * binary64 exact product the engine uses, kind 1 a 64-bit integer Shoup
 * product.  The larger is the roofline denominator for the batched forms.
 * ygpu_copy_device: a plain streaming copy (bytes % 16 == 0), the calibration
 * kernel for HBM byte counters. */
int ygpu_modmul_peak(ygpu_ctx* ctx, int32_t kind, double* modmul_per_sec);
int ygpu_copy_device(ygpu_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
/* Bound on the polls of each team-barrier wait inside the blind-rotation
 * kernels (default 65536, milliseconds against microseconds of real waiting).
 * A wait that runs out marks the workgroup failed: its accumulators read the
 * most negative torus value and the call reports YGPU_ERR_DEVICE.  polls = 0
 * fails every team barrier at once without waiting, which exercises that
 * report path with every kernel still finishing normally (a test hook; no
 * counterpart in the reference). */
int ygpu_set_barrier_polls(ygpu_ctx* ctx, int32_t polls);

/* Seeded key generation on the host (genTlweKey, genTrlweKey,
 * genBootstrappingKeyMP, genTlweKeySwitchingKey: tlwe.cpp, trlwe.cpp:70-81,
 * bootstrapping.cpp:150-174, keyswitching.cpp:19-29), with Gaussian noise of
 * standard deviation 2^sigma_log2_lwe / 2^sigma_log2_glwe (fractions of the
 * torus) rounded to integers.  Mask polynomials of the BSK are uniform mod qNtt,
 * as the reference's NTT-domain sampling makes them.  Outputs:
 * lwe_key int32[n] (bits), glwe_key int32[kN] (bits), bsk_coeff (layout above),
 * ksk (layout above).  Deterministic in `seed` for any thread count. */
int ygpu_keygen(const ygpu_params* params, uint64_t seed, double sigma_log2_lwe, double sigma_log2_glwe,
                int32_t* lwe_key, int32_t* glwe_key, uint64_t* bsk_coeff, int32_t* ksk, int32_t threads);

/* Seeded classic bootstrapping key (BootstrappingKey, bootstrapping.h:16-45;
 * genBootstrappingKey -> genBootstrappingKeyNormal -> encryptTrgswNtt,
 * bootstrapping.cpp:40-44, 98-104, trgsw.cpp:185-191, 227-246, 281-284): masks
 * uniform over the torus, mu_lvl = s_i 2^(w - (lvl+1) b) added in the torus to
 * a_row[0] of row r < k (the body keeps the mask's product from before) and to
 * the torus value of the body of row k (centred and reduced to the torus, so
 * that row's body residues are lifted torus values).  Layout, secret keys and
 * KSK as ygpu_keygen; bootstraps through ygpu_pbs like any coefficient-form
 * key (functionalBootstrappingNtt with l = lApprox). */
int ygpu_keygen_classic(const ygpu_params* params, uint64_t seed, double sigma_log2_lwe, double sigma_log2_glwe,
                        int32_t* lwe_key, int32_t* glwe_key, uint64_t* bsk_coeff, int32_t* ksk, int32_t threads);

/* Seeded WWL+24 succinct key generation (genBootstrappingKeyWWL24,
 * bootstrapping.cpp:106-126): cPrime[lvl] of step i encrypts s_i * 2^(w -
 * (lvl+1) b) at coefficient 0 (symEncTrlweMultiSampleNtt, trlwe.cpp:103-106:
 * mask uniform over the torus, body completed mod qNtt); s2[l2] is the TRLEV
 * (a - s * 2^(w - (l2+1) b), a s + e) of symEncTrlevWithKeyNtt (trlev.cpp:75-94),
 * formed in torus arithmetic.  Secret keys and KSK equal ygpu_keygen's for the
 * same seed.  Outputs as ygpu_key_upload_succinct's inputs. */
int ygpu_keygen_succinct(const ygpu_params* params, uint64_t seed, double sigma_log2_lwe, double sigma_log2_glwe, int32_t s2_level,
                         int32_t* lwe_key, int32_t* glwe_key, uint64_t* bsk_half, uint64_t* s2, int32_t* ksk, int32_t threads);

/* LWE encryption of torus messages mu[count] (symEncTlwe, tlwe.cpp:20-31) under
 * lwe_key, Gaussian noise 2^sigma_log2, seeded.  out: int32[count][n+1]. */
int ygpu_lwe_encrypt(int32_t n, const int32_t* lwe_key, uint64_t seed, double sigma_log2, int32_t count,
                     const int32_t* mu, int32_t* out);

/* Key generation and encryption with the reference's own noise law,
 * TUniform (genTUniformNoise / addTUniformNoise, numeric.cpp:47-60): integers
 * in [-2^b, 2^b] in torus units, the end points with probability 2^-(b+2) and
 * every other point with probability 2^-(b+1); b is YatfheParameters::
 * lweNoiseB for LWE and KSK rows and rlweNoiseB for the GLWE rows of the
 * bootstrapping key (yatfhe_parameters.h:31-51).  `form` selects the key:
 * YGPU_KEY_MP (outputs as ygpu_keygen), YGPU_KEY_CLASSIC (as
 * ygpu_keygen_classic) or YGPU_KEY_SUCCINCT (as ygpu_keygen_succinct: bsk_out
 * is bsk_half, s2_out and s2_level are used).  Same seeded streams as the
 * Gaussian forms: the secret keys equal theirs for the same seed. */
#define YGPU_KEY_MP 0
#define YGPU_KEY_CLASSIC 1
#define YGPU_KEY_SUCCINCT 2
int ygpu_keygen_tuniform(const ygpu_params* params, uint64_t seed, int32_t lwe_bound_bits, int32_t glwe_bound_bits, int32_t form,
                         int32_t s2_level, int32_t* lwe_key, int32_t* glwe_key, uint64_t* bsk_out, uint64_t* s2_out,
                         int32_t* ksk, int32_t threads);
int ygpu_lwe_encrypt_tuniform(int32_t n, const int32_t* lwe_key, uint64_t seed, int32_t bound_bits, int32_t count,
                              const int32_t* mu, int32_t* out);

/* ------------------------------------------------------ key interchange file
 * The canonical coefficient form (the layouts of ygpu_key_upload) as one file:
 * a 128-byte header (magic "YGPUKEY1", version, ygpu_params, word counts,
 * digests of each section and of the header) followed by the BSK residues
 * (uint64) and the KSK (int32).  Replaces the reference's NTT-form key stream
 * (ya_serializer.cpp:21-30, 76-98, 176-218), whose words are in Intel HEXL's
 * internal order; the reference has no KSK serializer.  Writes go to
 * `path`.tmp and are renamed into place. */
int ygpu_key_file_write(const char* path, const ygpu_params* params, const uint64_t* bsk_coeff, const int32_t* ksk);
/* Reads and validates the header only. */
int ygpu_key_file_info(const char* path, ygpu_params* params);
/* Reads a key file into caller buffers sized for its parameters; `expect`
 * (optional) must equal the file's parameters.  YGPU_ERR_CORRUPT on a bad
 * magic, version, size, digest or out-of-range residue. */
int ygpu_key_file_read(const char* path, const ygpu_params* expect, uint64_t* bsk_coeff, int32_t* ksk);
/* The WWL+24 succinct form (inputs as ygpu_key_upload_succinct) as a version-2
 * key file: the BSK section holds the body-digit rows, then the s^2 TRLEV; the
 * header records s2_level.  Replaces the reference's serializeBskWWL24 /
 * deserializeBskWWL24 streams (ya_serializer.cpp), which hold HEXL-order words. */
int ygpu_key_file_write_succinct(const char* path, const ygpu_params* params, const uint64_t* bsk_half, const uint64_t* s2,
                                 int32_t s2_level, const int32_t* ksk);
/* The key form of a file: *s2_level = 0 for a full key, > 0 for a succinct key. */
int ygpu_key_file_form(const char* path, int32_t* s2_level);
int ygpu_key_file_read_succinct(const char* path, const ygpu_params* expect, uint64_t* bsk_half, uint64_t* s2, int32_t* ksk);
/* Reads a key file of either form and uploads it (ygpu_key_upload or
 * ygpu_key_upload_succinct). */
int ygpu_key_upload_file(ygpu_ctx* ctx, const char* path, int32_t* key_id);

/* ------------------------------------------------------- lazy-pipe keys
 * The lazy-pipe ("OURS") key format, BootstrappingKeyMPLazyPipe
 * (yatfhe/include/yatfhe/bootstrapping.h:250-282), TORUS32, binary keys.  Its
 * rows have R = params->level gadget levels (param.lApprox, the accumulator's
 * decomposition); each row is stored as D = digit_levels zero-mean digits
 * (param.l; the reference's TORUS32 default 2^4 x 8 keeps all 32 bits).  For
 * every LWE key bit s_{i+1} (i < n-1) and row level lvl the key holds the D
 * digits of the half RGSW row cPrime[lvl] (bskDecompA / bskDecompB), bskFirst
 * (a TRLWE of v s_0: the key belongs to one test polynomial v) and the TRLEV
 * of s^2 with D levels (s2Dft).  Layouts:
 *   first     int32 [k+1][N]            bskFirst[0]: a[0..k-1], b
 *   digits_a  int8  [n-1][R][D][k][N]   digit l of cPrime[lvl].a[k] of s_{i+1}
 *   digits_b  int8  [n-1][R][D][N]      digit l of cPrime[lvl].b
 *   s2        uint64 residues mod q_ntt [D][k+1][N]
 * ygpu_keygen_lazy replaces genBootstrappingKeyMPLazyPipe
 * (bootstrapping.cpp:269-320): the half rows are ygpu_keygen_succinct's for the
 * same seed (their torus values, cut into digits), the secret keys and KSK
 * equal ygpu_keygen's.  The bootstrap replaces blindRotateLazyPipeNtt
 * (blind_rotate.cpp:822-902) followed by the extraction and keyswitch of
 * functionalBootstrapping: acc = X^{-b} (R(v) + (X^{a_0} - 1) bskFirst), then
 * for every step with a_i != 0 the stored digits of s_i are rotated by
 * (X^{a_i} - 1) (narrowed to int8, the gadget's 1 added on the diagonal digit),
 * scheme-switched through s^2 into RGSW rows, and acc := rows (x) acc.  The
 * rows are rebuilt for every ciphertext (in windows of steps on the device):
 * 2 R D forward transforms per step where an MP key's step needs 2 R.
 * Instantiated for N = 1024, k = 1, bases up to 2^7, R <= D <= 8, D b <= 32
 * (YGPU_ERR_UNSUPPORTED otherwise; gadget 2^7 x 3 has its own kernel form);
 * the engine-form entry points reject lazy-pipe key ids and these reject
 * other keys. */
int ygpu_keygen_lazy(const ygpu_params* params, uint64_t seed, double sigma_log2_lwe, double sigma_log2_glwe, int32_t digit_levels,
                     const int32_t* v, int32_t* lwe_key, int32_t* glwe_key, int32_t* first, int8_t* digits_a, int8_t* digits_b,
                     uint64_t* s2, int32_t* ksk, int32_t threads);
int ygpu_key_upload_lazy(ygpu_ctx* ctx, int32_t digit_levels, const int32_t* first, const int8_t* digits_a, const int8_t* digits_b,
                         const uint64_t* s2, const int32_t* ksk, int32_t* key_id);
/* One ciphertext, host buffers (acc_out and extract_out optional). */
int ygpu_pbs_lazy(ygpu_ctx* ctx, int32_t key_id, const int32_t* lwe_in, const void* v, int32_t* lwe_out, void* acc_out,
                  int32_t* extract_out);
/* blindRotateLazyPipeNtt itself: abar is the rescaled input (ScaledTlwe: a_0
 * .. a_{n-1}, b, already mod 2N), acc_out the accumulator. */
int ygpu_blind_rotate_lazy(ygpu_ctx* ctx, int32_t key_id, const int32_t* abar, const void* v, void* acc_out);
/* A batch on device buffers, as ygpu_pbs_batch_device. */
int ygpu_pbs_lazy_batch_device(ygpu_ctx* ctx, int32_t key_id, int32_t count, const int32_t* d_lwe_in, const void* d_v,
                               int32_t v_per_ct, int32_t* d_lwe_out, void* d_acc_out, int32_t* d_extract_out);
/* The Alt form, BootstrappingKeyMPLazyPipeAlt (bootstrapping.h:284-306): the
 * lazy-pipe format the reference's serving benchmark runs as "OURS"
 * (blindrotate_serving.cpp:480-488).  It stores the half RGSW rows of every s_i
 * in torus form, rows int32 [n][R][k+1][N] (bskPrime[i].cPrime[lvl]), and the
 * TRLEV of s^2 with D = digit_levels levels; it does not depend on v.  The
 * bootstrap replaces blindRotateLazyPipeAltNtt (blind_rotate.cpp:996-1136):
 * the first component GD(v) (x) RLEV(s_0) (deriveFirstComponentNtt with
 * prepareGdV, trgsw.cpp:911-955, R digits of v) is formed once per test
 * polynomial; every step rotates the torus rows by (X^{a_i} - 1), adds the
 * gadget to the body (rotateTrlweMinusOneBPlusOne), cuts the mask into D
 * digits and scheme-switches them through s^2 (the body is transformed whole),
 * D + 1 forward transforms per row level.  The lazy-pipe entry points above
 * run Alt keys too.  ygpu_keygen_lazy_alt replaces
 * genBootstrappingKeyMPLazyPipeAlt (bootstrapping.cpp:322-339): the rows are
 * ygpu_keygen_succinct's for the same seed in torus form. */
int ygpu_keygen_lazy_alt(const ygpu_params* params, uint64_t seed, double sigma_log2_lwe, double sigma_log2_glwe, int32_t digit_levels,
                         int32_t* lwe_key, int32_t* glwe_key, int32_t* rows, uint64_t* s2, int32_t* ksk, int32_t threads);
int ygpu_key_upload_lazy_alt(ygpu_ctx* ctx, int32_t digit_levels, const int32_t* rows, const uint64_t* s2, const int32_t* ksk,
                             int32_t* key_id);
/* Both lazy-pipe forms with the reference's TUniform noise (bounds as
 * ygpu_keygen_tuniform: lweNoiseB, rlweNoiseB). */
int ygpu_keygen_lazy_tuniform(const ygpu_params* params, uint64_t seed, int32_t lwe_bound_bits, int32_t glwe_bound_bits,
                              int32_t digit_levels, const int32_t* v, int32_t* lwe_key, int32_t* glwe_key, int32_t* first,
                              int8_t* digits_a, int8_t* digits_b, uint64_t* s2, int32_t* ksk, int32_t threads);
int ygpu_keygen_lazy_alt_tuniform(const ygpu_params* params, uint64_t seed, int32_t lwe_bound_bits, int32_t glwe_bound_bits,
                                  int32_t digit_levels, int32_t* lwe_key, int32_t* glwe_key, int32_t* rows, uint64_t* s2, int32_t* ksk,
                                  int32_t threads);
/* Steps per window of derived rows (0: the most that fit a 4 GiB workspace for
 * the batch); results do not depend on it.  No counterpart in the reference. */
int ygpu_set_lazy_window(ygpu_ctx* ctx, int32_t steps);

/* ---------------------------------------------------------------- serving
 * Continuous-batching server on one GPU.  Replaces the reference's serving
 * loop (ya_benchmark/blindrotate_serving.cpp:121-353): its shared key cache
 * with single-flight loads (ConcurrentKeyCache,
 * ya_benchmark/include/blindrotate_serving.h:25-111), its request scheduler
 * (Scheduler, 132-273) and its per-request worker bootstrap.  Requests queue
 * per parameter set; whenever one of a parameter set's lanes (streams with
 * their own staging buffers) is free, everything pending is launched as one
 * key-grouped batch, so batches overlap on the device.  Tenant keys live in a
 * device arena of key_budget_bytes under LRU eviction, loaded on demand from
 * their device-ready form in pinned host memory; a tenant whose key cannot be
 * placed while other batches hold the memory waits in the queue (admission by
 * device memory).  Results are bit-identical to the single-context entry
 * points: each ciphertext is computed independently of its batch. */
typedef struct ygpu_server ygpu_server;

typedef struct ygpu_server_config {
    int32_t device;
    int32_t lanes;            /* in-flight batches per parameter set (0: fitted to the
                                 hardware queues, see ygpu_server_lanes) */
    int32_t max_batch;        /* ciphertexts per launch (0: 4096) */
    int32_t queue_max;        /* queued requests per parameter set before submit sheds (0: 2^20) */
    int32_t max_tenants;      /* tenant table size (0: 4096) */
    int32_t max_loads;        /* key loads in flight per parameter set (0: 4) */
    int32_t eviction;         /* YGPU_EVICT_FREQ_BYTES (0) or YGPU_EVICT_LRU (1) */
    int32_t reserved;
    int64_t key_budget_bytes; /* device bytes for resident keys (0: 3/4 of free memory) */
    int32_t sched;            /* YGPU_SCHED_SHARED (0), YGPU_SCHED_AFFINITY (1) or YGPU_SCHED_SOFT (2) */
    int32_t spill_threshold;  /* soft: owner-lane depth (requests) at which a request spills (0: 2) */
    int32_t steal_threshold;  /* soft: depth (requests) a lane needs before an idle lane steals (0: 2) */
    int32_t reserved2;
} ygpu_server_config;

/* Request-to-lane scheduling of the server, the modes of the reference's
 * serving Scheduler (ya_benchmark/include/blindrotate_serving.h:132-273) with
 * the server's lanes (in-flight batches of a parameter set) as its workers.
 * SHARED: any free lane batches any tenant's requests.  AFFINITY: tenant t's
 * requests are batched only by lane t % lanes of its parameter set.  SOFT:
 * affinity, but a request whose owner lane already has spill_threshold
 * requests waiting goes to the shortest lane when that is shorter by more
 * than one, and a lane with nothing of its own takes (steals) the requests of
 * the longest lane once that holds steal_threshold or more. */
#define YGPU_SCHED_SHARED 0
#define YGPU_SCHED_AFFINITY 1
#define YGPU_SCHED_SOFT 2

/* Key eviction policies of the server's device cache.  FREQ_BYTES evicts the
 * keys whose loss costs the least host-to-device traffic: decayed request rate
 * (time constant 1 s) times the key's transfer bytes; LRU evicts the least
 * recently used.  Both evict only idle keys and keep keys contiguous. */
#define YGPU_EVICT_FREQ_BYTES 0
#define YGPU_EVICT_LRU 1

/* One bootstrap request.  The caller owns the struct and the buffers, which
 * must stay valid until the request completes (ygpu_server_wait). */
typedef struct ygpu_request {
    int32_t tenant;
    int32_t batch_size;       /* out: ciphertexts in the launch that served it */
    const int32_t* lwe_in;    /* int32[n+1] */
    const void* v;            /* test polynomial, Torus[N] */
    int32_t* lwe_out;         /* int32[n+1], written on completion */
    int64_t arrival_ns;       /* in: steady-clock arrival (0: time of submit) */
    int64_t launch_ns;        /* out: when its batch was launched */
    int64_t done_ns;          /* out: when its result was copied out */
    uint64_t seq;             /* internal */
    int32_t status;           /* out: 1 pending, YGPU_OK done, < 0 error */
    int32_t key_hit;          /* out: the tenant's key was resident when the request arrived */
    int32_t lane;             /* out: lane of its parameter set that served it */
    int32_t coalesced;        /* out: its key was absent and already awaited by an earlier request
                                 (it shared that load: the reference's waitedForPeer) */
} ygpu_request;

typedef struct ygpu_server_stats {
    int64_t submitted, completed, failed, shed;
    int64_t batches, batched_cts, request_hits;
    int64_t key_loads, key_evictions, bytes_loaded, memory_waits;
    int64_t resident_keys, resident_bytes, peak_resident_bytes, key_budget_bytes, host_key_bytes;
    int64_t coalesced, spilled, stolen, peak_queue_depth;
    int64_t loads_in_flight, peak_loads_in_flight;  /* key loads (copy + expansion) under way */
    int64_t load_inflight_ns;                       /* integral over time of loads in flight (ns) */
    int64_t file_reads, file_read_bytes, file_read_ns;  /* key files read on a miss: count, transfer-form bytes, time */
} ygpu_server_stats;

int ygpu_server_create(const ygpu_server_config* cfg, ygpu_server** out);
int ygpu_server_destroy(ygpu_server* s);
/* Parameter sets (one engine context each); both torus widths may coexist. */
int ygpu_server_add_params(ygpu_server* s, const ygpu_params* params, int32_t* pset);
/* Lanes of a parameter set (made at the first request or prefetch).  With
 * lanes = 0 in the configuration the server's streams are fitted to the
 * device's hardware queues (GPU_MAX_HW_QUEUES, 4 by default, read when the
 * server starts): one copy stream for every key load plus lanes, one queue
 * each, so no lane waits behind another's kernels; every parameter set has
 * one lane and the remaining queues go to the sets of the cheapest blind
 * rotation first (at most 3 a set).  The server assumes no other stream of
 * the process is busy on the device while it serves.  (No counterpart in the
 * reference, whose workers are CPU threads.) */
int ygpu_server_lanes(ygpu_server* s, int32_t pset, int32_t* lanes);
int ygpu_server_set_tile(ygpu_server* s, int32_t pset, int32_t cpw, int32_t twl);
/* Extraction coefficient (param.driftPhase) of a parameter set's bootstraps;
 * before the first request.  See ygpu_set_extract_index. */
int ygpu_server_set_extract_index(ygpu_server* s, int32_t pset, int32_t index);
/* Registers a tenant's keys (coefficient form, layouts as ygpu_key_upload):
 * transformed once into device-ready form in pinned host memory, loaded into
 * the device arena on first use. */
int ygpu_server_add_tenant(ygpu_server* s, int32_t pset, const uint64_t* bsk_coeff, const int32_t* ksk, int32_t* tenant);
/* A tenant with a WWL+24 succinct key (inputs as ygpu_key_upload_succinct):
 * the host holds and transfers the succinct form, about half the full key's
 * bytes; each load derives the full rows on the device. */
int ygpu_server_add_tenant_succinct(ygpu_server* s, int32_t pset, const uint64_t* bsk_half, const uint64_t* s2, int32_t s2_level,
                                    const int32_t* ksk, int32_t* tenant);
/* A new tenant sharing another tenant's host key material but cached on the
 * device on its own (separate residency, loads and evictions). */
int ygpu_server_add_tenant_shared(ygpu_server* s, int32_t src_tenant, int32_t* tenant);
/* A tenant whose key stays in a key file (either form, ygpu_key_file_write*)
 * until a load needs it: the lowest tier, below pinned host memory.  Replaces
 * the reference's DiskReader behind its serving key cache
 * (ya_benchmark/include/cache_manager.h:13-51; the loader reads the tenant's
 * key from file on a miss, ya_benchmark/blindrotate_serving.cpp:210-219).
 * Registration reads and checks the header only (parameters must equal the
 * set's).  On a miss one reader thread reads the whole file (digests checked)
 * into one of the set's pinned slots, packs it as ygpu_server_add_tenant
 * would, and the device load follows; a slot keeps its form until it is
 * reused (least recently used first, never while a load reads it).  A file
 * that fails to read fails the tenant's pending requests with its error
 * (YGPU_ERR_CORRUPT, YGPU_ERR_IO ...) and is read again on the next miss. */
int ygpu_server_add_tenant_file(ygpu_server* s, int32_t pset, const char* path, int32_t* tenant);
/* Pinned slots of a parameter set for file-backed keys (default 2, at most
 * 64), each a full key's transfer form; allocated with the set's first
 * file-backed tenant, so set before it. */
int ygpu_server_set_file_slots(ygpu_server* s, int32_t pset, int32_t slots);
int ygpu_server_tenant_params(ygpu_server* s, int32_t tenant, int32_t* pset);
/* Loads a tenant's key into the device cache now (evicting idle keys if
 * needed) and waits for it; YGPU_ERR_BUSY when no memory can be freed. */
int ygpu_server_prefetch(ygpu_server* s, int32_t tenant);
/* Asynchronous submit (thread-safe).  YGPU_ERR_BUSY when the queue is full. */
int ygpu_server_submit(ygpu_server* s, ygpu_request* req);
/* Blocks until the request completes; returns its status. */
int ygpu_server_wait(ygpu_server* s, ygpu_request* req);
/* Waits at most timeout_ms; returns the request's status (1 while pending). */
int ygpu_server_wait_for(ygpu_server* s, ygpu_request* req, int32_t timeout_ms);
int ygpu_server_stats_get(ygpu_server* s, ygpu_server_stats* out);

/* Open-loop load generator over the server (no counterpart in the reference
 * library; it mirrors the reference's serving benchmark driver): Poisson
 * arrivals at `rate` requests/s over tenants 0..n_tenants-1 with Zipf(zipf_s)
 * popularity (rank r is tenant r), every request of tenant t bootstrapping
 * lwe_in[t] with v[t]; when expected[t] is given each result is compared with
 * it.  Latency runs from a request's scheduled arrival to its completion. */
typedef struct ygpu_loadgen_config {
    double rate;
    int64_t requests;
    int64_t warm;             /* burst of warm-up requests before the measured run */
    double zipf_s;
    uint64_t seed;
    int32_t n_tenants;
    int32_t reserved;
    const int32_t* const* lwe_in;
    const void* const* v;
    const int32_t* const* expected;   /* may be NULL, or hold NULL entries */
    double slo_ms;
} ygpu_loadgen_config;

typedef struct ygpu_loadgen_result {
    double offered_rps, throughput_rps, wall_sec;
    int64_t completed, shed, errors, mismatches;
    double mean_ms, p50_ms, p90_ms, p99_ms, p99_9_ms, max_ms;
    double queue_p50_ms, queue_p99_ms, slo_violation_rate;
    int64_t batches;
    double mean_batch, hit_rate;
    int64_t key_loads, key_evictions, bytes_loaded, memory_waits;
    double hit_p50_ms, hit_p99_ms;      /* requests whose key was resident on arrival */
    double miss_p50_ms, miss_p99_ms;    /* requests that waited for a key load */
    double pset_p50_ms[4], pset_p99_ms[4];  /* by parameter set (first four) */
    /* the rest of the reference's serving record (blindrotate_serving.cpp:289-353) */
    double sd_ms;                       /* latency (sojourn) standard deviation */
    double queue_mean_ms, queue_p90_ms, queue_max_ms;
    double service_mean_ms, service_p50_ms, service_p90_ms, service_p99_ms, service_max_ms;  /* launch to result */
    double slo_violation_rate_100, slo_violation_rate_300, slo_violation_rate_500;
    int64_t coalesced_loads, peak_queue_depth, spilled, stolen;
    double load_concurrency_mean;       /* key loads under way, time-averaged over the run */
    int64_t load_concurrency_peak;
    double tenant_fairness_jain;        /* Jain's index of the per-tenant mean latencies */
    double locality;                    /* share served by the tenant's owner lane (0 under SHARED) */
    int64_t lane_requests[16];          /* requests served, by parameter set * lanes + lane (first 16) */
} ygpu_loadgen_result;

int ygpu_server_loadgen(ygpu_server* s, const ygpu_loadgen_config* cfg, ygpu_loadgen_result* out);

#ifdef __cplusplus
}
#endif
#endif /* YATFHE_GPU_H */
