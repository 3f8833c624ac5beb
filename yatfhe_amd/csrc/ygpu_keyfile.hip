// Key interchange file: the canonical coefficient form of a bootstrapping key
// and its keyswitching key, as one file.
//
// The reference serializes its NTT-form BSK words (ya_serializer.cpp:21-30,
// 76-98, 176-218) in Intel HEXL's word order, which is internal to HEXL, and has
// no KSK serializer.  This file format carries what crosses the engine's
// boundary instead: every key polynomial's residues mod qNtt (the inverse
// transform without centring), which any correct NTT implementation can
// produce and consume, plus the KSK.
//
// Layout (little-endian):
//   offset  0  char[8]  magic "YGPUKEY1"
//           8  uint32   format version (1)
//          12  uint32   header bytes (128)
//          16  ygpu_params (48 bytes: ten int32 fields, then q_ntt as uint64)
//          64  uint64   BSK words  = n * level * (k+1)^2 * N
//          72  uint64   KSK words  = k * N * ks_level * (n+1)
//          80  uint64   BSK digest
//          88  uint64   KSK digest
//          96  uint64   header digest (of bytes 0..95)
//         104  zero padding to 128
//         128  uint64[BSK words]   bsk[i][lvl][row][col][N], residues in [0, q_ntt)
//              int32[KSK words]    ksk[i][j][n+1]
// Digest of a word sequence w_0..w_{m-1}: XOR over i of
// mix(w_i XOR i * 0x9e3779b97f4a7c15), mix = the SplitMix64 finalizer; it
// detects truncation, reordering and bit rot (it is not an authenticator).
#include "ygpu_internal.h"

#include <cstdio>
#include <cstring>
#include <string>

namespace {

using namespace ygpu_rt;

constexpr char MAGIC[8] = {'Y', 'G', 'P', 'U', 'K', 'E', 'Y', '1'};
constexpr uint32_t VERSION = 1;
constexpr uint32_t HEADER_BYTES = 128;

struct FileHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    ygpu_params params;
    uint64_t bsk_words;
    uint64_t ksk_words;
    uint64_t bsk_digest;
    uint64_t ksk_digest;
    uint64_t header_digest;
    uint8_t pad[24];
};
static_assert(sizeof(ygpu_params) == 48, "ygpu_params layout");
static_assert(sizeof(FileHeader) == HEADER_BYTES, "header layout");

inline uint64_t mix(uint64_t z) {
    z ^= z >> 30;
    z *= 0xbf58476d1ce4e5b9ULL;
    z ^= z >> 27;
    z *= 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

template <class W>
uint64_t digest(const W* w, size_t count, size_t first_index = 0) {
    uint64_t h = 0;
    for (size_t i = 0; i < count; i++)
        h ^= mix(static_cast<uint64_t>(static_cast<typename std::make_unsigned<W>::type>(w[i])) ^
                 ((first_index + i) * 0x9e3779b97f4a7c15ULL));
    return h;
}

uint64_t header_digest(const FileHeader& h) {
    return digest(reinterpret_cast<const uint8_t*>(&h), offsetof(FileHeader, header_digest));
}

void sizes(const ygpu_params& p, uint64_t* bsk, uint64_t* ksk) {
    *bsk = static_cast<uint64_t>(p.n) * p.level * (p.k + 1) * (p.k + 1) * p.N;
    *ksk = static_cast<uint64_t>(p.k) * p.N * p.ks_level * (p.n + 1);
}

bool same_params(const ygpu_params& a, const ygpu_params& b) {
    return a.torus_bits == b.torus_bits && a.n == b.n && a.k == b.k && a.N == b.N && a.radix_bits == b.radix_bits &&
           a.level == b.level && a.ks_radix_bits == b.ks_radix_bits && a.ks_level == b.ks_level &&
           a.ks_width_bits == b.ks_width_bits && a.lwe_q_bits == b.lwe_q_bits && a.q_ntt == b.q_ntt;
}

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

int read_header(FILE* f, const char* path, FileHeader& h) {
    if (std::fread(&h, 1, sizeof(h), f) != sizeof(h))
        return fail(YGPU_ERR_CORRUPT, std::string(path) + ": shorter than a key file header");
    if (std::memcmp(h.magic, MAGIC, sizeof(MAGIC)) != 0) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": not a key file");
    if (h.version != VERSION || h.header_bytes != HEADER_BYTES)
        return fail(YGPU_ERR_CORRUPT, std::string(path) + ": unsupported key file version " + std::to_string(h.version));
    if (header_digest(h) != h.header_digest) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": header digest mismatch");
    uint64_t b = 0, k = 0;
    sizes(h.params, &b, &k);
    if (b != h.bsk_words || k != h.ksk_words) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": sizes disagree with parameters");
    return YGPU_OK;
}

}  // namespace

extern "C" {

int ygpu_key_file_write(const char* path, const ygpu_params* params, const uint64_t* bsk_coeff, const int32_t* ksk) {
    if (!path || !params || !bsk_coeff || !ksk) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    int rc = check_params(*params);
    if (rc) return rc;
    FileHeader h{};
    std::memcpy(h.magic, MAGIC, sizeof(MAGIC));
    h.version = VERSION;
    h.header_bytes = HEADER_BYTES;
    h.params = *params;
    sizes(*params, &h.bsk_words, &h.ksk_words);
    for (uint64_t i = 0; i < h.bsk_words; i++)
        if (bsk_coeff[i] >= params->q_ntt) return fail(YGPU_ERR_INVALID_ARG, "BSK coefficient " + std::to_string(i) + " is not below q_ntt");
    h.bsk_digest = digest(bsk_coeff, h.bsk_words);
    h.ksk_digest = digest(ksk, h.ksk_words);
    h.header_digest = header_digest(h);
    const std::string tmp = std::string(path) + ".tmp";
    {
        File f;
        f.f = std::fopen(tmp.c_str(), "wb");
        if (!f.f) return fail(YGPU_ERR_IO, tmp + ": cannot open for writing");
        if (std::fwrite(&h, 1, sizeof(h), f.f) != sizeof(h) ||
            std::fwrite(bsk_coeff, sizeof(uint64_t), h.bsk_words, f.f) != h.bsk_words ||
            std::fwrite(ksk, sizeof(int32_t), h.ksk_words, f.f) != h.ksk_words || std::fflush(f.f) != 0) {
            std::remove(tmp.c_str());
            return fail(YGPU_ERR_IO, tmp + ": write failed");
        }
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return fail(YGPU_ERR_IO, std::string(path) + ": rename failed");
    }
    return YGPU_OK;
}

int ygpu_key_file_info(const char* path, ygpu_params* params) {
    if (!path || !params) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) return fail(YGPU_ERR_IO, std::string(path) + ": cannot open");
    FileHeader h{};
    int rc = read_header(f.f, path, h);
    if (rc) return rc;
    *params = h.params;
    return YGPU_OK;
}

int ygpu_key_file_read(const char* path, const ygpu_params* expect, uint64_t* bsk_coeff, int32_t* ksk) {
    if (!path || !bsk_coeff || !ksk) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) return fail(YGPU_ERR_IO, std::string(path) + ": cannot open");
    FileHeader h{};
    int rc = read_header(f.f, path, h);
    if (rc) return rc;
    if (expect && !same_params(*expect, h.params))
        return fail(YGPU_ERR_INVALID_ARG, std::string(path) + ": key parameters differ from the expected set");
    if (std::fread(bsk_coeff, sizeof(uint64_t), h.bsk_words, f.f) != h.bsk_words ||
        std::fread(ksk, sizeof(int32_t), h.ksk_words, f.f) != h.ksk_words)
        return fail(YGPU_ERR_CORRUPT, std::string(path) + ": truncated");
    char extra;
    if (std::fread(&extra, 1, 1, f.f) != 0) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": trailing bytes");
    if (digest(bsk_coeff, h.bsk_words) != h.bsk_digest) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": BSK digest mismatch");
    if (digest(ksk, h.ksk_words) != h.ksk_digest) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": KSK digest mismatch");
    for (uint64_t i = 0; i < h.bsk_words; i++)
        if (bsk_coeff[i] >= h.params.q_ntt) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": BSK residue not below q_ntt");
    return YGPU_OK;
}

int ygpu_key_upload_file(ygpu_ctx* ctx, const char* path, int32_t* key_id) {
    if (!ctx || !path || !key_id) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    ygpu_params p{};
    int rc = ygpu_get_params(ctx, &p);
    if (rc) return rc;
    uint64_t b = 0, k = 0;
    sizes(p, &b, &k);
    std::vector<uint64_t> bsk(b);
    std::vector<int32_t> ksk(k);
    rc = ygpu_key_file_read(path, &p, bsk.data(), ksk.data());
    if (rc) return rc;
    return ygpu_key_upload(ctx, bsk.data(), ksk.data(), key_id);
}

}  // extern "C"
