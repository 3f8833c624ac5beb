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
//           8  uint32   format version: 1 full key, 2 WWL+24 succinct key
//          12  uint32   header bytes (128)
//          16  ygpu_params (48 bytes: ten int32 fields, then q_ntt as uint64)
//          64  uint64   BSK words: n * level * (k+1)^2 * N (full), or
//                       (n * level + s2_level) * (k+1) * N (succinct)
//          72  uint64   KSK words  = k * N * ks_level * (n+1)
//          80  uint64   BSK digest
//          88  uint64   KSK digest
//          96  uint64   header digest (of bytes 0..95)
//         104  uint32   s2_level (version 2; 0 in version 1)
//         108  uint32   zero
//         112  uint64   digest of bytes 104..111 (version 2; 0 in version 1)
//         120  zero padding to 128
//         128  uint64[BSK words]   full: bsk[i][lvl][row][col][N]; succinct: the
//                                  body-digit rows bsk_half[i][lvl][col][N], then
//                                  s2[l2][col][N]; residues in [0, q_ntt)
//              int32[KSK words]    ksk[i][j][n+1]
// The succinct form is the reference's BootstrappingKeyWWL24 (bootstrapping.h:
// 116-136; its stream is serializeBskWWL24, ya_serializer.cpp) in coefficient form.
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
constexpr uint32_t VERSION_FULL = 1, VERSION_SUCCINCT = 2;
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
    uint32_t s2_level;
    uint32_t zero;
    uint64_t ext_digest;
    uint8_t pad[8];
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
uint64_t ext_digest(const FileHeader& h) {
    return digest(reinterpret_cast<const uint8_t*>(&h) + offsetof(FileHeader, s2_level), 8);
}

// Words of the BSK section (s2_level 0: full key) and of the KSK section.
void sizes(const ygpu_params& p, int s2_level, uint64_t* bsk, uint64_t* ksk) {
    *bsk = s2_level > 0 ? (static_cast<uint64_t>(p.n) * p.level + s2_level) * (p.k + 1) * p.N
                        : static_cast<uint64_t>(p.n) * p.level * (p.k + 1) * (p.k + 1) * p.N;
    *ksk = static_cast<uint64_t>(p.k) * p.N * p.ks_level * (p.n + 1);
}

bool same_params(const ygpu_params& a, const ygpu_params& b) { return params_equal(a, b); }

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
    if ((h.version != VERSION_FULL && h.version != VERSION_SUCCINCT) || h.header_bytes != HEADER_BYTES)
        return fail(YGPU_ERR_CORRUPT, std::string(path) + ": unsupported key file version " + std::to_string(h.version));
    if (header_digest(h) != h.header_digest) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": header digest mismatch");
    const bool succinct = h.version == VERSION_SUCCINCT;
    if (succinct ? (ext_digest(h) != h.ext_digest || h.zero != 0) : (h.s2_level != 0 || h.zero != 0 || h.ext_digest != 0))
        return fail(YGPU_ERR_CORRUPT, std::string(path) + ": key form fields damaged");
    if (succinct && check_s2_level(h.params, static_cast<int>(h.s2_level)) != YGPU_OK)
        return fail(YGPU_ERR_CORRUPT, std::string(path) + ": bad s2_level " + std::to_string(h.s2_level));
    uint64_t b = 0, k = 0;
    sizes(h.params, static_cast<int>(h.s2_level), &b, &k);
    if (b != h.bsk_words || k != h.ksk_words) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": sizes disagree with parameters");
    return YGPU_OK;
}

}  // namespace

namespace {

// Writes a key file whose BSK section is the concatenation of `parts`
// (s2_level 0: one part, the full key; otherwise the body-digit rows and s^2).
int write_file(const char* path, const ygpu_params* params, int s2_level, const uint64_t* const* parts, const uint64_t* part_words,
               int nparts, const int32_t* ksk) {
    int rc = check_params(*params);
    if (rc) return rc;
    if (s2_level > 0 && (rc = check_s2_level(*params, s2_level)) != YGPU_OK) return rc;
    FileHeader h{};
    std::memcpy(h.magic, MAGIC, sizeof(MAGIC));
    h.version = s2_level > 0 ? VERSION_SUCCINCT : VERSION_FULL;
    h.header_bytes = HEADER_BYTES;
    h.params = *params;
    sizes(*params, s2_level, &h.bsk_words, &h.ksk_words);
    uint64_t first = 0;
    for (int pi = 0; pi < nparts; pi++) {
        for (uint64_t i = 0; i < part_words[pi]; i++)
            if (parts[pi][i] >= params->q_ntt)
                return fail(YGPU_ERR_INVALID_ARG, "BSK coefficient " + std::to_string(first + i) + " is not below q_ntt");
        h.bsk_digest ^= digest(parts[pi], part_words[pi], first);
        first += part_words[pi];
    }
    if (first != h.bsk_words) return fail(YGPU_ERR_INVALID_ARG, "BSK section size");
    h.ksk_digest = digest(ksk, h.ksk_words);
    if (s2_level > 0) {
        h.s2_level = static_cast<uint32_t>(s2_level);
        h.ext_digest = ext_digest(h);
    }
    h.header_digest = header_digest(h);
    const std::string tmp = std::string(path) + ".tmp";
    {
        File f;
        f.f = std::fopen(tmp.c_str(), "wb");
        if (!f.f) return fail(YGPU_ERR_IO, tmp + ": cannot open for writing");
        bool ok = std::fwrite(&h, 1, sizeof(h), f.f) == sizeof(h);
        for (int pi = 0; ok && pi < nparts; pi++) ok = std::fwrite(parts[pi], sizeof(uint64_t), part_words[pi], f.f) == part_words[pi];
        ok = ok && std::fwrite(ksk, sizeof(int32_t), h.ksk_words, f.f) == h.ksk_words && std::fflush(f.f) == 0;
        if (!ok) {
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

// Reads the BSK section into `parts` and the KSK, checking form, sizes, digests
// and residue range.
int read_file(const char* path, const ygpu_params* expect, int want_succinct, uint64_t* const* parts, const uint64_t* part_words,
              int nparts, int32_t* ksk) {
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) return fail(YGPU_ERR_IO, std::string(path) + ": cannot open");
    FileHeader h{};
    int rc = read_header(f.f, path, h);
    if (rc) return rc;
    if ((h.version == VERSION_SUCCINCT) != (want_succinct != 0))
        return fail(YGPU_ERR_INVALID_ARG, std::string(path) + (want_succinct ? ": holds a full key" : ": holds a succinct (WWL+24) key"));
    if (expect && !same_params(*expect, h.params))
        return fail(YGPU_ERR_INVALID_ARG, std::string(path) + ": key parameters differ from the expected set");
    uint64_t total = 0;
    for (int pi = 0; pi < nparts; pi++) total += part_words[pi];
    if (total != h.bsk_words) return fail(YGPU_ERR_INVALID_ARG, std::string(path) + ": buffers do not match the file's key form");
    uint64_t first = 0, dig = 0;
    for (int pi = 0; pi < nparts; pi++) {
        if (std::fread(parts[pi], sizeof(uint64_t), part_words[pi], f.f) != part_words[pi])
            return fail(YGPU_ERR_CORRUPT, std::string(path) + ": truncated");
        dig ^= digest(parts[pi], part_words[pi], first);
        first += part_words[pi];
    }
    if (std::fread(ksk, sizeof(int32_t), h.ksk_words, f.f) != h.ksk_words) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": truncated");
    char extra;
    if (std::fread(&extra, 1, 1, f.f) != 0) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": trailing bytes");
    if (dig != h.bsk_digest) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": BSK digest mismatch");
    if (digest(ksk, h.ksk_words) != h.ksk_digest) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": KSK digest mismatch");
    for (int pi = 0; pi < nparts; pi++)
        for (uint64_t i = 0; i < part_words[pi]; i++)
            if (parts[pi][i] >= h.params.q_ntt) return fail(YGPU_ERR_CORRUPT, std::string(path) + ": BSK residue not below q_ntt");
    return YGPU_OK;
}

}  // namespace

extern "C" {

int ygpu_key_file_write(const char* path, const ygpu_params* params, const uint64_t* bsk_coeff, const int32_t* ksk) {
    if (!path || !params || !bsk_coeff || !ksk) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    uint64_t b = 0, k = 0;
    sizes(*params, 0, &b, &k);
    const uint64_t* parts[1] = {bsk_coeff};
    return write_file(path, params, 0, parts, &b, 1, ksk);
}

int ygpu_key_file_write_succinct(const char* path, const ygpu_params* params, const uint64_t* bsk_half, const uint64_t* s2,
                                 int32_t s2_level, const int32_t* ksk) {
    if (!path || !params || !bsk_half || !s2 || !ksk) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    int rc = check_s2_level(*params, s2_level);
    if (rc) return rc;
    const uint64_t words[2] = {static_cast<uint64_t>(params->n) * params->level * (params->k + 1) * params->N,
                               static_cast<uint64_t>(s2_level) * (params->k + 1) * params->N};
    const uint64_t* parts[2] = {bsk_half, s2};
    return write_file(path, params, s2_level, parts, words, 2, ksk);
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

int ygpu_key_file_form(const char* path, int32_t* s2_level) {
    if (!path || !s2_level) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) return fail(YGPU_ERR_IO, std::string(path) + ": cannot open");
    FileHeader h{};
    int rc = read_header(f.f, path, h);
    if (rc) return rc;
    *s2_level = static_cast<int32_t>(h.s2_level);
    return YGPU_OK;
}

int ygpu_key_file_read(const char* path, const ygpu_params* expect, uint64_t* bsk_coeff, int32_t* ksk) {
    if (!path || !bsk_coeff || !ksk) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    ygpu_params p{};
    int rc = ygpu_key_file_info(path, &p);
    if (rc) return rc;
    uint64_t b = 0, k = 0;
    sizes(p, 0, &b, &k);
    uint64_t* parts[1] = {bsk_coeff};
    return read_file(path, expect, 0, parts, &b, 1, ksk);
}

int ygpu_key_file_read_succinct(const char* path, const ygpu_params* expect, uint64_t* bsk_half, uint64_t* s2, int32_t* ksk) {
    if (!path || !bsk_half || !s2 || !ksk) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    ygpu_params p{};
    int32_t s2_level = 0;
    int rc = ygpu_key_file_info(path, &p);
    if (rc == YGPU_OK) rc = ygpu_key_file_form(path, &s2_level);
    if (rc) return rc;
    if (s2_level == 0) return fail(YGPU_ERR_INVALID_ARG, std::string(path) + ": holds a full key");
    const uint64_t words[2] = {static_cast<uint64_t>(p.n) * p.level * (p.k + 1) * p.N, static_cast<uint64_t>(s2_level) * (p.k + 1) * p.N};
    uint64_t* parts[2] = {bsk_half, s2};
    return read_file(path, expect, 1, parts, words, 2, ksk);
}

int ygpu_key_upload_file(ygpu_ctx* ctx, const char* path, int32_t* key_id) {
    if (!ctx || !path || !key_id) return fail(YGPU_ERR_INVALID_ARG, "null argument");
    ygpu_params p{};
    int rc = ygpu_get_params(ctx, &p);
    if (rc) return rc;
    int32_t s2_level = 0;
    rc = ygpu_key_file_form(path, &s2_level);
    if (rc) return rc;
    uint64_t b = 0, k = 0;
    sizes(p, s2_level, &b, &k);
    std::vector<uint64_t> bsk(b);
    std::vector<int32_t> ksk(k);
    if (s2_level == 0) {
        rc = ygpu_key_file_read(path, &p, bsk.data(), ksk.data());
        if (rc) return rc;
        return ygpu_key_upload(ctx, bsk.data(), ksk.data(), key_id);
    }
    const size_t half = static_cast<size_t>(p.n) * p.level * (p.k + 1) * p.N;
    rc = ygpu_key_file_read_succinct(path, &p, bsk.data(), bsk.data() + half, ksk.data());
    if (rc) return rc;
    return ygpu_key_upload_succinct(ctx, bsk.data(), bsk.data() + half, s2_level, ksk.data(), key_id);
}

}  // extern "C"
