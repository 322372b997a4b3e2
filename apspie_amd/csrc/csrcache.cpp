// Binary CSR cache: a graph written once and loaded without parsing.
//
// The reference re-parses MatrixMarket text on every run (readEdge / readMtx,
// util.cuh:354-430) and converts COO to CSR with cusparseXcoo2csr
// (util.cuh:456-471).  Here a graph built once (read_mtx + from_coo on the
// device, or a generator) can be saved as its CSR arrays and loaded straight
// into HBM.
//
// Layout (little-endian, 64-byte header, arrays packed back to back):
//   0  char[8]  magic "APSPCSR1"
//   8  uint32   version (1)
//  12  uint32   flags: bit 0 = values present
//  16  int64    m (rows)
//  24  int64    n (columns)
//  32  int64    nnz
//  40  uint64   checksum of the three arrays' bytes (below)
//  48  16 bytes zero
//  64  int32    rowptr[m + 1], int32 col[nnz], float val[nnz] if flag bit 0
// Checksum: h = 0xcbf29ce484222325; for each 8-byte little-endian word x of
// the array bytes (the last word zero-padded): h = (h ^ x) * 0x100000001b3,
// then h ^= h >> 29 at the end -- FNV-1a's constants over words instead of
// bytes (eight times fewer multiplies; enough to catch truncation and
// corruption, not an integrity guarantee against tampering).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace {

constexpr char kMagic[8] = {'A', 'P', 'S', 'P', 'C', 'S', 'R', '1'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 64;

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t flags;
    int64_t m, n, nnz;
    uint64_t checksum;
    uint8_t pad[16];
};
static_assert(sizeof(Header) == kHeader, "header is 64 bytes");

struct Hasher {
    uint64_t h = 0xcbf29ce484222325ull;
    uint8_t tail[8];
    size_t nt = 0;
    void add(const void *p, size_t bytes) {
        const uint8_t *b = (const uint8_t *)p;
        while (nt && bytes) {  // finish a word begun by the previous array
            tail[nt++] = *b++;
            --bytes;
            if (nt == 8) {
                word(tail);
                nt = 0;
            }
        }
        size_t i = 0;
        for (; i + 8 <= bytes; i += 8) word(b + i);
        for (; i < bytes; ++i) tail[nt++] = b[i];
    }
    void word(const uint8_t *w) {
        uint64_t x;
        std::memcpy(&x, w, 8);
        h = (h ^ x) * 0x100000001b3ull;
    }
    uint64_t done() {
        if (nt) {
            std::memset(tail + nt, 0, 8 - nt);
            word(tail);
            nt = 0;
        }
        return h ^ (h >> 29);
    }
};

uint64_t checksum(const int32_t *rowptr, int64_t m, const int32_t *col, const float *val,
                  int64_t nnz) {
    Hasher hs;
    hs.add(rowptr, sizeof(int32_t) * (size_t)(m + 1));
    hs.add(col, sizeof(int32_t) * (size_t)nnz);
    if (val) hs.add(val, sizeof(float) * (size_t)nnz);
    return hs.done();
}

// A read-only mapping of a cache file, checked against its header.
struct Mapped {
    void *base = MAP_FAILED;
    size_t size = 0;
    Header hdr{};
    const int32_t *rowptr = nullptr, *col = nullptr;
    const float *val = nullptr;
    ~Mapped() {
        if (base != MAP_FAILED) munmap(base, size);
    }
};

int map_file(const char *path, Mapped &mp, std::string &err) {
    const int fd = open(path, O_RDONLY);
    if (fd < 0) {
        err = std::string("cannot open ") + path;
        return APSPIE_ERR_INVALID;
    }
    struct stat st {};
    if (fstat(fd, &st) != 0 || (size_t)st.st_size < kHeader) {
        close(fd);
        err = std::string(path) + ": shorter than the header";
        return APSPIE_ERR_INVALID;
    }
    mp.size = (size_t)st.st_size;
    mp.base = mmap(nullptr, mp.size, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (mp.base == MAP_FAILED) {
        err = std::string("cannot map ") + path;
        return APSPIE_ERR_INVALID;
    }
    std::memcpy(&mp.hdr, mp.base, kHeader);
    const Header &h = mp.hdr;
    if (std::memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion) {
        err = std::string(path) + ": not an apspie CSR cache (magic / version)";
        return APSPIE_ERR_INVALID;
    }
    if (h.m < 0 || h.n < 0 || h.nnz < 0 || h.m >= INT32_MAX || h.n >= INT32_MAX ||
        h.nnz >= INT32_MAX) {
        err = std::string(path) + ": sizes out of range";
        return APSPIE_ERR_INVALID;
    }
    const bool hv = h.flags & 1u;
    const size_t want = kHeader + sizeof(int32_t) * (size_t)(h.m + 1) +
                        sizeof(int32_t) * (size_t)h.nnz + (hv ? sizeof(float) * (size_t)h.nnz : 0);
    if (mp.size != want) {
        err = std::string(path) + ": file size " + std::to_string(mp.size) + " != " +
              std::to_string(want) + " from the header";
        return APSPIE_ERR_INVALID;
    }
    const uint8_t *p = (const uint8_t *)mp.base + kHeader;
    mp.rowptr = (const int32_t *)p;
    mp.col = mp.rowptr + h.m + 1;
    mp.val = hv ? (const float *)(mp.col + h.nnz) : nullptr;
    if (checksum(mp.rowptr, h.m, mp.col, mp.val, h.nnz) != h.checksum) {
        err = std::string(path) + ": checksum mismatch";
        return APSPIE_ERR_INVALID;
    }
    return APSPIE_OK;
}

int write_file(const char *path, const apspie_csr *h) {
    if (!path || !h || !h->rowptr || (h->nnz && !h->col) || h->m < 0 || h->n < 0 || h->nnz < 0)
        return APSPIE_ERR_INVALID;
    Header hdr{};
    std::memcpy(hdr.magic, kMagic, 8);
    hdr.version = kVersion;
    hdr.flags = h->val ? 1u : 0u;
    hdr.m = h->m;
    hdr.n = h->n;
    hdr.nnz = h->nnz;
    hdr.checksum = checksum(h->rowptr, h->m, h->col, h->val, h->nnz);
    // written to a temporary name and renamed, so a reader never sees half a file
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return APSPIE_ERR_INVALID;
    bool ok = std::fwrite(&hdr, kHeader, 1, f) == 1;
    ok = ok && std::fwrite(h->rowptr, sizeof(int32_t), (size_t)h->m + 1, f) == (size_t)h->m + 1;
    if (h->nnz) {
        ok = ok && std::fwrite(h->col, sizeof(int32_t), (size_t)h->nnz, f) == (size_t)h->nnz;
        if (h->val)
            ok = ok && std::fwrite(h->val, sizeof(float), (size_t)h->nnz, f) == (size_t)h->nnz;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return APSPIE_ERR_INVALID;
    }
    return APSPIE_OK;
}

}  // namespace

extern "C" {

int apspie_csr_write(const char *path, const apspie_csr *host) { return write_file(path, host); }

int apspie_csr_read(const char *path, int32_t *m, int32_t *n, int64_t *nnz, int32_t *has_val,
                    int32_t *rowptr, int32_t *col, float *val) {
    if (!path || !m || !n || !nnz) return APSPIE_ERR_INVALID;
    Mapped mp;
    std::string err;
    const int st = map_file(path, mp, err);
    if (st != APSPIE_OK) return st;
    *m = (int32_t)mp.hdr.m;
    *n = (int32_t)mp.hdr.n;
    *nnz = mp.hdr.nnz;
    if (has_val) *has_val = mp.val ? 1 : 0;
    if (rowptr) std::memcpy(rowptr, mp.rowptr, sizeof(int32_t) * (size_t)(mp.hdr.m + 1));
    if (col && mp.hdr.nnz) std::memcpy(col, mp.col, sizeof(int32_t) * (size_t)mp.hdr.nnz);
    if (val && mp.val && mp.hdr.nnz) std::memcpy(val, mp.val, sizeof(float) * (size_t)mp.hdr.nnz);
    return APSPIE_OK;
}

int apspie_graph_save(apspie_ctx *c, const apspie_graph *g, const char *path) {
    if (!c || !g || !path) return APSPIE_ERR_INVALID;
    std::vector<int32_t> rowptr((size_t)g->m + 1), col((size_t)g->nnz);
    std::vector<float> val(g->val ? (size_t)g->nnz : 0);
    int st = apspie_graph_download(c, g, rowptr.data(), col.data(), g->val ? val.data() : nullptr);
    if (st != APSPIE_OK) return st;
    const apspie_csr h{g->m, g->n, g->nnz, rowptr.data(), col.data(), g->val ? val.data() : nullptr};
    st = write_file(path, &h);
    if (st != APSPIE_OK) c->last_error = std::string("cannot write ") + path;
    return st;
}

int apspie_graph_load(apspie_ctx *c, const char *path, apspie_graph **out) {
    if (!c || !path || !out) return APSPIE_ERR_INVALID;
    *out = nullptr;
    Mapped mp;
    std::string err;
    int st = map_file(path, mp, err);
    if (st != APSPIE_OK) {
        c->last_error = err;
        return st;
    }
    const apspie_csr h{(int32_t)mp.hdr.m, (int32_t)mp.hdr.n, mp.hdr.nnz, mp.rowptr, mp.col,
                       mp.val};
    return apspie_graph_upload(c, &h, out);
}

}  // extern "C"
