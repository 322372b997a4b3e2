// apspie_bfs_noalloc: the BFS entry point for callers that cannot allocate
// device memory (code under HIP graph capture elsewhere in the process,
// serving loops, harnesses that forbid hipMalloc).  Its frontier, bitmaps and
// control block live in one static __device__ array per GPU instead of the
// context's scratch pool; the kernels are those of apspie_bfs.
//
// The array is a host-side stand-in in host code: its device address is taken
// once per device with hipGetSymbolAddress (at context creation, with that
// device current) and only that address is passed on.  Every call on one GPU
// shares that GPU's array; a host mutex per device serialises the calls of
// this process (each call is synchronous), and what the array held from the
// previous call is overwritten by the next call's initialisation kernel.
#include <mutex>
#include <string>

#include "common.hpp"

namespace apspie_noalloc {

// 96 MiB: R-MAT scale 22 (4,194,304 vertices, 128M entries) fits with the
// depths staged for a host output (about 84 MB), and with device output
// (about 70 MB)
constexpr int64_t kScratchBytes = int64_t(96) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceScratch {
    std::once_flag once;
    void *ptr = nullptr;  // device address of `scratch` on this GPU
    hipError_t err = hipSuccess;
    std::mutex run;  // one call at a time on this GPU (per process)
};

DeviceScratch g_dev[kMaxDevices];

}  // namespace apspie_noalloc

namespace apspie {

// Called by apspie_ctx_create with `device` current: looks the array's
// address up once per device.  A failed lookup is kept and reported by the
// entry point; the other entry points do not depend on it.
void noalloc_init(int device) {
    using namespace apspie_noalloc;
    if (device < 0 || device >= kMaxDevices) return;
    DeviceScratch &d = g_dev[device];
    std::call_once(d.once, [&] {
        void *p = nullptr;
        d.err = hipGetSymbolAddress(&p, HIP_SYMBOL(apspie_noalloc::scratch));
        d.ptr = d.err == hipSuccess ? p : nullptr;
    });
}

}  // namespace apspie

using namespace apspie;

extern "C" int apspie_bfs_noalloc(apspie_ctx *c, const apspie_graph *g, int32_t source,
                                  int32_t *depth, int mem) {
    using apspie_noalloc::g_dev;
    using apspie_noalloc::kMaxDevices;
    using apspie_noalloc::kScratchBytes;
    if (!c || !depth) return APSPIE_ERR_INVALID;
    return guard(c, [&] {
        APSPIE_CHECK(g, APSPIE_ERR_INVALID, "null graph");
        APSPIE_CHECK(g->device == c->device, APSPIE_ERR_INVALID, "graph lives on another device");
        APSPIE_CHECK(source >= 0 && source < g->m, APSPIE_ERR_INVALID, "source out of range");
        APSPIE_CHECK(g->m == g->n, APSPIE_ERR_INVALID, "BFS needs a square matrix");
        APSPIE_CHECK(mem == APSPIE_MEM_HOST || mem == APSPIE_MEM_DEVICE, APSPIE_ERR_INVALID,
                     "mem must be APSPIE_MEM_HOST or APSPIE_MEM_DEVICE");
        APSPIE_CHECK(c->device >= 0 && c->device < kMaxDevices, APSPIE_ERR_UNSUPPORTED,
                     "device index beyond the no-alloc table");
        set_device(c);
        apspie_noalloc::DeviceScratch &d = g_dev[c->device];
        APSPIE_CHECK(d.ptr, APSPIE_ERR_DEVICE,
                     std::string("no-alloc scratch has no device address: hipGetSymbolAddress: ") +
                         hipGetErrorString(d.err));
        std::lock_guard<std::mutex> lock(d.run);
        // a host output is staged in the array's first bytes
        const int64_t stage = mem == APSPIE_MEM_HOST ? ((int64_t)g->m * 4 + 255) / 256 * 256 : 0;
        const int64_t need = stage + bfs_block_bytes(c, g);
        APSPIE_CHECK(need <= kScratchBytes, APSPIE_ERR_OVERFLOW,
                     "graph too large for the no-alloc BFS: needs " + std::to_string(need) +
                         " bytes of scratch, the static array holds " +
                         std::to_string(kScratchBytes) + " (use apspie_bfs)");
        char *base = (char *)d.ptr;
        int32_t *out = mem == APSPIE_MEM_HOST ? (int32_t *)base : depth;
        timed(c, [&] { bfs_run_in(c, g, source, out, base + stage, kScratchBytes - stage); });
        if (mem == APSPIE_MEM_HOST) {
            copy_from_device(c, depth, out, sizeof(int32_t) * g->m, APSPIE_MEM_HOST);
            sync(c);
        }
    });
}
