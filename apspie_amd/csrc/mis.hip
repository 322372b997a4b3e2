// Maximal independent set by Luby rounds under the max-priority semiring.
//
// Semantics follow spmspvMis() (spmspvMis.cuh:149-344): every round, each
// undecided candidate compares its priority with the max priority over its
// undecided neighbours (the max-priority SpMSpV y = A^T x, x = candidate
// priorities, :277-289); candidates that beat all neighbours join
// (updateMis :63-68) and their neighbours leave (updateNeighbor :70-78).
// Priorities are uint32 (default: mt19937 seeded 5489, the default seed of
// mt19937ar.cpp:111-112) with ties broken to the lower vertex id, i.e. the
// key (prio << 32) | (0x7fffffff - v): a total order, so the rounds converge
// to the greedy MIS in descending key order, whatever the schedule.  The
// reference's threshold schedule (minimum -= delta) only changes the number
// of rounds, not the set, and is not used.
//
// Per round (no host round trip): k_mis_lose (first round: pulled over the
// candidate rows) or k_expand (later rounds: pushed over the candidate
// edges) marks candidates that lose to an undecided neighbour, k_build
// admits the winners and lists them, k_expand pushes from the winners to
// remove their undecided neighbours, and k_build compacts the survivors into
// the next candidate frontier.
#include <random>

#include "common.hpp"
#include "device.hpp"
#include "frontier.hpp"
#include "timing.hpp"

namespace apspie {

namespace {

constexpr int kBlock = 256;
constexpr int kTile = 2048;

__device__ __forceinline__ uint64_t mis_key(const uint32_t *prio, int32_t v) {
    return (uint64_t(prio[v]) << 32) | uint64_t(0x7fffffffu - (uint32_t)v);
}

// live[v]: v's key while v is undecided, 0 once decided (a key is never 0:
// its low word is 0x7fffffff - v and v < 2^31 - 1), so one 8-byte load per
// edge answers "undecided with a larger key"

// a candidate that did not lose joins (the join list is built from these)
struct JoinOp {
    const int32_t *lose;
    int32_t *state;
    uint64_t *live;
    int32_t round;
    __device__ __forceinline__ bool keep(int32_t v) const {
        if (lose[v] == round) return false;
        state[v] = 1;
        live[v] = 0;
        return true;
    }
};

// the undecided neighbours of a vertex that joined leave (no neighbour of a
// joiner joined in the same round: keys are a strict total order)
struct RemoveOp {
    int32_t *state;
    uint64_t *live;
    __device__ __forceinline__ bool edge(int32_t, int32_t, int32_t v) const {
        if (state[v] == -1) {
            state[v] = 0;
            live[v] = 0;
        }
        return false;
    }
};

// The lose test pulled over each candidate's row, stopping at the first
// undecided neighbour with a larger key (with random keys, after about two
// entries for most candidates), where pushing every candidate edge read all
// of them: a lane per candidate takes the first kLoseQ entries with their
// loads in flight; rows still open after them go to an LDS list that the
// block's waves then finish a row at a time, 256 entries per step.
constexpr int kLoseQ = 8;
constexpr int kLoseStep = 256;

__global__ __launch_bounds__(kBlock) void k_mis_lose(const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col,
                                                     const uint64_t *__restrict__ live,
                                                     int32_t *lose, int32_t round,
                                                     FrontierBufs cur, const uint64_t *cnt_cur,
                                                     uint64_t *cnt_clear, uint64_t *stats) {
    __shared__ int32_t s_u[kBlock], s_e[kBlock], s_end[kBlock];
    __shared__ int32_t s_n;
    const uint64_t cc = *cnt_cur;
    const int32_t F = (int32_t)cnt_entries(cc);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *cnt_clear = 0;
        if (F > 0) {
            stats[0] += (uint64_t)cnt_edges(cc);
            stats[1] += (uint64_t)F;
            stats[2] += 1;
        }
    }
    const int lane = dev::lane_id(), wave = threadIdx.x / 64;
    for (int64_t base = (int64_t)blockIdx.x * kBlock; base < F;
         base += (int64_t)gridDim.x * kBlock) {
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
        const int64_t j = base + threadIdx.x;
        if (j < F) {
            const int32_t u = cur.v[j], rs = cur.start[j], re = rowptr[u + 1];
            const uint64_t ku = live[u];
            int32_t c[kLoseQ];
#pragma unroll
            for (int t = 0; t < kLoseQ; ++t) c[t] = rs + t < re ? col[rs + t] : -1;
            bool lost = false;
#pragma unroll
            for (int t = 0; t < kLoseQ; ++t)
                if (c[t] >= 0 && live[c[t]] > ku) lost = true;
            if (lost) lose[u] = round;
            else if (re - rs > kLoseQ) {
                const int32_t q = atomicAdd(&s_n, 1);
                s_u[q] = u;
                s_e[q] = rs + kLoseQ;
                s_end[q] = re;
            }
        }
        __syncthreads();
        const int32_t n = s_n;
        for (int32_t q = wave; q < n; q += kBlock / 64) {
            const int32_t u = s_u[q], re = s_end[q];
            const uint64_t ku = live[u];
            bool lost = false;
            for (int32_t e0 = s_e[q]; e0 < re && !lost; e0 += kLoseStep) {
                bool mine = false;
#pragma unroll
                for (int t = 0; t < kLoseStep / 64; ++t) {
                    const int32_t e = e0 + t * 64 + lane;
                    if (e < re && live[col[e]] > ku) mine = true;
                }
                lost = __ballot(mine) != 0;
            }
            if (lost && lane == 0) lose[u] = round;
        }
        __syncthreads();  // s_n and the list are rewritten by the next pass
    }
}

// candidate u loses if some undecided neighbour has a larger key (pushed
// over the candidate edges, load-balanced by tiles: the later rounds' few
// candidates are the ones with long rows of mostly decided neighbours,
// which a pulled row would walk to its end)
struct LoseOp {
    const uint64_t *live;
    int32_t *lose;
    int32_t round;
    __device__ __forceinline__ bool edge(int32_t u, int32_t, int32_t v) const {
        if (v == u || lose[u] == round) return false;
        if (live[v] > live[u]) lose[u] = round;
        return false;
    }
};

struct UndecidedOp {
    const int32_t *state;
    __device__ __forceinline__ bool keep(int32_t v) const { return state[v] == -1; }
};

__global__ void k_mis_init(const int32_t *rowptr, const uint32_t *prio, int32_t m,
                           int32_t *state, uint64_t *live, int32_t *lose, uint64_t *ctl) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < m;
         v += (int64_t)gridDim.x * blockDim.x) {
        // a vertex without neighbours is in every maximal independent set
        const bool iso = rowptr[v + 1] == rowptr[v];
        state[v] = iso ? 1 : -1;
        live[v] = iso ? 0 : mis_key(prio, (int32_t)v);
        lose[v] = -1;
    }
    if (blockIdx.x == 0 && threadIdx.x < 16) ctl[threadIdx.x] = 0;
}

}  // namespace

void mis_run(Ctx *c, const Graph *g, const uint32_t *prio, int32_t *state) {
    if (g->m == 0) return;
    FrontierSet fs = frontier_alloc(c, "mis", g->m, g->nnz, kTile, true);
    // the vertices that join in a round, expanded to remove their neighbours
    FrontierSet js = frontier_alloc(c, "mis_join", g->m, g->nnz, kTile, true);
    uint64_t *jcnt = js.ctl;
    int32_t *lose = (int32_t *)scratch(c, "mis_lose", sizeof(int32_t) * g->m);
    uint64_t *live = (uint64_t *)scratch(c, "mis_live", sizeof(uint64_t) * g->m);
    const int grid = fill_grid(c, kBlock, 4);
    k_mis_init<<<grid, kBlock, 0, c->stream>>>(g->rowptr, prio, g->m, state, live, lose, fs.ctl);
    APSPIE_LAUNCH_CHECK();
    // round 0 candidates: every undecided vertex (degree > 0)
    k_build<kBlock, 8, kTile, true><<<grid, kBlock, 0, c->stream>>>(
        UndecidedOp{state}, g->rowptr, nullptr, g->m, nullptr, fs.f[0], fs.ctl + 0);
    APSPIE_LAUNCH_CHECK();
    KernelTimer kt(c);
    int round = 0;
    int chunk = 8;  // rounds between host checks: R-MAT graphs finish in 6-7
    int grow = 4;   // then 4, 8, 16, ... (an idle round still costs four launches)
    for (;;) {
        for (int k = 0; k < chunk; ++k, ++round) {
            const int cur = round & 1;
            uint64_t *cnt = fs.ctl + (round % 3), *nxt = fs.ctl + ((round + 1) % 3);
            kt.begin();
            // lose marks, pulled in the first round (every vertex a candidate,
            // random keys: most rows stop after a few entries; R-MAT-20: 173
            // -> 53 us) and pushed after it (84 / 47 / 36 us pulled against
            // 50 / 17 / 15 pushed in rounds 2-4); this launch also clears the
            // join list's counter
            if (round == 0)
                k_mis_lose<<<grid, kBlock, 0, c->stream>>>(g->rowptr, g->col, live, lose, round,
                                                          fs.f[cur], cnt, jcnt, fs.ctl + 4);
            else
                k_expand<kBlock, kTile, true, false, false><<<grid, kBlock, 0, c->stream>>>(
                    LoseOp{live, lose, round}, g->rowptr, g->col, (int32_t)g->nnz, fs.f[cur],
                    cnt, fs.f[cur ^ 1], nullptr, jcnt, fs.ctl + 4);
            kt.end();
            k_build<kBlock, 8, kTile, true><<<grid, kBlock, 0, c->stream>>>(
                JoinOp{lose, state, live, round}, g->rowptr, fs.f[cur].v, 0, cnt, js.f[0], jcnt);
            // removal pushes from the joiners (few, mostly short rows) instead of
            // scanning every candidate's row; this launch clears the next
            // round's counter
            k_expand<kBlock, kTile, true, false, false><<<grid, kBlock, 0, c->stream>>>(
                RemoveOp{state, live}, g->rowptr, g->col, (int32_t)g->nnz, js.f[0], jcnt, js.f[1],
                nullptr, nxt, nullptr);
            k_build<kBlock, 8, kTile, true><<<grid, kBlock, 0, c->stream>>>(
                UndecidedOp{state}, g->rowptr, fs.f[cur].v, 0, cnt, fs.f[cur ^ 1], nxt);
            APSPIE_LAUNCH_CHECK();
        }
        uint64_t *h = (uint64_t *)c->pinned;
        APSPIE_HIP(hipMemcpyAsync(h, fs.ctl + (round % 3), sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  c->stream));
        APSPIE_HIP(hipMemcpyAsync(h + 1, fs.ctl + 4, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  c->stream));
        sync(c);
        c->stats.host_syncs++;
        if (cnt_entries(h[0]) == 0) break;
        // the undecided vertex with the largest key joins every round
        APSPIE_CHECK(round <= g->m + 1, APSPIE_ERR_DEVICE, "MIS did not converge");
        chunk = grow;
        grow = std::min(grow * 2, 64);
    }
    uint64_t *h = (uint64_t *)c->pinned;
    c->stats.edges_traversed = (int64_t)h[1];
    c->stats.vertices_reached = (int64_t)h[2];
    c->stats.iterations = (int32_t)h[3];
    kt.collect();
}

// mt19937(seed) stream, one draw per vertex: generated on the host (the
// generator is sequential by definition) and copied once.
void mt19937_priorities(Ctx *c, uint32_t seed, int32_t m, uint32_t *out) {
    std::vector<uint32_t> h((size_t)std::max(m, 1));
    std::mt19937 gen(seed);
    for (int32_t i = 0; i < m; ++i) h[i] = gen();
    APSPIE_HIP(hipMemcpyAsync(out, h.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice,
                              c->stream));
    sync(c);
}

}  // namespace apspie
