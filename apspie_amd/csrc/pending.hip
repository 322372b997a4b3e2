// Entry points whose kernels are not in this build yet: they fail loudly.
#include "common.hpp"

namespace apspie {

#define APSPIE_PENDING(what) throw Error{APSPIE_ERR_UNSUPPORTED, what " is not built yet"}

void sssp_run(Ctx *, const Graph *, int32_t, float *) { APSPIE_PENDING("sssp"); }
void mis_run(Ctx *, const Graph *, const uint32_t *, int32_t *) { APSPIE_PENDING("mis"); }
void tc_run(Ctx *, const Graph *, uint64_t *, int32_t *) { APSPIE_PENDING("tc"); }
int64_t spmspv_run(Ctx *, const Graph *, int, int64_t, const int32_t *, const float *,
                   const int32_t *, int32_t *, float *, int) {
    APSPIE_PENDING("spmspv");
}
Graph *spgemm_run(Ctx *, const Graph *, const Graph *) { APSPIE_PENDING("spgemm"); }
int64_t scan_exclusive_dev(Ctx *, const int32_t *, int32_t *, int64_t) { APSPIE_PENDING("scan"); }
void sort_pairs_dev(Ctx *, uint32_t *, uint32_t *, int64_t, int) { APSPIE_PENDING("sort"); }
void bfs_batch_run(Ctx *, const Graph *, const int32_t *, int32_t, int32_t *) {
    APSPIE_PENDING("bfs_batch");
}
void sssp_batch_run(Ctx *, const Graph *, const int32_t *, int32_t, float *) {
    APSPIE_PENDING("sssp_batch");
}
void graph_from_coo_dev(Ctx *, Graph *, const int32_t *, const int32_t *, const float *, int64_t,
                        int) {
    APSPIE_PENDING("graph_from_coo");
}
void graph_transpose_dev(Ctx *, const Graph *, Graph *) { APSPIE_PENDING("transpose"); }
void mt19937_priorities(Ctx *, uint32_t, int32_t, uint32_t *) { APSPIE_PENDING("mt19937"); }

}  // namespace apspie

extern "C" {
int apspie_autotune(apspie_ctx *, const apspie_graph *, int32_t) { return APSPIE_ERR_UNSUPPORTED; }
int apspie_server_create(int, const apspie_server_config *, apspie_server **) {
    return APSPIE_ERR_UNSUPPORTED;
}
int apspie_server_destroy(apspie_server *) { return APSPIE_ERR_UNSUPPORTED; }
int apspie_server_add_graph(apspie_server *, const apspie_csr *, int32_t *) {
    return APSPIE_ERR_UNSUPPORTED;
}
int apspie_server_submit(apspie_server *, int32_t, int, int32_t, int64_t *) {
    return APSPIE_ERR_UNSUPPORTED;
}
int apspie_server_wait(apspie_server *, int64_t, void *, int64_t, double *) {
    return APSPIE_ERR_UNSUPPORTED;
}
int apspie_server_drain(apspie_server *) { return APSPIE_ERR_UNSUPPORTED; }
}
