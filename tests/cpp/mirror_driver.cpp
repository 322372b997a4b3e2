// A reference-style driver written against include/apspie.hpp: it follows
// the flow of the reference's test/<alg>/main.cu programs (read .mtx ->
// symmetrize -> buildMatrix -> device arrays -> bfs / sssp / spmspvMis /
// TC / spgemm) using the reference's entry-point names, and prints the
// results so tests/test_gpu_mirror.py can compare them with the oracle.
//
//   mirror_driver <file.mtx> <source>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "apspie.hpp"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s file.mtx source\n", argv[0]);
        return 2;
    }
    const int source = std::atoi(argv[2]);
    try {
        apspie::Context ctx(0);
        // readEdge / readMtx (util.cuh:354-430)
        int32_t m = 0, n = 0;
        int64_t nnz = 0;
        apspie::check(apspie_read_mtx(argv[1], &m, &n, &nnz, nullptr, nullptr, nullptr), nullptr,
                      "read_mtx");
        std::vector<int> row(nnz), col(nnz);
        std::vector<float> val(nnz);
        apspie::check(apspie_read_mtx(argv[1], &m, &n, &nnz, row.data(), col.data(), val.data()),
                      nullptr, "read_mtx");
        const int nv = m > n ? m : n;
        // -undirected: makeSymmetric + buildMatrix
        apspie::d_matrix A;
        apspie::matrix_new(&A, nv, nv);
        apspie::buildSymmetricMatrix<float>(&A, (int)nnz, row.data(), col.data(), val.data(), ctx);

        // integer weights 1 + mt19937(2) % 255 per pair u < v, mirrored
        std::vector<int> hcol(A.nnz);
        apspie::check(apspie_graph_download(ctx.get(), A.g, nullptr, hcol.data(), nullptr),
                      ctx.get(), "download");
        std::vector<float> w(A.nnz);
        apspie::check(apspie_gen_weights(A.m, A.h_cscColPtr, hcol.data(), 2, w.data()), nullptr,
                      "weights");
        float *d_w = nullptr;
        int *d_out = nullptr;
        float *d_dist = nullptr;
        uint32_t *d_prio = nullptr;
        apspie::check(apspie_dev_alloc(ctx.get(), 4 * (int64_t)A.nnz, (void **)&d_w), ctx.get(), "alloc");
        apspie::check(apspie_dev_alloc(ctx.get(), 4 * (int64_t)A.m, (void **)&d_out), ctx.get(), "alloc");
        apspie::check(apspie_dev_alloc(ctx.get(), 4 * (int64_t)A.m, (void **)&d_dist), ctx.get(), "alloc");
        apspie::check(apspie_dev_alloc(ctx.get(), 4 * (int64_t)A.m, (void **)&d_prio), ctx.get(), "alloc");
        apspie::check(apspie_memcpy(ctx.get(), d_w, w.data(), 4 * (int64_t)A.nnz, 0), ctx.get(), "copy");

        std::vector<int> h(A.m);
        auto dump_int = [&](const char *tag, const int *d) {
            apspie::check(apspie_memcpy(ctx.get(), h.data(), d, 4 * (int64_t)A.m, 1), ctx.get(), "copy");
            std::printf("%s", tag);
            for (int i = 0; i < A.m; ++i) std::printf(" %d", h[i]);
            std::printf("\n");
        };

        apspie::bfs<float>(source, A.nnz, A.m, nullptr, A.d_cscColPtr, A.d_cscRowInd, d_out, 0, ctx);
        dump_int("bfs", d_out);
        apspie::spmspvBfs(source, A.nnz, A.m, A.h_cscColPtr, A.d_cscColPtr, A.d_cscRowInd, d_out,
                          0, ctx);
        dump_int("spmspvBfs", d_out);

        apspie::sssp<float>(source, A.nnz, A.m, d_w, A.d_cscColPtr, A.d_cscRowInd, d_dist, 0, ctx);
        std::vector<float> hd(A.m);
        apspie::check(apspie_memcpy(ctx.get(), hd.data(), d_dist, 4 * (int64_t)A.m, 1), ctx.get(), "copy");
        std::printf("sssp");
        for (int i = 0; i < A.m; ++i) std::printf(" %.9g", hd[i]);
        std::printf("\n");

        std::vector<uint32_t> prio(A.m);
        std::mt19937 gen(5489u);
        for (auto &p : prio) p = gen();
        apspie::check(apspie_memcpy(ctx.get(), d_prio, prio.data(), 4 * (int64_t)A.m, 0), ctx.get(), "copy");
        apspie::spmspvMis<uint32_t>(A.nnz, A.m, A.h_cscColPtr, A.d_cscColPtr, A.d_cscRowInd, d_prio,
                                    d_out, 0.1f, ctx);
        dump_int("mis", d_out);

        const uint64_t tri = apspie::tc(A.nnz, A.m, A.d_cscColPtr, A.d_cscRowInd, d_out, ctx);
        std::printf("tc_total %llu\n", (unsigned long long)tri);
        dump_int("tc_rows", d_out);

        apspie::d_matrix C;
        apspie::matrix_new(&C, A.m, A.m);
        apspie::spgemm(&C, &A, &A, 0, 0, ctx);
        std::printf("spgemm_nnz %d\n", C.nnz);
        apspie::matrix_delete(&C);

        // mXm (tc.cuh:314): the weighted matrix times itself, column by column
        apspie::d_matrix D;
        apspie::matrix_new(&D, A.m, A.m);
        const int dn = apspie::mXm<float>(A.nnz, A.m, d_w, A.d_cscColPtr, A.d_cscRowInd, d_w,
                                          A.h_cscColPtr, A.d_cscColPtr, A.d_cscRowInd, &D, ctx);
        std::vector<int> dcol(dn);
        std::vector<float> dval(dn);
        apspie::check(apspie_memcpy(ctx.get(), dcol.data(), D.d_cscRowInd, 4 * (int64_t)dn, 1),
                      ctx.get(), "copy");
        apspie::check(apspie_memcpy(ctx.get(), dval.data(), D.d_cscVal, 4 * (int64_t)dn, 1),
                      ctx.get(), "copy");
        std::printf("mxm_colptr");
        for (int i = 0; i <= A.m; ++i) std::printf(" %d", D.h_cscColPtr[i]);
        std::printf("\nmxm_rows");
        for (int i = 0; i < dn; ++i) std::printf(" %d", dcol[i]);
        std::printf("\nmxm_vals");
        for (int i = 0; i < dn; ++i) std::printf(" %.9g", dval[i]);
        std::printf("\n");
        apspie::matrix_delete(&D);

        // csr_to_dcsc (spgemm.cuh:138-139 call pattern) + copy_part
        const int pn = A.m < 4 ? A.m : 4, ps = A.m / pn;
        apspie::csr_to_dcsc<float>(&A, ps, pn, ctx, true);
        apspie::csr_to_dcsc<float>(&A, ps, pn, ctx);
        apspie::copy_part(&A);
        std::vector<int> cind(A.col_length), coff(A.col_length + 1), rind(A.nnz);
        apspie::check(apspie_memcpy(ctx.get(), cind.data(), A.d_dcscColPtr_ind,
                                    4 * (int64_t)A.col_length, 1), ctx.get(), "copy");
        apspie::check(apspie_memcpy(ctx.get(), coff.data(), A.d_dcscColPtr_off,
                                    4 * ((int64_t)A.col_length + 1), 1), ctx.get(), "copy");
        apspie::check(apspie_memcpy(ctx.get(), rind.data(), A.d_dcscRowInd, 4 * (int64_t)A.nnz, 1),
                      ctx.get(), "copy");
        std::printf("dcsc_parts %d %d\ndcsc_part_ptr", ps, pn);
        for (int i = 0; i <= pn; ++i) std::printf(" %d", A.h_dcscPartPtr[i]);
        std::printf("\ndcsc_col_ind");
        for (int v : cind) std::printf(" %d", v);
        std::printf("\ndcsc_col_off");
        for (int v : coff) std::printf(" %d", v);
        std::printf("\ndcsc_row_ind");
        for (int v : rind) std::printf(" %d", v);
        std::printf("\n");

        apspie_dev_free(ctx.get(), d_w, 4 * (int64_t)A.nnz);
        apspie_dev_free(ctx.get(), d_out, 4 * (int64_t)A.m);
        apspie_dev_free(ctx.get(), d_dist, 4 * (int64_t)A.m);
        apspie_dev_free(ctx.get(), d_prio, 4 * (int64_t)A.m);
        apspie::matrix_delete(&A);
    } catch (const apspie::Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
