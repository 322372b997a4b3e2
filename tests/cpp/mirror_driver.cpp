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

        // mXm with the reference's argument list: C into the caller's arrays
        {
            const int64_t cap = 16 * (int64_t)A.nnz;
            int *d_cp = nullptr, *d_ri = nullptr;
            float *d_cv = nullptr;
            apspie::check(apspie_dev_alloc(ctx.get(), 4 * ((int64_t)A.m + 1), (void **)&d_cp),
                          ctx.get(), "alloc");
            apspie::check(apspie_dev_alloc(ctx.get(), 4 * cap, (void **)&d_ri), ctx.get(), "alloc");
            apspie::check(apspie_dev_alloc(ctx.get(), 4 * cap, (void **)&d_cv), ctx.get(), "alloc");
            const int cn = apspie::mXm<float>(A.nnz, A.m, d_w, A.d_cscColPtr, A.d_cscRowInd, d_w,
                                              A.h_cscColPtr, A.d_cscColPtr, A.d_cscRowInd, d_cv,
                                              d_cp, d_ri, ctx);
            std::vector<int> cp(A.m + 1), ri(cn);
            std::vector<float> cv(cn);
            apspie::check(apspie_memcpy(ctx.get(), cp.data(), d_cp, 4 * ((int64_t)A.m + 1), 1),
                          ctx.get(), "copy");
            apspie::check(apspie_memcpy(ctx.get(), ri.data(), d_ri, 4 * (int64_t)cn, 1), ctx.get(),
                          "copy");
            apspie::check(apspie_memcpy(ctx.get(), cv.data(), d_cv, 4 * (int64_t)cn, 1), ctx.get(),
                          "copy");
            std::printf("mxm3_colptr");
            for (int v : cp) std::printf(" %d", v);
            std::printf("\nmxm3_rows");
            for (int v : ri) std::printf(" %d", v);
            std::printf("\nmxm3_vals");
            for (float v : cv) std::printf(" %.9g", v);
            std::printf("\n");
            apspie_dev_free(ctx.get(), d_cp, 4 * ((int64_t)A.m + 1));
            apspie_dev_free(ctx.get(), d_ri, 4 * cap);
            apspie_dev_free(ctx.get(), d_cv, 4 * cap);
        }

        // allocScratch (util.cuh:646) + the two mXv forms (mXv.cuh:110, :249)
        {
            apspie::d_scratch *d = nullptr;
            apspie::allocScratch(&d, A.nnz, A.m);
            // dense x: every fifth vertex in the frontier with value 1..3,
            // 0.5 and 0.25 (not above 0.5) elsewhere in a pattern
            std::vector<float> x(A.m), y(A.m);
            for (int i = 0; i < A.m; ++i)
                x[i] = i % 5 == 0 ? (float)(i % 3 + 1) : i % 7 == 1 ? 0.5f : i % 7 == 2 ? 0.25f : 0;
            float *d_x = nullptr, *d_y = nullptr;
            apspie::check(apspie_dev_alloc(ctx.get(), 4 * (int64_t)A.m, (void **)&d_x), ctx.get(),
                          "alloc");
            apspie::check(apspie_dev_alloc(ctx.get(), 4 * (int64_t)A.m, (void **)&d_y), ctx.get(),
                          "alloc");
            apspie::check(apspie_memcpy(ctx.get(), d_x, x.data(), 4 * (int64_t)A.m, 0), ctx.get(),
                          "copy");
            const int total = apspie::mXv<float>(d_x, A.nnz, A.m, d_w, A.d_cscColPtr,
                                                 A.d_cscRowInd, d_y, d, ctx);
            apspie::check(apspie_memcpy(ctx.get(), y.data(), d_y, 4 * (int64_t)A.m, 1), ctx.get(),
                          "copy");
            std::printf("mxv_dense_total %d\nmxv_dense_y", total);
            for (float v : y) std::printf(" %.9g", v);
            std::printf("\n");
            // no entry above 0.5: y keeps what it held, 0 is returned
            for (int i = 0; i < A.m; ++i) x[i] = i % 2 ? 0.5f : 0.0f;
            apspie::check(apspie_memcpy(ctx.get(), d_x, x.data(), 4 * (int64_t)A.m, 0), ctx.get(),
                          "copy");
            const int none = apspie::mXv<float>(d_x, A.nnz, A.m, d_w, A.d_cscColPtr,
                                                A.d_cscRowInd, d_y, d, ctx);
            std::vector<float> y2(A.m);
            apspie::check(apspie_memcpy(ctx.get(), y2.data(), d_y, 4 * (int64_t)A.m, 1), ctx.get(),
                          "copy");
            std::printf("mxv_dense_empty %d %d\n", none, (int)(y2 == y));

            // sparse x: the source's row with its weights (a column of A)
            const int r0 = A.h_cscColPtr[source], r1 = A.h_cscColPtr[source + 1];
            int nnz_x = r1 - r0;
            int *d_yi = nullptr;
            float *d_yv = nullptr;
            apspie::check(apspie_dev_alloc(ctx.get(), 4 * (int64_t)A.m, (void **)&d_yi), ctx.get(),
                          "alloc");
            apspie::check(apspie_dev_alloc(ctx.get(), 4 * (int64_t)A.m, (void **)&d_yv), ctx.get(),
                          "alloc");
            int nnz = nnz_x;
            apspie::mXv<float>(A.d_cscRowInd + r0, d_w + r0, A.nnz, A.m, nnz, d_w,
                               A.d_cscColPtr, A.d_cscRowInd, d_yi, d_yv, d, ctx);
            std::vector<int> yi(nnz);
            std::vector<float> yv(nnz);
            apspie::check(apspie_memcpy(ctx.get(), yi.data(), d_yi, 4 * (int64_t)nnz, 1), ctx.get(),
                          "copy");
            apspie::check(apspie_memcpy(ctx.get(), yv.data(), d_yv, 4 * (int64_t)nnz, 1), ctx.get(),
                          "copy");
            std::printf("mxv_sparse_nnz %d\nmxv_sparse_ind", nnz);
            for (int v : yi) std::printf(" %d", v);
            std::printf("\nmxv_sparse_val");
            for (float v : yv) std::printf(" %.9g", v);
            std::printf("\n");
            int zero = 0;  // nnz == 0 returns at once
            apspie::mXv<float>(A.d_cscRowInd, d_w, A.nnz, A.m, zero, d_w, A.d_cscColPtr,
                               A.d_cscRowInd, d_yi, d_yv, d, ctx);
            std::printf("mxv_sparse_zero %d\n", zero);
            apspie_dev_free(ctx.get(), d_x, 4 * (int64_t)A.m);
            apspie_dev_free(ctx.get(), d_y, 4 * (int64_t)A.m);
            apspie_dev_free(ctx.get(), d_yi, 4 * (int64_t)A.m);
            apspie_dev_free(ctx.get(), d_yv, 4 * (int64_t)A.m);
            apspie::freeScratch(d);
        }

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
