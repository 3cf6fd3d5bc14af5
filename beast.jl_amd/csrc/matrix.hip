// Resident system matrices: assembly into HBM, transfers, and (lu.hip) factorisation/solve.
#include <cstring>

#include "bh_internal.h"
#include "bh_matrix.h"

namespace bh {

// transpose-copy between the host column-major layout and the row-major HBM layout
__global__ void k_colmajor_to_rowmajor(const double2 *__restrict__ src, long long ld, double2 *__restrict__ dst,
                                       long long n)
{
    __shared__ double2 tile[32][33];
    const long long c0 = (long long)blockIdx.x * 32, r0 = (long long)blockIdx.y * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;   // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const long long r = r0 + tx, c = c0 + k;   // read column c, row r (contiguous in r)
        if (r < n && c < n) tile[k][tx] = src[r + c * ld];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const long long r = r0 + k, c = c0 + tx;   // write row r, column c (contiguous in c)
        if (r < n && c < n) dst[r * n + c] = tile[tx][k];
    }
}

__global__ void k_rowmajor_to_colmajor(const double2 *__restrict__ src, double2 *__restrict__ dst, long long ld,
                                       long long n)
{
    __shared__ double2 tile[32][33];
    const long long c0 = (long long)blockIdx.x * 32, r0 = (long long)blockIdx.y * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int k = ty; k < 32; k += 8) {
        const long long r = r0 + k, c = c0 + tx;
        if (r < n && c < n) tile[k][tx] = src[r * n + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const long long r = r0 + tx, c = c0 + k;
        if (r < n && c < n) dst[r + c * ld] = tile[tx][k];
    }
}

}  // namespace bh

using namespace bh;

extern "C" int bh_matrix_assemble(bh_ctx *ctx, const bh_operator *op, const bh_quadstrat *qs, const bh_space *X,
                                  bh_matrix **out)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx && X && out, "NULL argument");
    BH_REQUIRE(op != nullptr, "operator is NULL");
    *out = nullptr;
    if (op->kind != BH_OP_NCROSS && op->kind != BH_OP_IDENTITY) {
        // a batch of one: the same kernels as bh_matrix_assemble_batch, the same bits
        assemble_resident_batch(ctx, 1, op, qs, &X, out);
        return BH_OK;
    }
    std::unique_ptr<bh_matrix> M(new bh_matrix());
    M->ctx = ctx;
    M->n = X->ndofs;
    pool_alloc(*ctx, M->a64, (size_t)M->n * M->n * 2);
    bh_space *sp = const_cast<bh_space *>(X);
    // the tile kernel also writes the FP32 copy that an FP32 factorisation starts from
    const long long lda = padded_ld(M->n);
    if (M->n > 0) pool_alloc(*ctx, M->lu32, (size_t)lda * M->n * 2);
    M->lu32_fresh = assemble_device(*ctx, op, qs, false, *sp, *sp, M->a64.p, M->n, true, false, 1.0, 0.0,
                                    M->n > 0 ? M->lu32.p : nullptr, lda);
    if (!M->lu32_fresh) pool_release(*ctx, M->lu32);
    ctx_retain(ctx);
    *out = M.release();
    BH_API_END
}

extern "C" int bh_matrix_accumulate(bh_matrix *M, const bh_operator *op, const bh_quadstrat *qs, const bh_space *X,
                                    const double *scale)
{
    BH_API_BEGIN
    BH_REQUIRE(M && op && X, "NULL argument");
    BH_REQUIRE(X->ndofs == M->n, "space has %lld dofs, matrix is %lld x %lld", (long long)X->ndofs, (long long)M->n,
               (long long)M->n);
    BH_REQUIRE(X->mesh->ctx == M->ctx, "space belongs to another context");
    std::lock_guard<std::mutex> lock(M->mu);
    bh_space *sp = const_cast<bh_space *>(X);
    const double sr = scale ? scale[0] : 1.0, si = scale ? scale[1] : 0.0;
    assemble_device(*M->ctx, op, qs, false, *sp, *sp, M->a64.p, M->n, true, true, sr, si);
    M->factored = -1;   // earlier factors describe the old matrix
    M->lu32_fresh = false;
    BH_API_END
}

extern "C" int bh_matrix_from_host(bh_ctx *ctx, int64_t n, const double *A, int64_t ld, bh_matrix **out)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx && A && out, "NULL argument");
    BH_REQUIRE(n > 0 && ld >= n, "bad size / leading dimension");
    *out = nullptr;
    std::unique_ptr<bh_matrix> M(new bh_matrix());
    M->ctx = ctx;
    M->n = n;
    pool_alloc(*ctx, M->a64, (size_t)n * n * 2);
    // only the n x n block is read: rows [n, ld) of each caller column are padding
    DevBuf<double> tmp((size_t)n * n * 2);
    BH_CHECK_HIP(hipMemcpy2DAsync(tmp.p, sizeof(double) * 2 * n, A, sizeof(double) * 2 * ld, sizeof(double) * 2 * n,
                                  n, hipMemcpyHostToDevice, ctx->stream));
    dim3 grid((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32)), block(32, 8);
    k_colmajor_to_rowmajor<<<grid, block, 0, ctx->stream>>>((const double2 *)tmp.p, n, (double2 *)M->a64.p, n);
    BH_CHECK_HIP(hipGetLastError());
    BH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    ctx_retain(ctx);
    *out = M.release();
    BH_API_END
}

extern "C" int bh_matrix_destroy(bh_matrix *M)
{
    BH_API_BEGIN
    if (M) {
        Ctx *c = M->ctx;
        (void)hipStreamSynchronize(c->stream);
        delete M;
        ctx_release(c);
    }
    BH_API_END
}

extern "C" int bh_matrix_size(const bh_matrix *M, int64_t *n)
{
    BH_API_BEGIN
    BH_REQUIRE(M && n, "NULL argument");
    *n = M->n;
    BH_API_END
}

extern "C" int bh_matrix_download(bh_matrix *M, double *A, int64_t ld)
{
    BH_API_BEGIN
    BH_REQUIRE(M && A, "NULL argument");
    BH_REQUIRE(ld >= M->n, "ld < n");
    const int64_t n = M->n;
    // only the n x n block is written: rows [n, ld) of each caller column stay untouched
    DevBuf<double> tmp((size_t)n * n * 2);
    dim3 grid((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32)), block(32, 8);
    k_rowmajor_to_colmajor<<<grid, block, 0, M->ctx->stream>>>((const double2 *)M->a64.p, (double2 *)tmp.p, n, n);
    BH_CHECK_HIP(hipGetLastError());
    BH_CHECK_HIP(hipMemcpy2DAsync(A, sizeof(double) * 2 * ld, tmp.p, sizeof(double) * 2 * n, sizeof(double) * 2 * n,
                                  n, hipMemcpyDeviceToHost, M->ctx->stream));
    BH_CHECK_HIP(hipStreamSynchronize(M->ctx->stream));
    BH_API_END
}
