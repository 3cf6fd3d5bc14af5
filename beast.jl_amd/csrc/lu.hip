// Dense LU factorisation and solves (in progress).
#include "bh_internal.h"
#include "bh_matrix.h"

using namespace bh;

extern "C" int bh_matrix_factor(bh_matrix *M, int precision)
{
    BH_API_BEGIN
    BH_FAIL(BH_ERR_UNSUPPORTED, "bh_matrix_factor: not available in this build");
    BH_API_END
}

extern "C" int bh_matrix_solve(bh_matrix *M, int64_t nrhs, const double *B, double *X, double tol, int maxit,
                               bh_solve_stats *stats)
{
    BH_API_BEGIN
    BH_FAIL(BH_ERR_UNSUPPORTED, "bh_matrix_solve: not available in this build");
    BH_API_END
}

extern "C" int bh_matrix_pivots(bh_matrix *M, int64_t *ipiv)
{
    BH_API_BEGIN
    BH_FAIL(BH_ERR_UNSUPPORTED, "bh_matrix_pivots: not available in this build");
    BH_API_END
}
