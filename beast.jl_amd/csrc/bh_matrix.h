// Resident matrix object (row-major complex in HBM) and its factors.
#pragma once

#include <mutex>

#include "bh_internal.h"

namespace bh {
long long padded_ld(int64_t n);   // leading dimension of the LU factors
}

struct bh_matrix {
    bh::Ctx *ctx = nullptr;
    int64_t n = 0;
    bh::DevBuf<double> a64;        // FP64 matrix, row-major interleaved complex (n*n*2)
    bh::DevBuf<double> lu64;       // FP64 LU factors (when factored in FP64)
    bh::DevBuf<float> lu32;        // FP32 LU factors (when factored in FP32), or the rounded matrix (lu32_fresh)
    bool lu32_fresh = false;       // lu32 holds the current matrix rounded to FP32 (written by the assembly)
    bh::DevBuf<float> dinv32;      // FP32 factors: inverted 128 x 128 diagonal blocks of L and U, then G = Dinv x (block beside it)
    bh::DevBuf<float> x32;         // FP32 solve vector (n x 8 complex)
    bh::DevBuf<int> tsync;         // FP32 solves: ticket counter, error flag, one ready flag per 128-row block
    long long tsync_tickets = 0;   // tickets handed out so far
    int tsync_epoch = 0;           // value of the ready flags of the last solve
    bh::DevBuf<long long> ipiv;    // 0-based pivot rows (sequential interchanges)
    bh::DevBuf<int> perm;          // (P b)[i] = b[perm[i]]
    long long lda = 0;             // leading dimension of the factors
    int factored = -1;             // -1 none, BH_FP64 or BH_FP32
    double seconds_factor = 0.0;
    std::mutex mu;                 // factor and solve calls on one matrix run one at a time
    // the large buffers go back to the context's pool (the caller has synchronised)
    ~bh_matrix() {
        if (!ctx) return;
        bh::pool_release(*ctx, a64);
        bh::pool_release(*ctx, lu64);
        bh::pool_release(*ctx, lu32);
        bh::pool_release(*ctx, dinv32);
    }
};
