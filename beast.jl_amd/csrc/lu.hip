// Dense complex LU with partial pivoting and the solves behind Z \ b.
//
// Reference: the reference factors its Matrix{ComplexF64} with LAPACK getrf and
// solves with getrs through LinearAlgebra (src/solvers/lusolver.jl:26-34,
// BEAST.lu in src/solvers/lu.jl:29-47).  Here the matrix stays in HBM, row-major
// interleaved complex, and is factored by a right-looking blocked LU
// (Golub & Van Loan, Matrix Computations, Alg. 3.4.x).
//
// FP64 factors (block depth 256) keep LAPACK's pivots exactly:
//   k_panel         32-column panel as one cooperative kernel: each workgroup keeps
//                   its rows in LDS; per column one grid barrier exchanges the
//                   per-workgroup pivot candidates (value, row, row contents).
//                   Pivot choice is LAPACK's: largest |Re|+|Im|, first row on ties.
// FP32 factors (block depth 1024 while the trailing matrix is large, factored as
// two 512-column halves, then 512) pivot each 32-column panel by tournament
// (communication-avoiding LU, Grigori, Demmel & Xiang, SIAM J. Matrix Anal. Appl.
// 32 (2011)):
//   k_tslu_panel    one launch per panel: 256 rows per workgroup, one per thread
//                   in registers, reduced to 32 candidates by partial pivoting;
//                   the last workgroup of each group of 8 to finish merges the
//                   group one level up, and the last merge emits L11\U11, ipiv and
//                   a row plan.  Further workgroups carry the previous panel's
//                   update of the block.
//   k_tslu_apply    moves the planned rows in the block's other columns, solves the
//                   panel's block row and turns the panel rows into L21 = A21 U11^-1.
//   The next block column is factored on a second stream while the trailing
//   update of the current one runs (look-ahead by one block).
// Shared by both:
//   k_swap_plan, k_laswp_apply  row interchanges: the sequential swaps replayed
//                   once on row indices, then each moved row moved once.
//   k_trsm_unit     32-row unit-lower solve for U12, column in registers.
//   k_cgemm         C -= A B on MFMA (v_mfma_{f32,f64}_16x16x4): FP64 updates and
//                   the FP32 updates inside a block column.
//   k_cgemm_bf16x3  FP32 trailing update from split-bf16 planes (hi*hi + hi*lo +
//                   lo*hi on v_mfma_f32_16x16x32_bf16, FP32 accumulation).
// Solves: permutation gather, then for FP64 factors blocked substitution
// (k_trsv_step); for FP32 factors one launch per triangle with the inverted
// 128x128 diagonal blocks and their couplings (k_trinv, k_blk_coupling,
// k_trsv_flag).  FP32 factors precondition flexible GMRES
// inside iterative refinement in FP64 (GMRES-IR, Carson & Higham, SIAM J. Sci.
// Comput. 40 (2018)), whose residuals use the resident FP64 matrix.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <climits>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "bh_device.h"
#include "bh_internal.h"
#include "bh_matrix.h"

namespace bh {

template <typename T>
struct CT;
template <>
struct CT<float> {
    using t = float2;
};
template <>
struct CT<double> {
    using t = double2;
};

template <typename V>
__device__ __forceinline__ V cmulv(V a, V b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

template <typename V>
__device__ __forceinline__ V csubv(V a, V b)
{
    return {a.x - b.x, a.y - b.y};
}

// ---------------------------------------------------------------- grid barrier
// Two-level barrier by generation.  Workgroup g arrives on the counter of group
// g % 8 (the XCD it was dispatched to, each counter on its own 128-byte line);
// the last arrival of a group arrives on the top counter; the last group bumps
// the generation that everyone waits on.  Counters reset themselves, so the
// state carries over between launches.  Agent-scope release/acquire makes the
// data written before the barrier visible to every XCD after it.
constexpr int kBarStride = 32;   // unsigned words per cache line
constexpr int kBarWords = kBarStride * 10;

__device__ __forceinline__ void grid_barrier(unsigned *bar, unsigned nblocks, unsigned &gen)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned ngroups = nblocks < 8 ? nblocks : 8;
        const unsigned grp = blockIdx.x & 7;
        const unsigned gsize = (nblocks - grp + 7) / 8;
        unsigned *gcnt = bar + kBarStride * (1 + grp);
        unsigned *top = bar + kBarStride * 9;
        __threadfence();
        const unsigned a = __hip_atomic_fetch_add(gcnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (a == gsize - 1) {
            __hip_atomic_store(gcnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned t = __hip_atomic_fetch_add(top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (t == ngroups - 1) {
                __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        while (__hip_atomic_load(&bar[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen)
            __builtin_amdgcn_s_sleep(1);
        __threadfence();
    }
    gen += 1;
    __syncthreads();
}

// (value, row) argmax across a wave: larger value wins, smaller row on ties
template <typename T>
__device__ __forceinline__ void wave_argmax(T &v, int &r)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const T v2 = __shfl_xor(v, off, 64);
        const int r2 = __shfl_xor(r, off, 64);
        if (v2 > v || (v2 == v && r2 < r)) { v = v2; r = r2; }
    }
}

// ---------------------------------------------------------------- panel
constexpr int kPW = 32;   // panel width

template <typename T>
struct PanelWork {
    T *slot_val;                    // [2][G]
    int *slot_row;                  // [2][G]
    typename CT<T>::t *cand;        // [2][G][kPW]
    typename CT<T>::t *diag;        // [2][kPW]
    unsigned *bar;                  // [kBarWords]
    int *info;                      // first zero pivot (1-based), 0 if none
};

template <typename T>
__global__ __launch_bounds__(256) void k_panel(typename CT<T>::t *A, long long lda, int n, int k0, int pw, int rpb,
                                               long long *ipiv, PanelWork<T> w)
{
    using V = typename CT<T>::t;
    extern __shared__ char smem_raw[];
    V *L = (V *)smem_raw;   // [rpb][kPW]
    __shared__ V prow[kPW], drow[kPW];
    __shared__ T wv_val[4];
    __shared__ int wv_row[4];
    const int G = gridDim.x;
    const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int r0 = k0 + g * rpb;
    const int r1 = min(n, r0 + rpb);
    const int nrows = max(0, r1 - r0);
    for (int idx = tid; idx < nrows * kPW; idx += 256) {
        const int r = idx / kPW, c = idx % kPW;
        L[idx] = c < pw ? A[(long long)(r0 + r) * lda + k0 + c] : V{0, 0};
    }
    unsigned gen = __hip_atomic_load(&w.bar[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    for (int j = 0; j < pw; ++j) {
        const int c = k0 + j;
        const int par = j & 1;
        // local candidate: largest |Re|+|Im| in column j among rows >= c, first row on ties
        T best = (T)-1;
        int brow = INT_MAX;
        for (int r = tid; r < nrows; r += 256) {
            const int grw = r0 + r;
            if (grw < c) continue;
            const V a = L[r * kPW + j];
            const T v = fabs(a.x) + fabs(a.y);
            if (v > best || (v == best && grw < brow)) { best = v; brow = grw; }
        }
        wave_argmax(best, brow);
        if (lane == 0) { wv_val[wid] = best; wv_row[wid] = brow; }
        __syncthreads();
        best = wv_val[0];
        brow = wv_row[0];
#pragma unroll
        for (int q = 1; q < 4; ++q)
            if (wv_val[q] > best || (wv_val[q] == best && wv_row[q] < brow)) { best = wv_val[q]; brow = wv_row[q]; }
        if (tid == 0) {
            w.slot_val[par * G + g] = best;
            w.slot_row[par * G + g] = brow;
        }
        if (brow != INT_MAX && tid < kPW) w.cand[((size_t)par * G + g) * kPW + tid] = L[(brow - r0) * kPW + tid];
        // wave 0 (lanes 32 .. 63) also writes the diagonal row: the barrier's release fence
        // is thread 0's and orders only its own wave's stores
        if (c >= r0 && c < r1 && tid >= 32 && tid < 32 + kPW) w.diag[par * kPW + tid - 32] = L[(c - r0) * kPW + tid - 32];
        grid_barrier(w.bar, G, gen);
        // global winner over the G candidates
        best = (T)-1;
        brow = INT_MAX;
        for (int t = tid; t < G; t += 256) {
            const T v = w.slot_val[par * G + t];
            const int rw = w.slot_row[par * G + t];
            if (rw == INT_MAX) continue;
            if (v > best || (v == best && rw < brow)) { best = v; brow = rw; }
        }
        wave_argmax(best, brow);
        if (lane == 0) { wv_val[wid] = best; wv_row[wid] = brow; }
        __syncthreads();
        int p = wv_row[0];
        best = wv_val[0];
#pragma unroll
        for (int q = 1; q < 4; ++q)
            if (wv_val[q] > best || (wv_val[q] == best && wv_row[q] < p)) { best = wv_val[q]; p = wv_row[q]; }
        // the workgroup owning row p published it
        const int gp = (p - k0) / rpb;
        if (tid < kPW) prow[tid] = w.cand[((size_t)par * G + gp) * kPW + tid];
        else if (tid >= 64 && tid < 64 + kPW) drow[tid - 64] = w.diag[par * kPW + tid - 64];
        __syncthreads();
        const V piv = prow[j];
        if (g == 0 && tid == 0) {
            ipiv[c] = p;
            if (piv.x == (T)0 && piv.y == (T)0 && *w.info == 0) *w.info = c + 1;
        }
        if (p != c) {
            if (p >= r0 && p < r1 && tid < kPW) L[(p - r0) * kPW + tid] = drow[tid];
            if (c >= r0 && c < r1 && tid >= 64 && tid < 64 + kPW) L[(c - r0) * kPW + tid - 64] = prow[tid - 64];
        }
        __syncthreads();
        if (piv.x != (T)0 || piv.y != (T)0) {
            const T den = piv.x * piv.x + piv.y * piv.y;
            const V inv{piv.x / den, -piv.y / den};
            for (int r = tid; r < nrows; r += 256) {
                if (r0 + r <= c) continue;
                const V l = cmulv(L[r * kPW + j], inv);
                L[r * kPW + j] = l;
                for (int jj = j + 1; jj < pw; ++jj) L[r * kPW + jj] = csubv(L[r * kPW + jj], cmulv(l, prow[jj]));
            }
        }
        __syncthreads();
    }
    for (int idx = tid; idx < nrows * kPW; idx += 256) {
        const int r = idx / kPW, c = idx % kPW;
        if (c < pw) A[(long long)(r0 + r) * lda + k0 + c] = L[idx];
    }
}

// ---------------------------------------------------------------- tournament pivoting
// Panel factorisation with tournament pivoting (TSLU, the panel step of CALU:
// Grigori, Demmel & Xiang, SIAM J. Matrix Anal. Appl. 32 (2011)), used by the
// FP32 factorisations that GMRES-IR refines.  Partial pivoting needs one
// grid-wide exchange per column; here the pw <= 32 pivot rows of a panel are chosen
//   1. per workgroup of kTsluT rows: partial pivoting selects <= pw candidates,
//   2. by a reduction tree: a workgroup stacks kTsluFan candidate sets (their
//      rows as they entered the panel) and keeps the rows partial pivoting selects,
// until one set remains.  Partial pivoting on the final stack picks its
// pivots among the winners in order, and eliminating with winners only leaves
// the winners' rows exactly as partial pivoting on the winners alone would
// (no loser is ever a pivot row), so the final workgroup also emits
// L11 \ U11, LAPACK-style sequential interchanges and a row-move plan.
// k_tslu_apply then moves the rows and forms L21 = A21 U11^-1.
constexpr int kTsluT = 256;             // rows (threads) per tournament workgroup
constexpr int kTsluFan = kTsluT / kPW;  // candidate sets merged per workgroup
constexpr int kPlanMax = 2 * kPW;       // row moves of one panel

// Maximum over a wave of unsigned keys (DPP: quad, half-row and row mirrors, then
// the row broadcasts of lanes 15 and 31), read from lane 63.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v)
{
    int x = (int)v;
    x = (int)max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(x, x, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    x = (int)max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(x, x, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    x = (int)max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(x, x, 0x141, 0xF, 0xF, false));  // row_half_mirror
    x = (int)max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(x, x, 0x140, 0xF, 0xF, false));  // row_mirror
    x = (int)max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(x, x, 0x142, 0xA, 0xF, false));  // row_bcast:15
    x = (int)max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(x, x, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return (unsigned)__builtin_amdgcn_readlane(x, 63);
}

// Partial pivoting over up to kTsluT rows held in registers, one row per thread
// (v: the row's first kPW entries; rows with used set never pivot).  piv[j]
// (LDS) receives the thread whose row pivots column j (-1 if none is left).
// Pivot choice: largest |Re|+|Im| at 2^-12 relative resolution (key = value bits
// with the low 11 mantissa bits replaced by 2047 - thread), smaller thread on
// ties.  One barrier per column: each wave finds its best key by DPP and that
// lane stores its row in the wave's slot (slots and keys alternate between two
// buffers by column parity), so after the barrier every thread reads the
// winning wave's row and eliminates with it in registers.  The pivot rows keep
// their factored entries (L left of the diagonal, U from it).  The columns run
// in eight phases of four: phase P updates the columns [4P, kPW), those of the
// phase under a predicate (a full unroll of the 32 x 32 elimination does not fit
// the instruction cache).  The elimination has no branch (rows that take no part
// subtract a zero multiple of the pivot row), and the next column's key and its
// DPP reduction follow its own update, so they overlap the other columns' updates.
// The slots must hold finite values before the first column (a column without a
// pivot reads wave 0's slot).
__device__ __forceinline__ unsigned pivot_key(float2 x, bool used, int tid)
{
    const float mag = fabsf(x.x) + fabsf(x.y);
    return used ? 0u : ((__float_as_uint(mag) & ~2047u) | (unsigned)(2047 - tid));
}

template <int P>
__device__ __forceinline__ void gepp_phase(float2 (&v)[kPW], bool &used, int pw, int *piv, unsigned (*s_key)[4],
                                           float2 (*s_row)[4][kPW], unsigned &key, unsigned &wmax)
{
    constexpr int C0 = 4 * P, NW = kTsluT / 64;
    static_assert(NW <= 4, "one key slot per wave");
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int jend = min(C0 + 4, pw);
#pragma unroll 1
    for (int j = C0; j < jend; ++j) {
        const int par = j & 1;
        // key and wmax hold column j's key of this row and the wave's best
        if (key == wmax && wmax != 0u) {
#pragma unroll
            for (int cc = C0; cc < kPW; ++cc) s_row[par][wid][cc] = v[cc];
        }
        if (lane == 0) s_key[par][wid] = wmax;
        __syncthreads();
        unsigned kmax = s_key[par][0];
        int wbest = 0;
#pragma unroll
        for (int q = 1; q < NW; ++q) {
            const unsigned kq = s_key[par][q];
            wbest = kq > kmax ? q : wbest;
            kmax = kq > kmax ? kq : kmax;
        }
        const int p = kmax ? 2047 - (int)(kmax & 2047u) : -1;
        if (tid == 0) piv[j] = p;
        if (tid == p) used = true;
        float2 pr[kPW];   // the pivot row, read in one batch
#pragma unroll
        for (int cc = C0; cc < kPW; ++cc) pr[cc] = s_row[par][wbest][cc];
        float2 d = pr[C0], vj = v[C0];
        // column j of this row and of the pivot row (selects per component: a
        // conditional struct copy makes the compiler move v to scratch memory)
#pragma unroll
        for (int q = 1; q < 4; ++q) {
            d.x = (j == C0 + q) ? pr[C0 + q].x : d.x;
            d.y = (j == C0 + q) ? pr[C0 + q].y : d.y;
            vj.x = (j == C0 + q) ? v[C0 + q].x : vj.x;
            vj.y = (j == C0 + q) ? v[C0 + q].y : vj.y;
        }
        const float den = fmaf(d.x, d.x, d.y * d.y);
        const bool elim = !used && p >= 0 && den != 0.f;
        // multiplier vj / d through the hardware reciprocal of |d|^2 (zero for rows
        // that take no part)
        const float rden = elim ? __builtin_amdgcn_rcpf(den) : 0.f;
        const float2 l{fmaf(vj.x, d.x, vj.y * d.y) * rden, fmaf(vj.y, d.x, -vj.x * d.y) * rden};
#pragma unroll
        for (int cc = C0; cc < kPW; ++cc) {
            const float ux = fmaf(l.y, pr[cc].y, fmaf(-l.x, pr[cc].x, v[cc].x));
            const float uy = fmaf(-l.y, pr[cc].x, fmaf(-l.x, pr[cc].y, v[cc].y));
            if (cc < C0 + 4) {
                v[cc].x = cc == j ? (elim ? l.x : v[cc].x) : (cc > j ? ux : v[cc].x);
                v[cc].y = cc == j ? (elim ? l.y : v[cc].y) : (cc > j ? uy : v[cc].y);
            } else {
                v[cc].x = ux;
                v[cc].y = uy;
            }
        }
        // the next column's key
        if (j + 1 < pw) {
            float2 vn = v[C0 + 1];
#pragma unroll
            for (int q = 2; q <= 4; ++q) {
                if (C0 + q < kPW) {
                    vn.x = (j + 1 == C0 + q) ? v[C0 + q].x : vn.x;
                    vn.y = (j + 1 == C0 + q) ? v[C0 + q].y : vn.y;
                }
            }
            key = pivot_key(vn, used, tid);
            wmax = wave_max_u32(key);
        }
    }
}

__device__ __forceinline__ void block_gepp(float2 (&v)[kPW], bool used, int pw, int *piv, unsigned (*s_key)[4],
                                           float2 (*s_row)[4][kPW])
{
    unsigned key = pivot_key(v[0], used, threadIdx.x);
    unsigned wmax = wave_max_u32(key);
    gepp_phase<0>(v, used, pw, piv, s_key, s_row, key, wmax);
    gepp_phase<1>(v, used, pw, piv, s_key, s_row, key, wmax);
    gepp_phase<2>(v, used, pw, piv, s_key, s_row, key, wmax);
    gepp_phase<3>(v, used, pw, piv, s_key, s_row, key, wmax);
    gepp_phase<4>(v, used, pw, piv, s_key, s_row, key, wmax);
    gepp_phase<5>(v, used, pw, piv, s_key, s_row, key, wmax);
    gepp_phase<6>(v, used, pw, piv, s_key, s_row, key, wmax);
    gepp_phase<7>(v, used, pw, piv, s_key, s_row, key, wmax);
    __syncthreads();
}

// x := x U^-1 for one row x (registers) and U upper triangular (LDS), column by
// column, in four phases of eight; row i of U is read in one batch.
template <int P>
__device__ __forceinline__ void right_trsv_phase(float2 (&x)[kPW], int pw, const float2 (*U)[kPW + 1])
{
    constexpr int C0 = 8 * P;
    const int iend = min(C0 + 8, pw);
#pragma unroll 1
    for (int i = C0; i < iend; ++i) {
        float2 ur[kPW];
#pragma unroll
        for (int cc = C0; cc < kPW; ++cc) ur[cc] = U[i][cc];
        float2 xi = x[C0], d = ur[C0];
#pragma unroll
        for (int q = 1; q < 8; ++q) {
            xi.x = (i == C0 + q) ? x[C0 + q].x : xi.x;
            xi.y = (i == C0 + q) ? x[C0 + q].y : xi.y;
            d.x = (i == C0 + q) ? ur[C0 + q].x : d.x;
            d.y = (i == C0 + q) ? ur[C0 + q].y : d.y;
        }
        const float den = fmaf(d.x, d.x, d.y * d.y);
        const float rden = den != 0.f ? __builtin_amdgcn_rcpf(den) : 0.f;
        xi = float2{fmaf(xi.x, d.x, xi.y * d.y) * rden, fmaf(xi.y, d.x, -xi.x * d.y) * rden};
#pragma unroll
        for (int cc = C0; cc < kPW; ++cc) {
            const float ux = fmaf(xi.y, ur[cc].y, fmaf(-xi.x, ur[cc].x, x[cc].x));
            const float uy = fmaf(-xi.y, ur[cc].x, fmaf(-xi.x, ur[cc].y, x[cc].y));
            if (cc < C0 + 8) {
                x[cc].x = cc == i ? xi.x : (cc > i ? ux : x[cc].x);
                x[cc].y = cc == i ? xi.y : (cc > i ? uy : x[cc].y);
            } else {
                x[cc].x = ux;
                x[cc].y = uy;
            }
        }
    }
}

struct TsluFinal {
    long long *ipiv;   // rows c .. c+pw receive the sequential interchanges
    float2 *top;       // [kPW][kPW] L11 \ U11, row j = the row that goes to c + j
    float2 *oldtop;    // [kPW][kPW] panel entries of rows c .. c+pw before the interchanges
    int *plan;         // [0] = m, then m destinations, then m sources (rows before the moves)
    int *info;         // first zero pivot (1-based), 0 if none
};

template <typename T, int TM, int TN, bool M3>
__device__ __forceinline__ void cgemm_tile(int id, const typename CT<T>::t *__restrict__ Ap, long long lda,
                                           const typename CT<T>::t *__restrict__ Bp, long long ldb,
                                           typename CT<T>::t *Cp, long long ldc, int M, int N, int K);

// Counters and candidate sets of the tournament.  Level l holds n_l sets
// (n_0 = workgroups of the first stage, n_{l+1} = ceil(n_l / kTsluFan)); its sets
// start at set offset n_0 + ... + n_{l-1} and its group counters at
// ceil(n_0 / kTsluFan) + ... + ceil(n_{l-1} / kTsluFan).
struct TsluWork {
    int *cnt;      // arrivals per group; every counter is zero between panels
    int *idx;      // [sets][kPW] candidate row ids, -1 for none
    float2 *val;   // [sets][kPW][kPW] their panel entries (after the previous panel's update)
};

// The tournament of the panel [c, c + pw) as one launch, and the previous
// panel's update of the block columns to its right.
// Workgroups [0, nsets0) each take kTsluT rows (r >= c): with prev set they first
// apply the previous panel's update to their rows of the panel columns
// (A[r, c:c+pw) -= L21(r) U12, L21 = A[r, c-kPW:c), U12 = A[c-kPW:c, c:c+pw)) and
// write it back, then select <= pw candidates by partial pivoting.  A workgroup
// publishes its set and counts itself into its group of kTsluFan sets (agent-scope
// release / acquire); the last to arrive merges the group's sets at the next
// level, and so on, so no workgroup waits for another.  The workgroup that
// partial-pivots the last single set writes L11 \ U11, ipiv, the panel rows
// before the moves and the row-move plan.
// Workgroups [nsets0, ...) are the tiles of A[c:, c+kPW : c+kPW+upd_n) -=
// L21 U12 with the same L21 and U12 (the rest of the previous panel's update).
__device__ __forceinline__ void tslu_panel_body(float2 *A, long long lda, int n, int c, int pw, int nsets0, int prev,
                                                int upd_n, const TsluWork &tw, const TsluFinal &fin, int bid)
{
    if (bid >= nsets0) {
        cgemm_tile<float, 4, 2, true>(bid - nsets0, A + (long long)c * lda + c - kPW, lda,
                                      A + (long long)(c - kPW) * lda + c + kPW, lda, A + (long long)c * lda + c + kPW,
                                      lda, n - c, upd_n, kPW);
        return;
    }
    __shared__ unsigned s_key[2][4];
    __shared__ __align__(16) float2 s_row[2][4][kPW];
    __shared__ int piv[kPW], s_rowid[kTsluT], zero_at, s_next;
    __shared__ int s_win[kPW], s_at[kPW], s_pos[kPW];
    __shared__ __align__(16) float2 s_u[kPW][kPW];
    const int tid = threadIdx.x;
    for (int e = tid; e < 2 * 4 * kPW; e += kTsluT) (&s_row[0][0][0])[e] = float2{0, 0};   // finite slots
    __syncthreads();
    int g = bid, nl = nsets0, off = 0, coff = 0;
    int rowid = -1;
    float2 v[kPW], orig[kPW];
    {
        const long long r = (long long)c + (long long)g * kTsluT + tid;
        if (r < n) rowid = (int)r;
#pragma unroll
        for (int cc = 0; cc < kPW; ++cc) v[cc] = (rowid >= 0 && cc < pw) ? A[r * lda + c + cc] : float2{0, 0};
        if (prev) {
            for (int idx = tid; idx < kPW * kPW; idx += kTsluT) {
                const int i = idx / kPW, cc = idx % kPW;
                s_u[i][cc] = cc < pw ? A[(long long)(c - kPW + i) * lda + c + cc] : float2{0, 0};
            }
            __syncthreads();
            if (rowid >= 0) {
                const float2 *lrow = A + r * lda + c - kPW;
                float2 l = lrow[0];
#pragma unroll 1
                for (int i = 0; i < kPW; ++i) {
                    const float2 ln = lrow[min(i + 1, kPW - 1)];   // the next multiplier, loaded ahead
                    float2 ur[kPW];                                 // row i of U12, read in one batch
#pragma unroll
                    for (int cc = 0; cc < kPW; ++cc) ur[cc] = s_u[i][cc];
#pragma unroll
                    for (int cc = 0; cc < kPW; ++cc) {
                        v[cc].x = fmaf(l.y, ur[cc].y, fmaf(-l.x, ur[cc].x, v[cc].x));
                        v[cc].y = fmaf(-l.y, ur[cc].x, fmaf(-l.x, ur[cc].y, v[cc].y));
                    }
                    l = ln;
                }
#pragma unroll
                for (int cc = 0; cc < kPW; ++cc)
                    if (cc < pw) A[r * lda + c + cc] = v[cc];
            }
        }
    }
#pragma unroll
    for (int cc = 0; cc < kPW; ++cc) orig[cc] = v[cc];
    s_rowid[tid] = rowid;
    if (tid == 0) zero_at = INT_MAX;
    for (;;) {
        block_gepp(v, rowid < 0, pw, piv, s_key, s_row);
        if (nl == 1) break;
        // publish this set: the owner of pivot j stages its entries in LDS (s_u is free
        // after the first stage), and wave 0 writes the row ids and entries, so that
        // thread 0's release fence covers every store of the set (a barrier does not wait
        // for other waves' stores, and a fence orders its own wave's)
        const size_t set = (size_t)(off + g);
        int j_of = -1;
        for (int j = 0; j < pw; ++j)
            if (piv[j] == tid) j_of = j;
        if (j_of >= 0) {
#pragma unroll
            for (int cc = 0; cc < kPW; ++cc) s_u[j_of][cc] = orig[cc];
        }
        __syncthreads();
        if (tid < 64) {
            if (tid < kPW) {
                const int p = tid < pw ? piv[tid] : -1;
                tw.idx[set * kPW + tid] = p >= 0 ? s_rowid[p] : -1;
            }
            for (int e = tid; e < kPW * kPW; e += 64) {
                const int j = e / kPW;
                if (j < pw && piv[j] >= 0) tw.val[set * kPW * kPW + e] = s_u[j][e % kPW];
            }
        }
        if (tid == 0) {
            const int grp = g / kTsluFan, members = min(kTsluFan, nl - grp * kTsluFan);
            __threadfence();
            const int arrived = __hip_atomic_fetch_add(tw.cnt + coff + grp, 1, __ATOMIC_ACQ_REL,
                                                       __HIP_MEMORY_SCOPE_AGENT);
            int next = -1;
            if (arrived == members - 1) {
                __hip_atomic_store(tw.cnt + coff + grp, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __threadfence();
                next = grp;
            }
            s_next = next;
        }
        __syncthreads();
        const int grp = s_next;
        if (grp < 0) return;
        // the last of its group: merge the group's sets one level up
        const size_t first = (size_t)off + (size_t)grp * kTsluFan;
        const int nrows = min(kTsluFan, nl - grp * kTsluFan) * kPW;
        coff += (nl + kTsluFan - 1) / kTsluFan;
        off += nl;
        nl = (nl + kTsluFan - 1) / kTsluFan;
        g = grp;
        rowid = tid < nrows ? tw.idx[first * kPW + tid] : -1;
#pragma unroll
        for (int cc = 0; cc < kPW; ++cc) v[cc] = rowid >= 0 ? tw.val[(first * kPW + tid) * kPW + cc] : float2{0, 0};
#pragma unroll
        for (int cc = 0; cc < kPW; ++cc) orig[cc] = v[cc];
        s_rowid[tid] = rowid;
    }
    // last stage: the owners of the pivot rows write L11 \ U11
    int j_of = -1;
    for (int j = 0; j < pw; ++j)
        if (piv[j] == tid) j_of = j;
    if (j_of >= 0) {
#pragma unroll
        for (int cc = 0; cc < kPW; ++cc) {
            fin.top[j_of * kPW + cc] = cc < pw ? v[cc] : float2{0, 0};
            if (cc == j_of && v[cc].x == 0.f && v[cc].y == 0.f) atomicMin(&zero_at, j_of);
        }
    }
    for (int idx = tid; idx < kPW * kPW; idx += kTsluT) {
        const int j = idx / kPW, cc = idx % kPW;
        if (j >= pw) fin.top[idx] = float2{0, 0};
        const long long r = (long long)c + j;
        fin.oldtop[idx] = (j < pw && cc < pw) ? A[r * lda + c + cc] : float2{0, 0};
    }
    if (tid < kPW) {
        const int p = tid < pw ? piv[tid] : -1;
        s_win[tid] = p >= 0 ? s_rowid[p] : -1;
        s_at[tid] = c + tid;    // row now at position c + tid
        s_pos[tid] = c + tid;   // position of row c + tid
    }
    __syncthreads();
    if (tid == 0) {
        // sequential interchanges bringing winner j to position c + j: rows of the
        // panel's own positions are tracked in s_at / s_pos; a winner from further
        // down has not moved before its own step
        int first_bad = zero_at;
        for (int j = 0; j < pw; ++j) {
            const int here = c + j, w = s_win[j];
            if (w < 0) {   // fewer candidates than columns: cannot happen for pw <= n - c
                fin.ipiv[here] = here;
                first_bad = min(first_bad, j);
                continue;
            }
            const int p = (w < c + pw) ? s_pos[w - c] : w;
            fin.ipiv[here] = p;
            if (p != here) {
                const int occupant = s_at[j];   // a row of the panel's own positions
                s_at[j] = w;
                if (p < c + pw) s_at[p - c] = occupant;
                if (w < c + pw) s_pos[w - c] = here;
                s_pos[occupant - c] = p;
            }
        }
        // moves: positions c + j receive the winners; a panel row that ended
        // below the panel's positions moves there
        int m = 0;
        for (int j = 0; j < pw; ++j) {
            const int w = s_win[j] >= 0 ? s_win[j] : c + j;
            if (w != c + j) {
                fin.plan[1 + m] = c + j;
                fin.plan[1 + kPlanMax + m] = w;
                ++m;
            }
        }
        for (int t = 0; t < pw; ++t) {
            if (s_pos[t] >= c + pw) {
                fin.plan[1 + m] = s_pos[t];
                fin.plan[1 + kPlanMax + m] = c + t;
                ++m;
            }
        }
        fin.plan[0] = m;
        if (first_bad != INT_MAX && *fin.info == 0) *fin.info = c + first_bad + 1;
    }
}

__global__ __launch_bounds__(kTsluT, 2) void k_tslu_panel(float2 *A, long long lda, int n, int c, int pw, int nsets0,
                                                          int prev, int upd_n, TsluWork tw, TsluFinal fin)
{
    tslu_panel_body(A, lda, n, c, pw, nsets0, prev, upd_n, tw, fin, blockIdx.x);
}

// After a tournament: workgroups [0, nsw) move the planned rows in the columns
// [a0, a1) and [b0, b1) (16 columns each, all sources read before any
// destination is written); the others finish the panel rows r >= c: rows
// c .. c+pw become L11 \ U11, every other row x (its pre-move panel entries,
// from oldtop when it was a panel row) becomes L21 = x U11^-1.
__device__ __forceinline__ void tslu_apply_body(float2 *A, long long lda, int n, int c, int pw, const float2 *top,
                                                const float2 *oldtop, const int *plan, int a0, int a1, int b0, int b1,
                                                int nsw, int bid)
{
    __shared__ int s_dst[kPlanMax], s_src[kPlanMax], s_m;
    __shared__ float2 s_buf[kPlanMax][16];
    __shared__ float2 U[kPW][kPW + 1];
    const int tid = threadIdx.x;
    if (tid == 0) s_m = plan[0];
    if (tid < kPlanMax) {
        s_dst[tid] = plan[1 + tid];
        s_src[tid] = plan[1 + kPlanMax + tid];
    }
    __syncthreads();
    const int m = s_m;
    if (bid < nsw) {
        const int cl = tid & 15, rl = tid >> 4;
        const int t = bid * 16 + cl, w1 = a1 - a0;
        int col = -1;
        if (t < w1) col = a0 + t;
        else if (t < w1 + (b1 - b0)) col = b0 + (t - w1);
        for (int s = rl; s < m; s += 16)
            if (col >= 0) s_buf[s][cl] = A[(long long)s_src[s] * lda + col];
        __syncthreads();
        for (int s = rl; s < m; s += 16)
            if (col >= 0) A[(long long)s_dst[s] * lda + col] = s_buf[s][cl];
        if (bid * 16 + 16 <= w1 || b1 <= b0) return;   // no columns right of the panel
        // the panel's block row in the columns to its right: U12 = L11^-1 (moved rows c .. c+pw),
        // L11 the unit lower part of top; one thread per column, the column in registers
        for (int idx = tid; idx < kPW * kPW; idx += 256) U[idx / kPW][idx % kPW] = top[idx];
        __syncthreads();
        if (rl != 0 || t < w1 || col < 0) return;
        float2 x[kPW];
#pragma unroll
        for (int i = 0; i < kPW; ++i) x[i] = i < pw ? A[(long long)(c + i) * lda + col] : float2{0, 0};
#pragma unroll
        for (int i = 1; i < kPW; ++i) {
            if (i < pw) {   // uniform
                float2 acc = x[i];
#pragma unroll
                for (int j = 0; j < i; ++j) acc = csubv(acc, cmulv(U[i][j], x[j]));
                x[i] = acc;
            }
        }
#pragma unroll
        for (int i = 0; i < kPW; ++i)
            if (i < pw) A[(long long)(c + i) * lda + col] = x[i];
        return;
    }
    for (int idx = tid; idx < kPW * kPW; idx += 256) U[idx / kPW][idx % kPW] = top[idx];
    __syncthreads();
    const long long r = (long long)c + (long long)(bid - nsw) * 256 + tid;
    if (r >= n) return;
    float2 *row = A + r * lda + c;
    if (r < c + pw) {
        const int j = (int)(r - c);
        for (int cc = 0; cc < pw; ++cc) row[cc] = U[j][cc];
        return;
    }
    int from = -1;   // panel row whose entries move here
    for (int s = 0; s < m; ++s)
        if (s_dst[s] == r) from = s_src[s] - c;
    const float2 *srow = from >= 0 ? oldtop + (size_t)from * kPW : row;
    float2 x[kPW];
#pragma unroll
    for (int cc = 0; cc < kPW; ++cc) x[cc] = cc < pw ? srow[cc] : float2{0, 0};
    right_trsv_phase<0>(x, pw, U);
    right_trsv_phase<1>(x, pw, U);
    right_trsv_phase<2>(x, pw, U);
    right_trsv_phase<3>(x, pw, U);
#pragma unroll
    for (int cc = 0; cc < kPW; ++cc)
        if (cc < pw) row[cc] = x[cc];
}

__global__ __launch_bounds__(256) void k_tslu_apply(float2 *A, long long lda, int n, int c, int pw, const float2 *top,
                                                    const float2 *oldtop, const int *plan, int a0, int a1, int b0,
                                                    int b1, int nsw)
{
    tslu_apply_body(A, lda, n, c, pw, top, oldtop, plan, a0, a1, b0, b1, nsw, blockIdx.x);
}

// ---------------------------------------------------------------- row interchanges
// Sequential interchanges ipiv[k1..k2) (LAPACK laswp) on a set of columns.
// The same interchanges as one permutation of the touched rows.  Replaying
// them sequentially per column costs one dependent global round trip per
// interchange; instead k_swap_plan replays them once on row indices (LDS) and
// k_laswp_apply moves every touched row of every column in a single pass.
// Plan slots: [0, len) are rows k1 + i; [len, 2 len) the rows ipiv[k1 + i] >= k2
// (slot len + i for the first i naming that row).  plan[s] = the row whose
// content ends at slot s, plan[2 len + s] = the slot's row (-1: slot unused).
constexpr int kMaxSwaps = 256;
// FP32 LU block depth (trailing-update depth): while the trailing matrix is
// large the blocks are 1024 deep, factored as two halves of kNBInner one after
// the other on the block-column stream; that halves the C traffic of the
// trailing updates, which pays while the trailing update hides the longer
// block-column chain.  Smaller trailing matrices take 512-deep blocks.
constexpr int kNBFP32 = 1024, kNBFP32Small = 512, kNBDeepAbove = 24576;
constexpr int kNBInner = 512;
constexpr int kAheadFirstBelow = 20480;   // trailing width below which the next block column goes first

// The 1024-deep block columns join their 512-column halves with split-bf16 products,
// so the plain FP32 factorisation keeps 512 throughout.
inline int fp32_block_depth(int n, int k, bool split)
{
    return split && n - k > kNBDeepAbove ? kNBFP32 : kNBFP32Small;
}

__device__ __forceinline__ void swap_plan_body(const long long *ipiv, int k1, int k2, int *plan)
{
    __shared__ int s_piv[kMaxSwaps], s_partner[kMaxSwaps], s_src[2 * kMaxSwaps];
    const int len = k2 - k1, tid = threadIdx.x;
    for (int i = tid; i < len; i += 256) s_piv[i] = (int)ipiv[k1 + i];
    __syncthreads();
    for (int i = tid; i < len; i += 256) {
        const int p = s_piv[i];
        int partner = p - k1;
        if (p >= k2) {
            int first = i;
            for (int j = 0; j < i; ++j)
                if (s_piv[j] == p) { first = j; break; }
            partner = len + first;
        }
        s_partner[i] = partner;
        s_src[i] = k1 + i;
        const bool head = p >= k2 && partner == len + i;
        s_src[len + i] = head ? p : -1;
        plan[2 * len + i] = k1 + i;
        plan[3 * len + i] = head ? p : -1;
    }
    __syncthreads();
    if (tid == 0) {
        for (int i = 0; i < len; ++i) {
            const int q = s_partner[i];
            const int t = s_src[i];
            s_src[i] = s_src[q];
            s_src[q] = t;
        }
    }
    __syncthreads();
    for (int s = tid; s < 2 * len; s += 256) plan[s] = s_src[s];
}

__global__ __launch_bounds__(256) void k_swap_plan(const long long *ipiv, int k1, int k2, int *plan)
{
    swap_plan_body(ipiv, k1, k2, plan);
}

// Apply a swap plan of 2 len slots to columns [c0, c1) and [c2, c3): CW
// columns per workgroup (64 KiB of LDS at most), every moved row read into LDS
// before any is written.
template <typename T>
constexpr int laswp_cols() { return sizeof(T) == 4 ? 32 : 8; }

template <typename T>
__device__ __forceinline__ void laswp_apply_body(typename CT<T>::t *A, long long lda, const int *plan, int len, int c0,
                                                 int c1, int c2, int c3, int bid)
{
    using V = typename CT<T>::t;
    constexpr int CW = laswp_cols<T>(), RL = 256 / CW;
    extern __shared__ char smem_raw[];
    V *buf = (V *)smem_raw;   // [2 len][CW]
    const int tid = threadIdx.x, cl = tid % CW, rl = tid / CW;
    const int t = bid * CW + cl, w1 = c1 - c0;
    int col = -1;
    if (t < w1) col = c0 + t;
    else if (t < w1 + (c3 - c2)) col = c2 + (t - w1);
    const int *src = plan, *pos = plan + 2 * len;
    for (int s = rl; s < 2 * len; s += RL) {
        const int p = pos[s], q = src[s];
        if (col >= 0 && p >= 0 && p != q) buf[s * CW + cl] = A[(long long)q * lda + col];
    }
    __syncthreads();
    for (int s = rl; s < 2 * len; s += RL) {
        const int p = pos[s], q = src[s];
        if (col >= 0 && p >= 0 && p != q) A[(long long)p * lda + col] = buf[s * CW + cl];
    }
}

// FP32: a workgroup takes 32 columns, one per thread and eight slots at a time; every
// moved entry is held in registers (all its loads in flight at once, no LDS staging),
// then written after the barrier.
template <>
__device__ __forceinline__ void laswp_apply_body<float>(float2 *A, long long lda, const int *plan, int len, int c0,
                                                        int c1, int c2, int c3, int bid)
{
    constexpr int CW = laswp_cols<float>(), RL = 256 / CW, MAXK = 2 * kMaxSwaps / RL;
    __shared__ int s_src[2 * kMaxSwaps], s_pos[2 * kMaxSwaps];
    const int tid = threadIdx.x, cl = tid % CW, rl = tid / CW;
    for (int q = tid; q < 2 * len; q += 256) {
        s_src[q] = plan[q];
        s_pos[q] = plan[2 * len + q];
    }
    __syncthreads();
    const int t = bid * CW + cl, w1 = c1 - c0;
    int col = -1;
    if (t < w1) col = c0 + t;
    else if (t < w1 + (c3 - c2)) col = c2 + (t - w1);
    float2 v[MAXK];
#pragma unroll
    for (int k = 0; k < MAXK; ++k) {
        const int sl = rl + RL * k;
        if (sl < 2 * len && col >= 0) {
            const int p = s_pos[sl], q = s_src[sl];
            if (p >= 0 && p != q) v[k] = A[(long long)q * lda + col];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MAXK; ++k) {
        const int sl = rl + RL * k;
        if (sl < 2 * len && col >= 0) {
            const int p = s_pos[sl], q = s_src[sl];
            if (p >= 0 && p != q) A[(long long)p * lda + col] = v[k];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_laswp_apply(typename CT<T>::t *A, long long lda, const int *plan, int len,
                                                     int c0, int c1, int c2, int c3)
{
    laswp_apply_body<T>(A, lda, plan, len, c0, c1, c2, c3, blockIdx.x);
}

// ---------------------------------------------------------------- small triangular solve
// B (m rows x ncols, m <= 32) := L^{-1} B with L unit lower (m x m), both inside A.
// One wave per workgroup, lane = column: the column lives in registers (all
// its loads in flight at once), L in LDS (every read a broadcast).
template <typename T>
__global__ __launch_bounds__(64) void k_trsm_unit(typename CT<T>::t *A, long long lda, int lr, int lc, int m, int br,
                                                  int bc, int ncols)
{
    using V = typename CT<T>::t;
    __shared__ V Ls[kPW][kPW + 1];
    for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
        const int i = idx / m, j = idx % m;
        Ls[i][j] = A[(long long)(lr + i) * lda + lc + j];
    }
    const int lane = threadIdx.x;
    const int col = blockIdx.x * 64 + lane;
    const bool act = col < ncols;
    V x[kPW];
#pragma unroll
    for (int i = 0; i < kPW; ++i) x[i] = (act && i < m) ? A[(long long)(br + i) * lda + bc + col] : V{0, 0};
    __syncthreads();
#pragma unroll
    for (int i = 1; i < kPW; ++i) {
        if (i < m) {   // uniform
            V s = x[i];
#pragma unroll
            for (int j = 0; j < i; ++j) s = csubv(s, cmulv(Ls[i][j], x[j]));
            x[i] = s;
        }
    }
    if (act) {
#pragma unroll
        for (int i = 0; i < kPW; ++i)
            if (i < m) A[(long long)(br + i) * lda + bc + col] = x[i];
    }
}

// X := L^-1 X on the rows [k, k + kb) (kb <= kNBInner) of the columns
// [c0, c0 + ncols), L the unit lower triangle of A[k:k+kb, k:k+kb): the FP32
// block row in one launch.  Workgroup g keeps columns [c0 + 16 g, + 16) of the
// rows in LDS and walks the 32-row blocks p: threads 0..15 solve block p with the
// diagonal tile of L (their column in registers), then every thread holds
// column c = t % 16 of the solved block in registers and updates its rows
// below (rows 32 (p + 1) + t / 16 + 16 m), reading the rows of L from L2.
constexpr int kStripW = 16;
constexpr int kNBInnerMax = 512;
constexpr int kBlockRowOneLaunch = 8192;   // widest block row solved by k_blockrow_trsm

__global__ __launch_bounds__(256) void k_blockrow_trsm(float2 *A, long long lda, int k, int kb, int c0, int ncols)
{
    __shared__ float2 X[kNBInnerMax][kStripW];
    __shared__ float2 Lt[kPW][kPW + 1];
    const int tid = threadIdx.x, c = tid % kStripW, g = tid / kStripW;
    const int col0 = c0 + (int)blockIdx.x * kStripW;
    const int w = min(kStripW, c0 + ncols - col0);
    for (int idx = tid; idx < kb * kStripW; idx += 256) {
        const int r = idx / kStripW, cc = idx % kStripW;
        X[r][cc] = cc < w ? A[(long long)(k + r) * lda + col0 + cc] : float2{0, 0};
    }
    const int np = (kb + kPW - 1) / kPW;
    for (int p = 0; p < np; ++p) {
        const int r0 = p * kPW, pr = min(kPW, kb - r0);
        for (int idx = tid; idx < kPW * kPW; idx += 256) {
            const int i = idx / kPW, j = idx % kPW;
            Lt[i][j] = (i < pr && j < pr) ? A[(long long)(k + r0 + i) * lda + k + r0 + j] : float2{0, 0};
        }
        __syncthreads();
        if (tid < kStripW) {
            float2 x[kPW];
#pragma unroll
            for (int i = 0; i < kPW; ++i) x[i] = X[r0 + i < kb ? r0 + i : 0][tid];
#pragma unroll
            for (int i = 1; i < kPW; ++i) {
                if (i < pr) {   // uniform
                    float2 acc = x[i];
#pragma unroll
                    for (int j = 0; j < i; ++j) acc = csubv(acc, cmulv(Lt[i][j], x[j]));
                    x[i] = acc;
                }
            }
#pragma unroll
            for (int i = 0; i < kPW; ++i)
                if (i < pr) X[r0 + i][tid] = x[i];
        }
        __syncthreads();
        if (r0 + pr < kb) {
            float2 xp[kPW];   // column c of the solved block
#pragma unroll
            for (int j = 0; j < kPW; ++j) xp[j] = j < pr ? X[r0 + j][c] : float2{0, 0};
            // the rows below in tiles of 32: the tile of L through LDS (coalesced),
            // thread (c, g) updates rows g and g + 16 of the tile
            for (int q0 = r0 + pr; q0 < kb; q0 += kPW) {
                const int qr = min(kPW, kb - q0);
                __syncthreads();
                for (int idx = tid; idx < kPW * kPW; idx += 256) {
                    const int i = idx / kPW, j = idx % kPW;
                    Lt[i][j] = (i < qr && j < pr) ? A[(long long)(k + q0 + i) * lda + k + r0 + j] : float2{0, 0};
                }
                __syncthreads();
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int i = g + 16 * h;
                    if (i < qr) {
                        float2 acc = X[q0 + i][c];
#pragma unroll
                        for (int j = 0; j < kPW; ++j) {
                            const float2 l = Lt[i][j];
                            acc.x = fmaf(-l.x, xp[j].x, fmaf(l.y, xp[j].y, acc.x));
                            acc.y = fmaf(-l.x, xp[j].y, fmaf(-l.y, xp[j].x, acc.y));
                        }
                        X[q0 + i][c] = acc;
                    }
                }
            }
        }
        __syncthreads();
    }
    for (int idx = tid; idx < kb * kStripW; idx += 256) {
        const int r = idx / kStripW, cc = idx % kStripW;
        if (cc < w) A[(long long)(k + r) * lda + col0 + cc] = X[r][cc];
    }
}

// ---------------------------------------------------------------- MFMA complex GEMM
// C[M x N] -= A[M x K] B[K x N], row-major complex, K a multiple of 16.
// Workgroup: 4 waves in 2 x 2, wave tile (16*TM) x (16*TN) complex.
template <typename T>
struct MfmaVec;
template <>
struct MfmaVec<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
};
template <>
struct MfmaVec<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
};

template <typename T>
__device__ __forceinline__ typename MfmaVec<T>::v4 mfma16(T a, T b, typename MfmaVec<T>::v4 c);

template <>
__device__ __forceinline__ MfmaVec<float>::v4 mfma16<float>(float a, float b, MfmaVec<float>::v4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <>
__device__ __forceinline__ MfmaVec<double>::v4 mfma16<double>(double a, double b, MfmaVec<double>::v4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kBK = 16;

// C[M x N] -= A[M x K] B[K x N], row-major complex, K a multiple of kBK.
// 256 threads = 2 x 2 waves, wave tile (16 TM) x (16 TN) complex, workgroup
// tile (32 TM) x (32 TN).  K advances in chunks of 16 through two LDS buffers:
// the next chunk is loaded into registers while the current one feeds the
// MFMAs, then stored to the other buffer (one barrier per chunk).
// M3 = false: four real products per complex product (Cr += ArBr - AiBi,
// Ci += ArBi + AiBr); M3 = true: three (T1 = ArBr, T2 = AiBi,
// T3 = (Ar+Ai)(Br+Bi), Cr = T1 - T2, Ci = T3 - T1 - T2), normwise stable and used
// for the FP32 factorisation that GMRES-IR refines.
// Tile of a GEMM launched as a 1-D grid of nbm x nbn tiles.  Workgroups are
// dispatched round-robin over the 8 XCDs (workgroup id mod 8), each with its
// own L2: the id is first remapped so every XCD owns a contiguous range of
// tiles, which is then walked in groups of kGroupM tile rows (column-major
// inside a group), so the workgroups an XCD runs at once share A row panels
// and B column panels in its L2.  Four tile rows per group (the 32 tiles an XCD
// runs at once: 4 x 8) measured 1% faster than 8, 2 or 16 on the C3 factorisation.
constexpr int kXcds = 8, kGroupM = 4;
__device__ __forceinline__ void gemm_tile(int id, int nbm, int nbn, int &tm, int &tn)
{
    const int total = nbm * nbn;
    const int q = total / kXcds, r = total % kXcds, x = id % kXcds;
    const int t = x * q + min(x, r) + id / kXcds;
    const int per_group = kGroupM * nbn;
    const int first = (t / per_group) * kGroupM;
    const int gm = min(kGroupM, nbm - first);
    const int in = t % per_group;
    tm = first + in % gm;
    tn = in / gm;
}

// One tile (linear index id, 256 threads) of the GEMM below.
template <typename T, int TM, int TN, bool M3>
__device__ __forceinline__ void cgemm_tile(int id, const typename CT<T>::t *__restrict__ Ap, long long lda,
                                           const typename CT<T>::t *__restrict__ Bp, long long ldb,
                                           typename CT<T>::t *Cp, long long ldc, int M, int N, int K)
{
    using V = typename CT<T>::t;
    using v4 = typename MfmaVec<T>::v4;
    constexpr int BM = 32 * TM, BN = 32 * TN;
    constexpr int APT = BM * kBK / 256;          // A elements per thread
    constexpr int BPT = kBK * BN / 256;          // B elements per thread
    constexpr int AROW = kBK / APT;              // threads per A row
    constexpr int BROW = BN / BPT;               // threads per B row
    constexpr int NACC = M3 ? 3 : 2;
    __shared__ V As[2][BM][kBK + 1];
    __shared__ V Bs[2][kBK][BN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    int tm, tn;
    gemm_tile(id, (M + BM - 1) / BM, (N + BN - 1) / BN, tm, tn);
    const long long m0 = (long long)tm * BM, n0 = (long long)tn * BN;
    const int ar = tid / AROW, ac = (tid % AROW) * APT;
    const int br = tid / BROW, bc = (tid % BROW) * BPT;
    const bool arow_ok = m0 + ar < M;
    const V *ag = Ap + (m0 + ar) * lda + ac;
    const V *bg = Bp + (long long)br * ldb + n0 + bc;
    V ra[APT], rb[BPT];
    auto load = [&](int k0) {
#pragma unroll
        for (int j = 0; j < APT; ++j) ra[j] = arow_ok ? ag[k0 + j] : V{0, 0};
#pragma unroll
        for (int j = 0; j < BPT; ++j) rb[j] = (n0 + bc + j < N) ? bg[(long long)k0 * ldb + j] : V{0, 0};
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int j = 0; j < APT; ++j) As[buf][ar][ac + j] = ra[j];
#pragma unroll
        for (int j = 0; j < BPT; ++j) Bs[buf][br][bc + j] = rb[j];
    };
    v4 acc[NACC][TM][TN];
#pragma unroll
    for (int q = 0; q < NACC; ++q)
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b) acc[q][a][b] = v4{0, 0, 0, 0};
    const int li = lane & 15, lk = lane >> 4;
    load(0);
    store(0);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < K; k0 += kBK) {
        const bool more = k0 + kBK < K;
        if (more) load(k0 + kBK);
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 4) {
            V af[TM], bf[TN];
#pragma unroll
            for (int a = 0; a < TM; ++a) af[a] = As[cur][wm * 16 * TM + a * 16 + li][kk + lk];
#pragma unroll
            for (int b = 0; b < TN; ++b) bf[b] = Bs[cur][kk + lk][wn * 16 * TN + b * 16 + li];
            if constexpr (M3) {
                T as[TM], bs[TN];
#pragma unroll
                for (int a = 0; a < TM; ++a) as[a] = af[a].x + af[a].y;
#pragma unroll
                for (int b = 0; b < TN; ++b) bs[b] = bf[b].x + bf[b].y;
#pragma unroll
                for (int a = 0; a < TM; ++a)
#pragma unroll
                    for (int b = 0; b < TN; ++b) {
                        acc[0][a][b] = mfma16<T>(af[a].x, bf[b].x, acc[0][a][b]);
                        acc[1][a][b] = mfma16<T>(af[a].y, bf[b].y, acc[1][a][b]);
                        acc[2][a][b] = mfma16<T>(as[a], bs[b], acc[2][a][b]);
                    }
            } else {
#pragma unroll
                for (int a = 0; a < TM; ++a)
#pragma unroll
                    for (int b = 0; b < TN; ++b) {
                        acc[0][a][b] = mfma16<T>(af[a].x, bf[b].x, acc[0][a][b]);
                        acc[1][a][b] = mfma16<T>(af[a].x, bf[b].y, acc[1][a][b]);
                    }
#pragma unroll
                for (int a = 0; a < TM; ++a)
#pragma unroll
                    for (int b = 0; b < TN; ++b) {
                        acc[0][a][b] = mfma16<T>(-af[a].y, bf[b].y, acc[0][a][b]);
                        acc[1][a][b] = mfma16<T>(af[a].y, bf[b].x, acc[1][a][b]);
                    }
            }
        }
        if (more) store(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }
    // output row of register r in lane l (column l % 16): the f32 16x16x4 MFMA
    // keeps 4 consecutive rows per lane group (4 (l/16) + r), the f64 one
    // interleaves them (4 r + l/16)
    // all C reads are issued before any write, so they overlap (one memory
    // round trip per tile instead of one per element)
    V cv[TM][TN][4];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ri = sizeof(T) == 4 ? 4 * (lane >> 4) + r : 4 * r + (lane >> 4);
                const long long gi = m0 + wm * 16 * TM + a * 16 + ri;
                const long long gj = n0 + wn * 16 * TN + b * 16 + (lane & 15);
                cv[a][b][r] = (gi < M && gj < N) ? Cp[gi * ldc + gj] : V{0, 0};
            }
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ri = sizeof(T) == 4 ? 4 * (lane >> 4) + r : 4 * r + (lane >> 4);
                const long long gi = m0 + wm * 16 * TM + a * 16 + ri;
                const long long gj = n0 + wn * 16 * TN + b * 16 + (lane & 15);
                if (gi < M && gj < N) {
                    T re, im;
                    if constexpr (M3) {
                        re = acc[0][a][b][r] - acc[1][a][b][r];
                        im = acc[2][a][b][r] - acc[0][a][b][r] - acc[1][a][b][r];
                    } else {
                        re = acc[0][a][b][r];
                        im = acc[1][a][b][r];
                    }
                    V c = cv[a][b][r];
                    c.x -= re;
                    c.y -= im;
                    Cp[gi * ldc + gj] = c;
                }
            }
}

template <typename T, int TM, int TN, bool M3, int WPE>
__global__ __launch_bounds__(256, WPE) void k_cgemm(const typename CT<T>::t *__restrict__ Ap, long long lda,
                                                    const typename CT<T>::t *__restrict__ Bp, long long ldb,
                                                    typename CT<T>::t *Cp, long long ldc, int M, int N, int K)
{
    cgemm_tile<T, TM, TN, M3>(blockIdx.x, Ap, lda, Bp, ldb, Cp, ldc, M, N, K);
}

// ---------------------------------------------------------------- split-bf16 GEMM
// Mixed-path trailing update C -= A B with FP32 complex C.  Each FP32 value x is
// split into bf16 hi = bf16(x) and lo = bf16(x - hi); a real product is taken as
// hi*hi + hi*lo + lo*hi (the lo*lo term is below 2^-16 relative), i.e. three
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation, about 16 significant bits per
// product.  The FP32 LU that GMRES-IR refines tolerates this.  A and B are split
// once per panel step into planes (Re hi, Re lo, Im hi, Im lo), 32-deep k slices
// contiguous, so the GEMM streams bf16 through LDS and does no conversions.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr int kSplitBK = 32;
// LDS rows of a stage are unpadded (64 B); the 16-byte k group g of row r is stored at
// g ^ swz(r), which keeps both the staging writes (whole rows) and the fragment reads
// (16 rows, one k group) free of bank conflicts with 32 or 64 banks
__device__ __forceinline__ int split_swz(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo)
{
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// Plane layout: P[4][K / 32][rows][32], so that a k-step's 32-deep slice of a tile's rows
// is one contiguous run per plane (measured 1.3% faster products than P[4][rows][K]).
// A (rows x K, row-major complex, lda) -> P[4][K / 32][rows][32] bf16
__device__ __forceinline__ void split_rows_body(const float2 *__restrict__ A, long long lda, int rows, int K,
                                                __bf16 *__restrict__ P, long long idx)
{
    if (idx >= (long long)rows * K) return;
    const long long r = idx / K, k = idx % K;
    const float2 v = A[r * lda + k];
    const long long plane = (long long)rows * K;
    const long long o = ((k >> 5) * rows + r) * 32 + (k & 31);
    split_bf16(v.x, P[o], P[plane + o]);
    split_bf16(v.y, P[2 * plane + o], P[3 * plane + o]);
}

__global__ void k_split_rows(const float2 *__restrict__ A, long long lda, int rows, int K, __bf16 *__restrict__ P)
{
    split_rows_body(A, lda, rows, K, P, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// B (K x cols, row-major complex, ldb) -> P[4][K / 32][cols][32] bf16 (transposed)
__device__ __forceinline__ void split_cols_body(const float2 *__restrict__ B, long long ldb, int K, int cols,
                                                __bf16 *__restrict__ P, int bx, int by)
{
    __shared__ float2 t[32][33];
    const int c0 = bx * 32, k0 = by * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int k = k0 + r, c = c0 + tx;
        t[r][tx] = (k < K && c < cols) ? B[(long long)k * ldb + c] : float2{0, 0};
    }
    __syncthreads();
    const long long plane = (long long)cols * K;
    for (int r = ty; r < 32; r += 8) {
        const int c = c0 + r, k = k0 + tx;
        if (c < cols && k < K) {
            const float2 v = t[tx][r];
            const long long o = ((long long)(k >> 5) * cols + c) * 32 + (k & 31);
            split_bf16(v.x, P[o], P[plane + o]);
            split_bf16(v.y, P[2 * plane + o], P[3 * plane + o]);
        }
    }
}

__global__ __launch_bounds__(256) void k_split_cols(const float2 *__restrict__ B, long long ldb, int K, int cols,
                                                    __bf16 *__restrict__ P)
{
    split_cols_body(B, ldb, K, cols, P, blockIdx.x, blockIdx.y);
}

// C is read and written once per tile with nontemporal accesses, so that the tile
// rows of C streaming through do not evict the A and B planes, which the other
// tiles of the launch re-read, from L2.
__device__ __forceinline__ float2 ntload2(const float2 *p)
{
    return float2{__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y)};
}

__device__ __forceinline__ void ntstore2(float2 *p, float2 v)
{
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

// C[M x N] -= A B from split planes PA[4][K/32][M][32], PB[4][K/32][N][32]; K a multiple of 32.
// WM x WN waves, wave tile (16 TM) x (16 TN) complex.  Two LDS stages (128 KB for the
// 128 x 128 tile, unpadded rows with swizzled k groups): the global loads of step k+1 are in flight while step k's products run
// from the other stage, one barrier per step; the first row block of the C tile is read
// during the last step's products.
template <int TM, int TN, int WM, int WN>
__device__ __forceinline__ void cgemm_bf16x3_body(const __bf16 *__restrict__ PA, const __bf16 *__restrict__ PB,
                                                  float2 *C, long long ldc, int M, int N, int K, int tile)
{
    constexpr int NT = 64 * WM * WN;
    constexpr int BM = 16 * TM * WM, BN = 16 * TN * WN, BK = kSplitBK, RS = BK;
    __shared__ __align__(16) __bf16 As[2][4][BM][RS];
    __shared__ __align__(16) __bf16 Bs[2][4][BN][RS];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv / WN, wn = wv % WN;
    int tm, tn;
    gemm_tile(tile, (M + BM - 1) / BM, (N + BN - 1) / BN, tm, tn);
    const long long m0 = (long long)tm * BM, n0 = (long long)tn * BN;
    const long long planeA = (long long)M * K, planeB = (long long)N * K;
    // 16-byte chunks (8 bf16) per chunk of K: A has 4 * BM * 4, B has 4 * BN * 4
    constexpr int ACH = 4 * BM * (BK / 8), BCH = 4 * BN * (BK / 8);
    constexpr int AP = (ACH + NT - 1) / NT, BP = (BCH + NT - 1) / NT;
    uint4 ra[AP], rb[BP];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < AP; ++q) {
            const int c = tid + NT * q;
            if (c < ACH) {
                const int p = c / (BM * 4), rem = c % (BM * 4), r = rem / 4, g = rem % 4;
                const long long row = m0 + r;
                ra[q] = row < M ? *(const uint4 *)(PA + p * planeA + ((long long)(k0 >> 5) * M + row) * 32 + 8 * g)
                                : uint4{0, 0, 0, 0};
            }
        }
#pragma unroll
        for (int q = 0; q < BP; ++q) {
            const int c = tid + NT * q;
            if (c < BCH) {
                const int p = c / (BN * 4), rem = c % (BN * 4), r = rem / 4, g = rem % 4;
                const long long col = n0 + r;
                rb[q] = col < N ? *(const uint4 *)(PB + p * planeB + ((long long)(k0 >> 5) * N + col) * 32 + 8 * g)
                                : uint4{0, 0, 0, 0};
            }
        }
    };
    auto store = [&](int st) {
#pragma unroll
        for (int q = 0; q < AP; ++q) {
            const int c = tid + NT * q;
            if (c < ACH) {
                const int p = c / (BM * 4), rem = c % (BM * 4), r = rem / 4, g = rem % 4;
                *(uint4 *)&As[st][p][r][8 * (g ^ split_swz(r))] = ra[q];
            }
        }
#pragma unroll
        for (int q = 0; q < BP; ++q) {
            const int c = tid + NT * q;
            if (c < BCH) {
                const int p = c / (BN * 4), rem = c % (BN * 4), r = rem / 4, g = rem % 4;
                *(uint4 *)&Bs[st][p][r][8 * (g ^ split_swz(r))] = rb[q];
            }
        }
    };
    // real and imaginary parts of the tile; the real part takes Re A Re B and
    // Im A (-Im B), the B fragments' imaginary planes negated by their sign bits
    f32x4 arr[TM][TN], ari[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) {
            arr[a][b] = f32x4{0, 0, 0, 0};
            ari[a][b] = f32x4{0, 0, 0, 0};
        }
    const int li = lane & 15, lg = lane >> 4;   // row/col within a 16-block, k group of 8
    const int kofs = 8 * (lg ^ split_swz(li));    // the k group's place in LDS (rows start at multiples of 16)
    // products of one k-step from LDS stage st
    auto compute = [&](int st) {
        bf16x8 afr[TM][4];
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int p = 0; p < 4; ++p) afr[a][p] = *(const bf16x8 *)&As[st][p][wm * 16 * TM + a * 16 + li][kofs];
#pragma unroll
        for (int b = 0; b < TN; ++b) {
            bf16x8 bfr[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) bfr[p] = *(const bf16x8 *)&Bs[st][p][wn * 16 * TN + b * 16 + li][kofs];
            const bf16x8 nb2 = __builtin_bit_cast(bf16x8, __builtin_bit_cast(u32x4, bfr[2]) ^ 0x80008000u);
            const bf16x8 nb3 = __builtin_bit_cast(bf16x8, __builtin_bit_cast(u32x4, bfr[3]) ^ 0x80008000u);
#pragma unroll
            for (int a = 0; a < TM; ++a) {
                arr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a][0], bfr[0], arr[a][b], 0, 0, 0);
                ari[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a][0], bfr[2], ari[a][b], 0, 0, 0);
                arr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a][2], nb2, arr[a][b], 0, 0, 0);
                ari[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a][2], bfr[0], ari[a][b], 0, 0, 0);
                arr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a][0], bfr[1], arr[a][b], 0, 0, 0);
                ari[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a][0], bfr[3], ari[a][b], 0, 0, 0);
                arr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a][2], nb3, arr[a][b], 0, 0, 0);
                ari[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a][2], bfr[1], ari[a][b], 0, 0, 0);
                arr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a][1], bfr[0], arr[a][b], 0, 0, 0);
                ari[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a][1], bfr[2], ari[a][b], 0, 0, 0);
                arr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a][3], nb2, arr[a][b], 0, 0, 0);
                ari[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[a][3], bfr[0], ari[a][b], 0, 0, 0);
            }
        }
    };
    auto cidx = [&](int a, int b, int r, long long &gi, long long &gj) {
        gi = m0 + wm * 16 * TM + a * 16 + 4 * (lane >> 4) + r;
        gj = n0 + wn * 16 * TN + b * 16 + (lane & 15);
    };
    load(0);
    store(0);
    __syncthreads();
    int st = 0;
    for (int k0 = 0; k0 + BK < K; k0 += BK, st ^= 1) {
        load(k0 + BK);
        compute(st);
        store(st ^ 1);   // the other stage was last read before the previous barrier
        __syncthreads();
    }
    // last k-step: the first row block of this wave's C tile is read while its products run
    float2 c0[TN][4];
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            long long gi, gj;
            cidx(0, b, r, gi, gj);
            c0[b][r] = (gi < M && gj < N) ? ntload2(C + gi * ldc + gj) : float2{0, 0};
        }
    compute(st);
    float2 cv[TM][TN][4];
#pragma unroll
    for (int a = 1; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                long long gi, gj;
                cidx(a, b, r, gi, gj);
                cv[a][b][r] = (gi < M && gj < N) ? ntload2(C + gi * ldc + gj) : float2{0, 0};
            }
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) cv[0][b][r] = c0[b][r];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                long long gi, gj;
                cidx(a, b, r, gi, gj);
                if (gi < M && gj < N) {
                    float2 c = cv[a][b][r];
                    c.x -= arr[a][b][r];
                    c.y -= ari[a][b][r];
                    ntstore2(C + gi * ldc + gj, c);
                }
            }
}

template <int TM, int TN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN, 2) void k_cgemm_bf16x3(const __bf16 *__restrict__ PA,
                                                                 const __bf16 *__restrict__ PB, float2 *C,
                                                                 long long ldc, int M, int N, int K)
{
    cgemm_bf16x3_body<TM, TN, WM, WN>(PA, PB, C, ldc, M, N, K, blockIdx.x);
}

// ---------------------------------------------------------------- conversions
template <typename T>
__global__ void k_to_work(const double2 *__restrict__ src, int n, typename CT<T>::t *dst, long long lda)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * n) return;
    const long long i = idx / n, j = idx % n;
    const double2 v = src[idx];
    dst[i * lda + j] = {(T)v.x, (T)v.y};
}

// ---------------------------------------------------------------- solves
// X is n x nrhs row-major (rhs contiguous per row).
template <typename T>
__global__ void k_permute_in(const double2 *__restrict__ b, int n, int nrhs, const int *__restrict__ perm,
                             typename CT<T>::t *x)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * nrhs) return;
    const int i = (int)(idx / nrhs), r = (int)(idx % nrhs);
    const double2 v = b[(long long)perm[i] * nrhs + r];
    x[idx] = {(T)v.x, (T)v.y};
}

constexpr int kTB = 64;   // triangular-solve block
constexpr int kMaxRhs = 8;

// Solve the diagonal block `blk` on xs (LDS, kTB x nrhs, holding the updated
// right-hand side of the block): unit lower (forward) or upper (backward).
// The block is first staged in LDS (coalesced), so the serial row sweep runs at
// LDS latency.  All 256 threads call it; wave 0 does the sweep.
template <typename T, bool LOWER>
__device__ void trsv_diag(const typename CT<T>::t *A, long long lda, int n, typename CT<T>::t *xs, int nrhs, int blk,
                          typename CT<T>::t (*D)[kTB + 1])
{
    using V = typename CT<T>::t;
    const int lane = threadIdx.x & 63;
    const int b0 = blk * kTB, bs = min(kTB, n - b0);
    for (int idx = threadIdx.x; idx < bs * kTB; idx += blockDim.x) {
        const int i = idx / kTB, j = idx % kTB;
        D[i][j] = j < bs ? A[(long long)(b0 + i) * lda + b0 + j] : V{0, 0};
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        for (int r = 0; r < nrhs; ++r) {
            for (int s = 0; s < bs; ++s) {
                const int i = LOWER ? s : bs - 1 - s;
                V prod{0, 0};
                const bool use = LOWER ? (lane < i) : (lane > i && lane < bs);
                if (use) prod = cmulv(D[i][lane], xs[lane * nrhs + r]);
                for (int off = 32; off >= 1; off >>= 1) {
                    prod.x += __shfl_xor(prod.x, off, 64);
                    prod.y += __shfl_xor(prod.y, off, 64);
                }
                if (lane == 0) {
                    V v = csubv(xs[i * nrhs + r], prod);
                    if (!LOWER) {
                        const V d = D[i][i];
                        const T den = d.x * d.x + d.y * d.y;
                        v = {(v.x * d.x + v.y * d.y) / den, (v.y * d.x - v.x * d.y) / den};
                    }
                    xs[i * nrhs + r] = v;
                }
                // LDS accesses of one wave complete in order; the shuffle above already
                // orders lane 0's read after every lane's product
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
        }
    }
    __syncthreads();
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void k_trsv_first(const typename CT<T>::t *A, long long lda, int n,
                                                    typename CT<T>::t *X, int nrhs, int blk)
{
    using V = typename CT<T>::t;
    __shared__ V xs[kTB * kMaxRhs];
    __shared__ V D[kTB][kTB + 1];
    const int b0 = blk * kTB, bs = min(kTB, n - b0);
    for (int t = threadIdx.x; t < bs * nrhs; t += blockDim.x) xs[t] = X[(long long)b0 * nrhs + t];
    __syncthreads();
    trsv_diag<T, LOWER>(A, lda, n, xs, nrhs, blk, D);
    for (int t = threadIdx.x; t < bs * nrhs; t += blockDim.x) X[(long long)b0 * nrhs + t] = xs[t];
}

// Update the rows of the blocks after (LOWER) / before (upper) block `blk` with
// its solved values; workgroup 0 owns the next block, keeps its rows in LDS and
// then solves that diagonal block.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void k_trsv_step(const typename CT<T>::t *A, long long lda, int n,
                                                   typename CT<T>::t *X, int nrhs, int blk, int nblk)
{
    using V = typename CT<T>::t;
    __shared__ V xs[kTB * kMaxRhs];
    __shared__ V xb[kTB * kMaxRhs];
    __shared__ V D[kTB][kTB + 1];
    const int next = LOWER ? blk + 1 : blk - 1;
    const int ch = LOWER ? next + (int)blockIdx.x : next - (int)blockIdx.x;
    if (ch < 0 || ch >= nblk) return;   // uniform per workgroup
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c0 = blk * kTB, cs = min(kTB, n - c0);
    const int r0 = ch * kTB, rs = min(kTB, n - r0);
    for (int t = threadIdx.x; t < cs * nrhs; t += blockDim.x) xb[t] = X[(long long)c0 * nrhs + t];
    for (int t = threadIdx.x; t < rs * nrhs; t += blockDim.x) xs[t] = X[(long long)r0 * nrhs + t];
    __syncthreads();
    for (int i = wv; i < rs; i += 4) {
        const V a = lane < cs ? A[(long long)(r0 + i) * lda + c0 + lane] : V{0, 0};
        for (int r = 0; r < nrhs; ++r) {
            V prod = lane < cs ? cmulv(a, xb[lane * nrhs + r]) : V{0, 0};
            for (int off = 32; off >= 1; off >>= 1) {
                prod.x += __shfl_xor(prod.x, off, 64);
                prod.y += __shfl_xor(prod.y, off, 64);
            }
            if (lane == 0) xs[i * nrhs + r] = csubv(xs[i * nrhs + r], prod);
        }
    }
    __syncthreads();
    if (blockIdx.x == 0) trsv_diag<T, LOWER>(A, lda, n, xs, nrhs, next, D);
    for (int t = threadIdx.x; t < rs * nrhs; t += blockDim.x) X[(long long)r0 * nrhs + t] = xs[t];
}

// ---------------------------------------------------------------- FP32 solves with inverted diagonal blocks
// The FP32 factors only precondition GMRES-IR, so their triangular solves use
// explicit inverses of the 128 x 128 diagonal blocks (computed once per
// factorisation): each block step is then two small matrix-vector products with
// no serial sweep, one launch per block.
constexpr int kIB = 128;
constexpr int kTrinvParts = 4;   // lanes per column in k_trinv

// Inverse of the unit-lower 128 x 128 block held in S (LDS, zero above the diagonal,
// in place), by 32 x 32 blocks: the four diagonal blocks by forward substitution, a
// thread per column with the column in registers (no barrier); then block row I = 1..3
// as S_IJ = -S_II sum_{K=J..I-1} L_IK S_KJ (J < I), the sums first (into T, while
// L_IJ is still in place), then the products with S_II.  512 threads.
constexpr int kIB32 = 32;   // block of the blocked inverse
__device__ __forceinline__ void unit_lower_inverse_blocked(float2 *S, float2 (*T)[kIB32][kIB32])
{
    const int t = threadIdx.x;
    float2 x[kIB32];
    const int blk = t / kIB32, c = t % kIB32;
    if (t < kIB) {
        // column c of the inverse of diagonal block blk: x_r = [r == c] - sum_{k<r} L[r][k] x_k
        const float2 *L = S + (size_t)(kIB32 * blk) * kIB + kIB32 * blk;
#pragma unroll
        for (int r = 0; r < kIB32; ++r) {
            float2 acc{0, 0};
#pragma unroll
            for (int k = 0; k < r; ++k) {
                const float2 a = L[r * kIB + k];
                acc.x = fmaf(a.x, x[k].x, fmaf(-a.y, x[k].y, acc.x));
                acc.y = fmaf(a.x, x[k].y, fmaf(a.y, x[k].x, acc.y));
            }
            x[r] = r < c ? float2{0, 0} : (r == c ? float2{1, 0} : float2{-acc.x, -acc.y});
        }
    }
    __syncthreads();   // every diagonal block has been read
    if (t < kIB) {
        float2 *L = S + (size_t)(kIB32 * blk) * kIB + kIB32 * blk;
#pragma unroll
        for (int r = 0; r < kIB32; ++r) L[r * kIB + c] = x[r];
    }
    __syncthreads();
    for (int I = 1; I < kIB / kIB32; ++I) {
        const float2 *rowI = S + (size_t)(kIB32 * I) * kIB;
        for (int e = t; e < I * kIB32 * kIB32; e += kIB * kTrinvParts) {
            const int J = e / (kIB32 * kIB32), r = e / kIB32 % kIB32, cc = e % kIB32;
            float2 acc{0, 0};
#pragma unroll 16
            for (int k = kIB32 * J; k < kIB32 * I; ++k) {
                const float2 a = rowI[r * kIB + k], b = S[(size_t)k * kIB + kIB32 * J + cc];
                acc.x = fmaf(a.x, b.x, fmaf(-a.y, b.y, acc.x));
                acc.y = fmaf(a.x, b.y, fmaf(a.y, b.x, acc.y));
            }
            T[J][r][cc] = acc;
        }
        __syncthreads();
        for (int e = t; e < I * kIB32 * kIB32; e += kIB * kTrinvParts) {
            const int J = e / (kIB32 * kIB32), r = e / kIB32 % kIB32, cc = e % kIB32;
            float2 acc{0, 0};
#pragma unroll 8
            for (int k = 0; k <= r; ++k) {
                const float2 a = rowI[r * kIB + kIB32 * I + k], b = T[J][k][cc];
                acc.x = fmaf(a.x, b.x, fmaf(-a.y, b.y, acc.x));
                acc.y = fmaf(a.x, b.y, fmaf(a.y, b.x, acc.y));
            }
            S[(size_t)(kIB32 * I + r) * kIB + kIB32 * J + cc] = float2{-acc.x, -acc.y};
        }
        __syncthreads();
    }
}

// In-place inverse of diagonal block blockIdx.x of the unit-lower (LOWER) or
// upper factor in LDS; rows past the end of the matrix become identity rows.
// LOWER: blocked (unit_lower_inverse_blocked).  Upper: row by row,
// X[i][t] = ([i == t] - sum_k T[i][k] X[k][t]) / T[i][i] over the rows k already
// inverted (entries outside the triangle are zero), kTrinvParts threads per column
// sharing each row's dot product (partial sums combined through LDS).
template <bool LOWER>
__device__ __forceinline__ void trinv_body(const float2 *A, long long lda, int n, float2 *Dinv, int bid)
{
    __shared__ float2 S[kIB * kIB];
    const int t = threadIdx.x, col = t % kIB, part = t / kIB;
    const int b0 = bid * kIB, bs = min(kIB, n - b0);
    for (int idx = t; idx < kIB * kIB; idx += kIB * kTrinvParts) {
        const int i = idx / kIB, j = idx % kIB;
        float2 v{0, 0};
        if (i < bs && j < bs && (LOWER ? j < i : j >= i)) v = A[(long long)(b0 + i) * lda + b0 + j];
        if (i == j && (LOWER || i >= bs)) v = {1, 0};
        S[idx] = v;
    }
    __syncthreads();
    if constexpr (LOWER) {
        __shared__ float2 T[kIB / kIB32 - 1][kIB32][kIB32];
        unit_lower_inverse_blocked(S, T);
        float2 *out = Dinv + (size_t)bid * kIB * kIB;
        for (int idx = t; idx < kIB * kIB; idx += kIB * kTrinvParts) out[idx] = S[idx];
        return;
    }
    __shared__ float2 red[kTrinvParts][kIB];
    // part = t / kIB: the waves of a part read whole rows of S (consecutive columns,
    // no bank conflicts); part p sums the terms k = p mod kTrinvParts of each dot product
    for (int s = LOWER ? 1 : 0; s < kIB; ++s) {
        const int i = LOWER ? s : kIB - 1 - s;
        const int k0 = LOWER ? part : i + 1 + part, k1 = LOWER ? i : kIB;
        float2 u{0, 0}, w{0, 0};   // two chains, so consecutive terms overlap
        int k = k0;
        for (; k + kTrinvParts < k1; k += 2 * kTrinvParts) {
            const float2 a = S[i * kIB + k], x = S[k * kIB + col];
            const float2 b = S[i * kIB + k + kTrinvParts], y = S[(k + kTrinvParts) * kIB + col];
            u.x += a.x * x.x - a.y * x.y;
            u.y += a.x * x.y + a.y * x.x;
            w.x += b.x * y.x - b.y * y.y;
            w.y += b.x * y.y + b.y * y.x;
        }
        if (k < k1) {
            const float2 a = S[i * kIB + k], x = S[k * kIB + col];
            u.x += a.x * x.x - a.y * x.y;
            u.y += a.x * x.y + a.y * x.x;
        }
        red[part][col] = {u.x + w.x, u.y + w.y};
        const float2 d = S[i * kIB + i];
        __syncthreads();
        if (part == 0) {
            float2 acc = red[0][col];
#pragma unroll
            for (int p = 1; p < kTrinvParts; ++p) { acc.x += red[p][col].x; acc.y += red[p][col].y; }
            if (LOWER) {
                if (col < i) S[i * kIB + col] = {-acc.x, -acc.y};
            } else {
                const float den = d.x * d.x + d.y * d.y;
                const float2 inv{d.x / den, -d.y / den};
                if (col > i) S[i * kIB + col] = {-(acc.x * inv.x - acc.y * inv.y), -(acc.x * inv.y + acc.y * inv.x)};
                else if (col == i) S[i * kIB + i] = inv;
            }
        }
        __syncthreads();
    }
    float2 *out = Dinv + (size_t)bid * kIB * kIB;
    for (int idx = t; idx < kIB * kIB; idx += kIB * kTrinvParts) out[idx] = S[idx];
}

template <bool LOWER>
__global__ __launch_bounds__(kIB * kTrinvParts) void k_trinv(const float2 *A, long long lda, int n, float2 *Dinv)
{
    trinv_body<LOWER>(A, lda, n, Dinv, blockIdx.x);
}

// out[i] = (SUB ? out[i] : 0) - / + sum_c Mx[i][c] xin[c] for rows i < rs, columns
// c < cs (xin, out: distinct LDS vectors, row-major x nrhs).  The kTrsvT threads of a
// workgroup hold the 128 x 128 block in registers, loaded at once (one memory round
// trip per block instead of one per entry): thread t (wave w, lane l) takes the rows
// (kTrsvT / 8) p + 8 w + l / 8 (p < kBlkRows) and of each the 16-byte chunks
// l % 8 + 8 j (j < 8), i.e. the columns 2 (l % 8 + 8 j) and the one after.  Eight
// lanes read one 128-byte line of a row, so each load of a wave touches 8 cache
// lines (a layout with a row per lane touches 64, and the vector memory pipeline
// then takes four times as long over the block); the eight lanes of a row then sum
// their parts by DPP within the half-row.
constexpr int kTrsvT = 512, kBlkRowStep = kTrsvT / 8, kBlkRows = kIB / kBlkRowStep, kBlkChunks = 8;
static_assert(kBlkRows * kBlkRowStep == kIB, "rows of a block split evenly over the threads");
using BlkRegs = float4[kBlkRows][kBlkChunks];

__device__ __forceinline__ void blk_load(const float2 *Mx, long long ld, int rs, int cs, BlkRegs &a)
{
    const int t = threadIdx.x, w = t >> 6, l = t & 63, rl = 8 * w + (l >> 3), c8 = l & 7;
    if (rs == kIB && cs == kIB) {
#pragma unroll
        for (int p = 0; p < kBlkRows; ++p) {
            const float4 *row4 = (const float4 *)(Mx + (long long)(kBlkRowStep * p + rl) * ld);   // 16-byte aligned: ld and block origins are multiples of 64
#pragma unroll
            for (int j = 0; j < kBlkChunks; ++j) a[p][j] = row4[c8 + 8 * j];
        }
    } else {
#pragma unroll
        for (int p = 0; p < kBlkRows; ++p) {
            const int r = kBlkRowStep * p + rl;
            const float2 *row = Mx + (long long)min(r, max(rs - 1, 0)) * ld;
#pragma unroll
            for (int j = 0; j < kBlkChunks; ++j) {
                const int c = 2 * (c8 + 8 * j);
                const float2 u = (r < rs && c < cs) ? row[c] : float2{0, 0};
                const float2 v = (r < rs && c + 1 < cs) ? row[c + 1] : float2{0, 0};
                a[p][j] = float4{u.x, u.y, v.x, v.y};
            }
        }
    }
}

// Sum over the eight lanes 8 m .. 8 m + 7 of a value: DPP quad_perm 1,0,3,2 and
// 2,3,0,1 make each quad uniform, then the half-row mirror adds the other quad.
__device__ __forceinline__ float sum8_lanes(float v)
{
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, v), __builtin_bit_cast(int, v),
                                                               0xB1, 0xF, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, v), __builtin_bit_cast(int, v),
                                                               0x4E, 0xF, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, v), __builtin_bit_cast(int, v),
                                                               0x141, 0xF, 0xF, false));
    return v;
}

template <bool SUB, int NRHS>
__device__ __forceinline__ void blk_apply(const BlkRegs &a, int rs, const float2 *xin, float2 *out)
{
    const int t = threadIdx.x, w = t >> 6, l = t & 63, rl = 8 * w + (l >> 3), c8 = l & 7;
#pragma unroll
    for (int q = 0; q < NRHS; ++q) {
        float2 acc[kBlkRows];
#pragma unroll
        for (int p = 0; p < kBlkRows; ++p) acc[p] = {0, 0};
#pragma unroll
        for (int j = 0; j < kBlkChunks; ++j) {
            const int c = 2 * (c8 + 8 * j);
            const float2 x0 = xin[c * NRHS + q], x1 = xin[(c + 1) * NRHS + q];
#pragma unroll
            for (int p = 0; p < kBlkRows; ++p) {   // explicit FMAs: the same bits for every NRHS
                const float4 m = a[p][j];
                acc[p].x = fmaf(-m.w, x1.y, fmaf(m.z, x1.x, fmaf(-m.y, x0.y, fmaf(m.x, x0.x, acc[p].x))));
                acc[p].y = fmaf(m.w, x1.x, fmaf(m.z, x1.y, fmaf(m.y, x0.x, fmaf(m.x, x0.y, acc[p].y))));
            }
        }
#pragma unroll
        for (int p = 0; p < kBlkRows; ++p) {
            const float2 v{sum8_lanes(acc[p].x), sum8_lanes(acc[p].y)};
            const int i = kBlkRowStep * p + rl;
            if (c8 == 0 && i < rs) {
                float2 &o = out[i * NRHS + q];
                o = SUB ? float2{o.x - v.x, o.y - v.y} : v;
            }
        }
    }
    __syncthreads();
}

// G[i] = Dinv[i] M[i][i -/+ 1] for the blocks with a neighbour (LOWER: the
// block left of the diagonal of L, else the block right of the diagonal of U),
// so the solve step that waits for the neighbour is one product.  Workgroup
// (i, y) forms rows [16 y, 16 y + 16) of G[i]; thread t: row t / 16, columns
// 8 (t % 16) .. + 8.
template <bool LOWER>
__device__ __forceinline__ void blk_coupling_body(const float2 *A, long long lda, int n, const float2 *Dinv, float2 *G,
                                                  int bx, int by)
{
    __shared__ float2 Ds[16][kIB];
    __shared__ float2 Ms[32][kIB];
    const int nb = (n + kIB - 1) / kIB;
    const int i = bx, j = LOWER ? i - 1 : i + 1;
    if (j < 0 || j >= nb) return;   // uniform
    const int r0 = i * kIB, rs = min(kIB, n - r0), c0 = j * kIB, cs = min(kIB, n - c0);
    const int tid = threadIdx.x, rr = tid >> 4, cg = (tid & 15) * 8, y0 = by * 16;
    for (int idx = tid; idx < 16 * kIB; idx += 256) {
        const int r = idx / kIB, k = idx % kIB;
        Ds[r][k] = (y0 + r < rs && k < rs) ? Dinv[(size_t)i * kIB * kIB + (size_t)(y0 + r) * kIB + k] : float2{0, 0};
    }
    float2 acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = {0, 0};
    for (int k0 = 0; k0 < kIB; k0 += 32) {
        __syncthreads();
        for (int idx = tid; idx < 32 * kIB; idx += 256) {
            const int k = idx / kIB, c = idx % kIB;
            Ms[k][c] = (k0 + k < rs && c < cs) ? A[(long long)(r0 + k0 + k) * lda + c0 + c] : float2{0, 0};
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < 32; ++k) {
            const float2 d = Ds[rr][k0 + k];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const float2 m = Ms[k][cg + c];
                acc[c].x = fmaf(d.x, m.x, fmaf(-d.y, m.y, acc[c].x));
                acc[c].y = fmaf(d.x, m.y, fmaf(d.y, m.x, acc[c].y));
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) G[(size_t)i * kIB * kIB + (size_t)(y0 + rr) * kIB + cg + c] = acc[c];
}

template <bool LOWER>
__global__ __launch_bounds__(256) void k_blk_coupling(const float2 *A, long long lda, int n, const float2 *Dinv,
                                                      float2 *G)
{
    blk_coupling_body<LOWER>(A, lda, n, Dinv, G, blockIdx.x, blockIdx.y);
}

// One launch per triangle, one workgroup per CU.  A workgroup takes a block in the
// order of a ticket counter, solves it and takes the next, until none is left (so a
// block never waits for a workgroup that has not started: every block it waits for
// has an earlier ticket, held by a running workgroup).  Block i subtracts the
// products with the blocks it depends on as their ready flags appear, applies its
// inverted diagonal block and publishes its rows with a ready flag (agent scope
// release / acquire).  A wait that exceeds kTrsvSpin polls sets the error word and
// gives up, so every workgroup ends.  Tickets handed out per launch: the blocks plus
// one final ticket per workgroup.
constexpr long long kTrsvSpin = 1LL << 24;

template <bool LOWER, int NRHS>
__global__ __launch_bounds__(kTrsvT) void k_trsv_flag(const float2 *A, long long lda, int n, const float2 *Dinv,
                                                      const float2 *G, float2 *X, int *sync, int ticket0, int epoch)
{
    __shared__ float2 xs[kIB * NRHS];
    __shared__ float2 xb[kIB * NRHS];
    // the inverted diagonal block, each thread's share at a place of its own (read back
    // by the thread that wrote it: no barrier, no bank conflicts)
    __shared__ float4 dsh[kBlkRows][kBlkChunks][kTrsvT];
    __shared__ int s_idx, s_ready;
    int *counter = sync, *err = sync + 1, *flag = sync + 2;
    const int nblk = (n + kIB - 1) / kIB;
    for (;;) {
        __syncthreads();   // the previous block's LDS reads are done
        if (threadIdx.x == 0)
            s_idx = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - ticket0;
        __syncthreads();
        const int idx = s_idx;
        if (idx < 0 || idx >= nblk) return;   // no block left
        const int blk = LOWER ? idx : nblk - 1 - idx;
        const int r0 = blk * kIB, rs = min(kIB, n - r0);
        const int jadj = LOWER ? blk - 1 : blk + 1, cadj = jadj * kIB, csadj = min(kIB, n - cadj);
        auto wait_flag = [&](int j) {   // thread 0
            long long it = 0;
            while (__hip_atomic_load(flag + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
                if (++it > kTrsvSpin) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                // once a wait has given up, the others do not wait out their own limits
                if ((it & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                __builtin_amdgcn_s_sleep(1);
            }
        };
        // the block's own operands, held for the whole solve: the inverted diagonal block
        // (LDS) and the coupling G to the adjacent block (registers), so that neither is
        // loaded after a wait
        BlkRegs a, g;
        blk_load(Dinv + (size_t)blk * kIB * kIB, kIB, rs, rs, a);
        if (idx > 0) blk_load(G + (size_t)blk * kIB * kIB, kIB, rs, csadj, g);
#pragma unroll
        for (int p = 0; p < kBlkRows; ++p)
#pragma unroll
            for (int j = 0; j < kBlkChunks; ++j) dsh[p][j][threadIdx.x] = a[p][j];
        for (int t = threadIdx.x; t < kIB * NRHS; t += blockDim.x) xs[t] = t < rs * NRHS ? X[(long long)r0 * NRHS + t] : float2{0, 0};
        __syncthreads();
        // every dependency but the adjacent block, taken as their ready flags appear; the
        // next block of L (or U) is loaded before its flag is waited for
        const int early = max(idx - 1, 0);
        int done = 0;
        auto load_early = [&](int d) {
            const int c0 = (LOWER ? d : nblk - 1 - d) * kIB;
            blk_load(A + (long long)r0 * lda + c0, lda, rs, min(kIB, n - c0), a);
        };
        if (early > 0) load_early(0);
        while (done < early) {
            if (threadIdx.x == 0) {
                wait_flag(LOWER ? done : nblk - 1 - done);
                // blocks publish in order, so every further flag already set is usable too
                int ready = done + 1;
                while (ready < early &&
                       __hip_atomic_load(flag + (LOWER ? ready : nblk - 1 - ready), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT) == epoch)
                    ++ready;
                __threadfence();   // acquire: the rows of those blocks
                s_ready = ready;
            }
            __syncthreads();
            const int ready = s_ready;
            for (; done < ready; ++done) {
                const int j = LOWER ? done : nblk - 1 - done;
                const int c0 = j * kIB, cs = min(kIB, n - c0);
                for (int t = threadIdx.x; t < kIB * NRHS; t += blockDim.x)
                    xb[t] = t < cs * NRHS ? X[(long long)c0 * NRHS + t] : float2{0, 0};
                __syncthreads();
                blk_apply<true, NRHS>(a, rs, xb, xs);
                if (done + 1 < early) load_early(done + 1);
            }
        }
        // y = Dinv x without the adjacent block (into xb); then x = y - G x_adjacent
#pragma unroll
        for (int p = 0; p < kBlkRows; ++p)
#pragma unroll
            for (int j = 0; j < kBlkChunks; ++j) a[p][j] = dsh[p][j][threadIdx.x];
        blk_apply<false, NRHS>(a, rs, xs, xb);
        if (idx > 0) {
            if (threadIdx.x == 0) {
                wait_flag(jadj);
                __threadfence();   // acquire: the adjacent block's rows
            }
            __syncthreads();
            for (int t = threadIdx.x; t < kIB * NRHS; t += blockDim.x)
                xs[t] = t < csadj * NRHS ? X[(long long)cadj * NRHS + t] : float2{0, 0};
            __syncthreads();
            blk_apply<true, NRHS>(g, rs, xs, xb);
        }
        // wave 0 writes the rows, so that thread 0's release fence covers every store (a
        // barrier does not wait for other waves' stores, and a fence orders its own wave's)
        if (threadIdx.x < 64)
            for (int t = threadIdx.x; t < rs * NRHS; t += 64) X[(long long)r0 * NRHS + t] = xb[t];
        if (threadIdx.x == 0) {
            __threadfence();   // release: this block's rows
            __hip_atomic_store(flag + blk, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// r = b - A x in FP64 (A row-major n x n), one wave per row; x, b, r are n x nrhs row-major.
// r = b - A x (RESID) or r = A x, for nrhs columns (x, b, r row-major n x nrhs)
template <bool RESID>
__device__ __forceinline__ void apply_a_body(const double2 *__restrict__ A, int n, const double2 *__restrict__ x,
                                             const double2 *__restrict__ b, double2 *__restrict__ r, int nrhs)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const int lane = threadIdx.x & 63;
    const double2 *a = A + (size_t)row * n;
    double2 acc[kMaxRhs];
#pragma unroll
    for (int q = 0; q < kMaxRhs; ++q) acc[q] = {0, 0};
    auto step = [&](const double2 av, int cc) {
#pragma unroll
        for (int q = 0; q < kMaxRhs; ++q) {
            if (q < nrhs) {
                const double2 xv = x[(size_t)cc * nrhs + q];
                acc[q].x += av.x * xv.x - av.y * xv.y;
                acc[q].y += av.x * xv.y + av.y * xv.x;
            }
        }
    };
    int c = lane;
    // four row loads in flight per lane, summed in the same order as one at a time
    for (; c + 192 < n; c += 256) {
        double2 av[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) av[u] = a[c + 64 * u];
#pragma unroll
        for (int u = 0; u < 4; ++u) step(av[u], c + 64 * u);
    }
    for (; c < n; c += 64) step(a[c], c);
#pragma unroll
    for (int q = 0; q < kMaxRhs; ++q) {
        if (q >= nrhs) break;
        double sx = acc[q].x, sy = acc[q].y;
        for (int off = 32; off >= 1; off >>= 1) {
            sx += __shfl_xor(sx, off, 64);
            sy += __shfl_xor(sy, off, 64);
        }
        if (lane == 0) {
            if constexpr (RESID) {
                const double2 bv = b[(size_t)row * nrhs + q];
                r[(size_t)row * nrhs + q] = {bv.x - sx, bv.y - sy};
            } else {
                r[(size_t)row * nrhs + q] = {sx, sy};
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_residual(const double2 *__restrict__ A, int n, const double2 *__restrict__ x,
                                                  const double2 *__restrict__ b, double2 *__restrict__ r, int nrhs)
{
    apply_a_body<true>(A, n, x, b, r, nrhs);
}

__global__ __launch_bounds__(256) void k_matvec_cols(const double2 *__restrict__ A, int n,
                                                     const double2 *__restrict__ x, double2 *__restrict__ y, int nrhs)
{
    apply_a_body<false>(A, n, x, nullptr, y, nrhs);
}

// y = A x for one vector (FP64), used by GMRES.
// The same for q columns side by side (row-major n x q blocks): one Krylov space per
// column, all advanced by the same launches.
struct ColScale {
    double s[kMaxRhs];
};

__global__ void k_scale_cols(const double2 *__restrict__ src, int n, int q, ColScale sc, double2 *__restrict__ dst)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * q) return;
    const double f = sc.s[i % q];
    dst[i] = {src[i].x * f, src[i].y * f};
}

// out[block][j][c] = sum over the block's rows i of conj(V_j[i][c]) w[i][c], j < m, c < q
__global__ __launch_bounds__(256) void k_dots_cols(const double2 *__restrict__ Vb, int n, int m, int q,
                                                   const double2 *__restrict__ w, double2 *__restrict__ out)
{
    __shared__ double2 red[4][kMaxRhs];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j = 0; j < m; ++j) {
        const double2 *v = Vb + (size_t)j * n * q;
        double ax[kMaxRhs], ay[kMaxRhs];
#pragma unroll
        for (int c = 0; c < kMaxRhs; ++c) ax[c] = ay[c] = 0;
        for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
#pragma unroll
            for (int c = 0; c < kMaxRhs; ++c) {
                if (c < q) {
                    const double2 a = v[(size_t)i * q + c], b = w[(size_t)i * q + c];
                    ax[c] += a.x * b.x + a.y * b.y;
                    ay[c] += a.x * b.y - a.y * b.x;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < kMaxRhs; ++c) {
            if (c < q) {
                double sx = ax[c], sy = ay[c];
                for (int off = 32; off >= 1; off >>= 1) {
                    sx += __shfl_xor(sx, off, 64);
                    sy += __shfl_xor(sy, off, 64);
                }
                if (lane == 0) red[wid][c] = {sx, sy};
            }
        }
        __syncthreads();
        if ((int)threadIdx.x < q) {
            const int c = threadIdx.x;
            out[((size_t)blockIdx.x * m + j) * q + c] = {red[0][c].x + red[1][c].x + red[2][c].x + red[3][c].x,
                                                        red[0][c].y + red[1][c].y + red[2][c].y + red[3][c].y};
        }
        __syncthreads();
    }
}

// w[i][c] -= sum_j h[j][c] V_j[i][c]
__global__ void k_axpys_cols(const double2 *__restrict__ Vb, int n, int m, int q, const double2 *__restrict__ h,
                             double2 *w)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * q) return;
    const int c = (int)(i % q);
    double2 acc = w[i];
    for (int j = 0; j < m; ++j) {
        const double2 v = Vb[(size_t)j * n * q + i], cf = h[(size_t)j * q + c];
        acc.x -= cf.x * v.x - cf.y * v.y;
        acc.y -= cf.x * v.y + cf.y * v.x;
    }
    w[i] = acc;
}

template <typename T>
__global__ void k_gather_lo(const double2 *__restrict__ src, int n, const int *__restrict__ perm,
                            typename CT<T>::t *dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double2 v = src[perm[i]];
        dst[i] = {(T)v.x, (T)v.y};
    }
}

template <typename T>
__global__ void k_widen(const typename CT<T>::t *__restrict__ src, long long cnt, double2 *__restrict__ dst)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) dst[i] = {(double)src[i].x, (double)src[i].y};
}

__global__ void k_add(double2 *x, const double2 *__restrict__ d, long long cnt)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) { x[i].x += d[i].x; x[i].y += d[i].y; }
}

// ---------------------------------------------------------------- host driver
void check(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) BH_FAIL(BH_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
}

template <typename T>
struct LUWork {
    DevBuf<T> slot_val;
    DevBuf<int> slot_row;
    DevBuf<typename CT<T>::t> cand, diag;
    DevBuf<unsigned> bar;
    DevBuf<int> info;
    int maxG = 0;
    int rpb_max = 0;
    DevBuf<int> plan;                      // row-interchange plan (k_swap_plan)
    DevBuf<int> tcnt;                      // tournament group counters (zero between panels)
    DevBuf<int> tidx;                      // tournament candidate sets (row ids), all levels
    DevBuf<float2> tval;                   // their panel rows
    DevBuf<float2> top, oldtop;            // factored winners; panel rows before the moves
    DevBuf<int> tplan;                     // row moves of the last tournament
    DevBuf<int> plan_side;                 // row-interchange plan of the block-column stream
    DevBuf<__bf16> pa_side, pb_side;       // split planes of the block-column stream
    DevBuf<__bf16> pa_u12;                 // split planes of L for the block row
    DevBuf<float2> eblk, eblk_side;        // I - inv(L_bb) of the 128-row diagonal blocks of a block row
                                           // (context stream, block-column stream)
    int eblk_of = -1, eblk_side_of = -1;   // first row of the block row whose inverses they hold
};

template <typename T, int TM, int TN, bool M3, int WPE>
void gemm_launch(Ctx &ctx, hipStream_t s, typename CT<T>::t *A, long long lda, int ar, int ac, int br, int bc, int cr,
                 int cc, int M, int N, int K)
{
    constexpr int BM = 32 * TM, BN = 32 * TN;
    const long long tiles = (long long)((N + BN - 1) / BN) * ((M + BM - 1) / BM);
    BH_REQUIRE(tiles < (1LL << 31), "gemm grid too large");
    const dim3 grid((unsigned)tiles);
    ProfScope prof(ctx, "k_cgemm", 8.0 * M * N * K, s);   // real flops of the complex update
    k_cgemm<T, TM, TN, M3, WPE><<<grid, 256, 0, s>>>(A + (long long)ar * lda + ac, lda,
                                                               A + (long long)br * lda + bc, lda,
                                                               A + (long long)cr * lda + cc, lda, M, N, K);
    check("k_cgemm");
}

// Tile shapes chosen by measurement on MI355X (N = 12288 LU): FP32 3M products,
// 128 x 64 complex tiles, two waves per SIMD; FP64 4M, 64 x 64, two waves per SIMD.
// FP32 products of one tile row (M <= 128, the block rows' updates) take 128 x 32
// tiles at three waves per SIMD: twice the tiles, so a 40k-wide row of them fills
// the device in whole rounds (C3 block rows: 46.5 -> 42.0 ms of FP32 products; each
// entry's sum over k runs in the same order as with 128 x 64 tiles).
template <typename T>
void gemm(Ctx &ctx, hipStream_t s, typename CT<T>::t *A, long long lda, int ar, int ac, int br, int bc, int cr, int cc,
          int M, int N, int K)
{
    if (M <= 0 || N <= 0 || K <= 0) return;
    BH_REQUIRE(K % kBK == 0, "gemm K=%d not a multiple of %d", K, kBK);
    if constexpr (sizeof(T) == 4) {
        if (M <= 128) gemm_launch<T, 4, 1, true, 3>(ctx, s, A, lda, ar, ac, br, bc, cr, cc, M, N, K);
        else gemm_launch<T, 4, 2, true, 2>(ctx, s, A, lda, ar, ac, br, bc, cr, cc, M, N, K);
    } else {
        gemm_launch<T, 2, 2, false, 2>(ctx, s, A, lda, ar, ac, br, bc, cr, cc, M, N, K);
    }
}

template <typename T>
void laswp(Ctx &ctx, hipStream_t s, typename CT<T>::t *A, long long lda, int k1, int k2, const long long *ipiv, int c0,
           int c1, int c2, int c3, int *plan)
{
    const int w = std::max(0, c1 - c0) + std::max(0, c3 - c2);
    const int len = k2 - k1;
    if (w <= 0 || len <= 0) return;
    ProfScope prof(ctx, "k_laswp", 2.0 * 2 * len * w * sizeof(typename CT<T>::t), s);   // upper bound: every slot moves
    BH_REQUIRE(len <= kMaxSwaps, "laswp: %d interchanges exceed %d", len, kMaxSwaps);
    k_swap_plan<<<1, 256, 0, s>>>(ipiv, k1, k2, plan);
    check("k_swap_plan");
    constexpr int CW = laswp_cols<T>();
    const size_t shm = sizeof(T) == 4 ? 0 : (size_t)2 * len * CW * sizeof(typename CT<T>::t);
    k_laswp_apply<T><<<(w + CW - 1) / CW, 256, shm, s>>>(A, lda, plan, len, c0, std::max(c0, c1), c2, std::max(c2, c3));
    check("k_laswp_apply");
}

template <typename T>
void trsm_unit(hipStream_t s, typename CT<T>::t *A, long long lda, int lr, int m, int br, int bc, int ncols)
{
    if (ncols <= 0 || m <= 0) return;
    k_trsm_unit<T><<<(ncols + 63) / 64, 64, 0, s>>>(A, lda, lr, lr, m, br, bc, ncols);
    check("k_trsm_unit");
}

template <typename T>
void panel(Ctx &ctx, LUWork<T> &W, typename CT<T>::t *A, long long lda, int n, int k0, int pw, long long *ipiv)
{
    const int rows = n - k0;
    int G = (rows + W.rpb_max - 1) / W.rpb_max;
    G = std::max(1, std::min(G, W.maxG));
    const int rpb = (rows + G - 1) / G;
    BH_REQUIRE(rpb <= W.rpb_max, "panel: %d rows per workgroup exceed %d", rpb, W.rpb_max);
    PanelWork<T> pw_{W.slot_val.p, W.slot_row.p, W.cand.p, W.diag.p, W.bar.p, W.info.p};
    const size_t shm = (size_t)rpb * kPW * sizeof(typename CT<T>::t);
    void *args[] = {&A, (void *)&lda, (void *)&n, (void *)&k0, (void *)&pw, (void *)&rpb, &ipiv, &pw_};
    ProfScope prof(ctx, "k_panel", (double)rows * pw * sizeof(typename CT<T>::t) * 2);
    // A grid barrier needs all its workgroups resident at once.  Two cooperative
    // panels on different streams could each hold part of the device and wait
    // for the rest, so with several live contexts the panels run one at a time.
    if (live_contexts() > 1) {
        std::lock_guard<std::mutex> lock(coop_mutex());
        BH_CHECK_HIP(hipLaunchCooperativeKernel((const void *)k_panel<T>, dim3(G), dim3(256), args, shm, ctx.stream));
        BH_CHECK_HIP(hipStreamSynchronize(ctx.stream));
    } else {
        BH_CHECK_HIP(hipLaunchCooperativeKernel((const void *)k_panel<T>, dim3(G), dim3(256), args, shm, ctx.stream));
    }
}

// Panel of columns [c, c + pw) of the block column [k, k + kb) by tournament
// pivoting.  One launch applies the previous panel's update to the block's
// columns right of it (prev) and runs the tournament, which writes ipiv[c .. c+pw);
// the apply launch moves the rows in the block's other columns, finishes the
// panel rows and forms the panel's block row U12 right of the panel.
void tslu_panel(Ctx &ctx, hipStream_t s, LUWork<float> &W, float2 *A, long long lda, int n, int k, int kb, int c,
                int pw, long long *ipiv, bool prev)
{
    const TsluFinal fin{ipiv, W.top.p, W.oldtop.p, W.tplan.p, W.info.p};
    const TsluWork tw{W.tcnt.p, W.tidx.p, W.tval.p};
    {
        const int nsets0 = (n - c + kTsluT - 1) / kTsluT;
        const int upd_n = prev ? std::max(0, k + kb - (c + kPW)) : 0;
        const long long tiles = upd_n > 0 ? (long long)((n - c + 127) / 128) * ((upd_n + 63) / 64) : 0;
        BH_REQUIRE(nsets0 + tiles < (1LL << 31), "tournament grid too large");
        ProfScope prof(ctx, "k_tslu_panel", (double)(n - c) * pw * sizeof(float2) * (prev ? 3 : 1), s);
        k_tslu_panel<<<(unsigned)(nsets0 + tiles), kTsluT, 0, s>>>(A, lda, n, c, pw, nsets0, prev ? 1 : 0, upd_n, tw,
                                                                  fin);
        check("k_tslu_panel");
    }
    ProfScope prof(ctx, "k_tslu_apply", (double)(n - c) * pw * sizeof(float2) * 2, s);
    const int w = (c - k) + (k + kb - (c + pw));
    const int nsw = (w + 15) / 16;
    const int grid = nsw + (n - c + 255) / 256;
    k_tslu_apply<<<grid, 256, 0, s>>>(A, lda, n, c, pw, W.top.p, W.oldtop.p, W.tplan.p, k, c, c + pw, k + kb, nsw);
    check("k_tslu_apply");
}

// Trailing update C -= A B through split-bf16 planes (FP32 path only).
// With reuse_a, PA already holds the planes of the same A (same ar, ac, M, K).
void gemm_split(Ctx &ctx, hipStream_t s, float2 *A, long long lda, int ar, int ac, int br, int bc, int cr, int cc, int M,
                int N, int K, DevBuf<__bf16> &PA, DevBuf<__bf16> &PB, bool reuse_a = false)
{
    if (M <= 0 || N <= 0 || K <= 0) return;
    BH_REQUIRE(K % kSplitBK == 0, "split gemm K=%d not a multiple of %d", K, kSplitBK);
    BH_REQUIRE(!reuse_a || PA.n >= (size_t)4 * M * K, "split gemm: no planes to reuse");
    PA.ensure((size_t)4 * M * K);
    PB.ensure((size_t)4 * N * K);
    {
        ProfScope prof(ctx, "k_split", 8.0 * ((reuse_a ? 0.0 : (double)M) + N) * K * 2.0, s);
        const long long na = (long long)M * K;
        if (!reuse_a) {
            k_split_rows<<<(unsigned)((na + 255) / 256), 256, 0, s>>>(A + (long long)ar * lda + ac, lda, M, K, PA.p);
            check("k_split_rows");
        }
        dim3 g2((unsigned)((N + 31) / 32), (unsigned)((K + 31) / 32));
        k_split_cols<<<g2, 256, 0, s>>>(A + (long long)br * lda + bc, lda, K, N, PB.p);
        check("k_split_cols");
    }
    // work: bf16 MFMA flops executed (3 products per real product, 4 real per complex)
    ProfScope prof(ctx, "k_cgemm_bf16x3", 24.0 * M * N * K, s);
    // 128 x 128 complex tiles, 8 waves (4 x 2) of 32 x 64
    constexpr int TM = 2, TN = 4, WM = 4, WN = 2, BM = 16 * TM * WM, BN = 16 * TN * WN;
    const long long tiles = (long long)((N + BN - 1) / BN) * ((M + BM - 1) / BM);
    BH_REQUIRE(tiles < (1LL << 31), "gemm grid too large");
    k_cgemm_bf16x3<TM, TN, WM, WN><<<dim3((unsigned)tiles), 64 * WM * WN, 0, s>>>(
        PA.p, PB.p, A + (long long)cr * lda + cc, lda, M, N, K);
    check("k_cgemm_bf16x3");
}

// Block row of rows [k, k + kb): A[k:k+kb, c0:c1) := L11^-1 A[k:k+kb, c0:c1) (with PB,
// FP32 blocks deeper than kNBInner solve their halves and join them with one
// split-bf16 product through PB's planes).
template <typename T>
void block_row_u12(Ctx &ctx, hipStream_t s, LUWork<T> &W, typename CT<T>::t *A, long long lda, int n, int k, int kb,
                   int c0, int c1, DevBuf<__bf16> *PB = nullptr, bool one_launch = true);

// Block column [k, k + kb): its 32-column panels and the updates inside the block.
template <typename T>
void factor_block(Ctx &ctx, hipStream_t s, LUWork<T> &W, typename CT<T>::t *A, long long lda, int n, int k, int kb,
                  long long *ipiv)
{
    if constexpr (sizeof(T) == 4) {
        if (kb > kNBInner) {
            // left half; its interchanges, block row and update reach the right half; right
            // half; its interchanges reach the left half
            const int h = kNBInner;
            factor_block<T>(ctx, s, W, A, lda, n, k, h, ipiv);
            for (int q = k; q < k + h; q += kMaxSwaps)
                laswp<T>(ctx, s, A, lda, q, std::min(q + kMaxSwaps, k + h), ipiv, k + h, k + kb, 0, 0, W.plan_side.p);
            block_row_u12<T>(ctx, s, W, A, lda, n, k, h, k + h, k + kb);
            gemm_split(ctx, s, A, lda, k + h, k, k, k + h, k + h, k + h, n - (k + h), kb - h, h, W.pa_side, W.pb_side);
            factor_block<T>(ctx, s, W, A, lda, n, k + h, kb - h, ipiv);
            for (int q = k + h; q < k + kb; q += kMaxSwaps)
                laswp<T>(ctx, s, A, lda, q, std::min(q + kMaxSwaps, k + kb), ipiv, k, k + h, 0, 0, W.plan_side.p);
            return;
        }
    }
    for (int sp = 0; sp < kb; sp += kPW) {
        const int c = k + sp, pw = std::min(kPW, kb - sp);
        if constexpr (sizeof(T) == 4) {
            // the tournament launch carries the previous panel's update, the apply
            // launch this panel's block row
            tslu_panel(ctx, s, W, A, lda, n, k, kb, c, pw, ipiv, sp > 0);
        } else {
            panel<T>(ctx, W, A, lda, n, c, pw, ipiv);
            laswp<T>(ctx, s, A, lda, c, c + pw, ipiv, k, c, c + pw, k + kb, W.plan.p);
            if (c + pw < k + kb) {
                trsm_unit<T>(s, A, lda, c, pw, c, c + pw, k + kb - (c + pw));
                gemm<T>(ctx, s, A, lda, c + pw, c, c, c + pw, c + pw, c + pw, n - (c + pw), k + kb - (c + pw), pw);
            }
        }
    }
}

// E = I - S for kb/kIB blocks S = inv(L_bb) as k_trinv<true> leaves them (unit
// diagonal, zero above it): zero diagonal, negated strictly lower part.
__global__ void k_inv_to_e(float2 *E, int nblk)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nblk * kIB * kIB) return;
    const int i = (int)(idx / kIB % kIB), j = (int)(idx % kIB);
    const float2 v = E[idx];
    E[idx] = i == j ? float2{0, 0} : float2{-v.x, -v.y};
}

// Rows [r, r + kIB) of a block row, columns [c0, c1): C <- inv(L_bb) C, formed as
// C - E C with E = I - inv(L_bb).  The product has one tile row (M = kIB = BM), so
// each workgroup reads, as B, exactly the columns of C it writes, and reads all of
// them (main loop, then the epilogue's reads of C) before its first write.  128 x 32
// tiles, as for the other one-row products.
void blockrow_apply_inverse(Ctx &ctx, hipStream_t s, const float2 *E, float2 *A, long long lda, int r, int c0, int c1)
{
    const int N = c1 - c0;
    static_assert(32 * 4 == kIB, "one tile row of the FP32 product must cover the block");
    ProfScope prof(ctx, "k_cgemm", 8.0 * kIB * (double)N * kIB, s);
    float2 *C = A + (long long)r * lda + c0;
    k_cgemm<float, 4, 1, true, 3><<<(unsigned)((N + 31) / 32), 256, 0, s>>>(E, kIB, C, lda, C, lda, kIB, N, kIB);
    check("k_cgemm");
}

// U12 = L11^-1 A12 for block row [k, k + kb) over the columns [c0, n).
template <typename T>
void block_row_u12(Ctx &ctx, hipStream_t s, LUWork<T> &W, typename CT<T>::t *A, long long lda, int n, int k, int kb,
                   int c0, int c1, DevBuf<__bf16> *PB, bool one_launch)
{
    if (c1 <= c0) return;
    if constexpr (sizeof(T) == 4) {
        if (kb > kNBInner && PB) {
            // top half, its update of the bottom half (split-bf16 products), bottom half
            const int h = kNBInner;
            block_row_u12<T>(ctx, s, W, A, lda, n, k, h, c0, c1, nullptr, one_launch);
            gemm_split(ctx, s, A, lda, k + h, k, k, c0, k + h, c0, kb - h, c1 - c0, h, W.pa_u12, *PB);
            block_row_u12<T>(ctx, s, W, A, lda, n, k + h, kb - h, c0, c1, nullptr, one_launch);
            return;
        }
    }
    if constexpr (sizeof(T) == 4) {
        // FP32 block rows by 128-row blocks: the inverses of L11's diagonal blocks, then
        // per block one product with I - inv(L_bb) and one K = 128 update below it
        // (measured faster than the one-launch VALU solve at every width)
        if (kb % kIB == 0 && kb <= kNBInnerMax) {
            const int nb = kb / kIB;
            const bool side = s != ctx.stream;
            float2 *E = (side ? W.eblk_side : W.eblk).p;
            int &of = side ? W.eblk_side_of : W.eblk_of;
            if (of != k) {   // a block row taken in two column ranges reuses its inverses
                ProfScope prof(ctx, "k_trinv", 4.0 * kIB * (double)kIB * kIB * nb / 3.0, s);   // kIB^3/6 complex MACs per block
                k_trinv<true><<<nb, kIB * kTrinvParts, 0, s>>>(A + (long long)k * lda + k, lda, kb, E);
                check("k_trinv");
                k_inv_to_e<<<(unsigned)(((long long)nb * kIB * kIB + 255) / 256), 256, 0, s>>>(E, nb);
                check("k_inv_to_e");
                of = k;
            }
            for (int b = 0; b < nb; ++b) {
                const int r = k + b * kIB;
                blockrow_apply_inverse(ctx, s, E + (size_t)b * kIB * kIB, A, lda, r, c0, c1);
                if (b + 1 < nb) gemm<T>(ctx, s, A, lda, r + kIB, r, r, c0, r + kIB, c0, k + kb - (r + kIB), c1 - c0, kIB);
            }
            return;
        }
        // other depths: narrow block rows in one launch (latency-bound as separate
        // 32-row steps); wide ones, and block rows that run beside the block-column
        // chain (the one-launch kernel would take the whole device from it), keep the
        // MFMA updates below
        const int ncols = c1 - c0;
        if (one_launch && kb <= kNBInnerMax && ncols <= kBlockRowOneLaunch) {
            ProfScope prof(ctx, "k_blockrow_trsm", 4.0 * kb * (double)kb * ncols, s);   // kb^2 / 2 complex MACs per column
            k_blockrow_trsm<<<(ncols + kStripW - 1) / kStripW, 256, 0, s>>>(A, lda, k, kb, c0, ncols);
            check("k_blockrow_trsm");
            return;
        }
    }
    for (int sp = 0; sp < kb; sp += kPW) {
        const int pw = std::min(kPW, kb - sp);
        trsm_unit<T>(s, A, lda, k + sp, pw, k + sp, c0, c1 - c0);
        if (sp + pw < kb)
            gemm<T>(ctx, s, A, lda, k + sp + pw, k + sp, k + sp, c0, k + sp + pw, c0, kb - (sp + pw), c1 - c0, pw);
    }
}

// Right-looking blocked LU (block columns of NB = 256, panels of 32).  FP64
// (LAPACK's partial pivoting) runs on the context stream.  FP32 (tournament
// pivoting) looks one block ahead: after block k's U12 is formed, the next
// block column is updated first and factored on a second stream while the rest
// of the trailing matrix is updated on the context stream; the two touch
// disjoint columns.
void invert_diagonal_blocks(hipStream_t s, const float2 *LU, long long lda, int n, float2 *Dinv);

template <typename T>
struct LUBuffers {
    LUWork<T> W;
    DevBuf<__bf16> PA, PB;   // split planes of the trailing updates
};

// The FP32 factorisation's workspace and block-column stream, kept in the context.
struct LUCache {
    std::mutex mu;
    LUBuffers<float> b;
    hipStream_t s2 = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    ~LUCache() {
        if (s2) { (void)hipStreamSynchronize(s2); (void)hipStreamDestroy(s2); }
        if (ready) (void)hipEventDestroy(ready);
        if (done) (void)hipEventDestroy(done);
    }
};

LUCache &lu_cache(Ctx &ctx)
{
    static std::mutex create;
    std::lock_guard<std::mutex> g(create);
    if (!ctx.lu_cache) ctx.lu_cache = std::make_shared<LUCache>();
    return *static_cast<LUCache *>(ctx.lu_cache.get());
}

// dinv (FP32 only, may be NULL): the diagonal-block inverses and couplings of the
// factors are formed on the context stream before the workspace is released, so
// the device does not idle while the host frees it.
template <typename T>
void lu_factor(Ctx &ctx, typename CT<T>::t *A, long long lda, int n, long long *ipiv, int *info_host,
               bool split = false, float2 *dinv = nullptr)
{
    hipStream_t s = ctx.stream;
    // FP32: the context's cached workspace unless another factorisation holds it
    LUCache *cache = nullptr;
    std::unique_lock<std::mutex> held;
    std::unique_ptr<LUBuffers<T>> own;
    LUBuffers<T> *buf = nullptr;
    if constexpr (sizeof(T) == 4) {
        cache = &lu_cache(ctx);
        held = std::unique_lock<std::mutex>(cache->mu, std::try_to_lock);
        if (held.owns_lock()) {
            buf = &cache->b;
            if (cache->s2) BH_CHECK_HIP(hipStreamSynchronize(cache->s2));   // after a failed call
        } else {
            cache = nullptr;
        }
    }
    if (!buf) {
        own = std::make_unique<LUBuffers<T>>();
        buf = own.get();
    }
    LUWork<T> &W = buf->W;
    DevBuf<__bf16> &PA = buf->PA, &PB = buf->PB;
    W.eblk_of = W.eblk_side_of = -1;
    // rows per workgroup kept in LDS: 64 KiB of panel slice
    W.rpb_max = (int)(65536 / (kPW * sizeof(typename CT<T>::t)));
    int per_cu = 0;
    BH_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_panel<T>, 256,
                                                              (size_t)W.rpb_max * kPW * sizeof(typename CT<T>::t)));
    BH_REQUIRE(per_cu >= 1, "panel kernel cannot be resident");
    W.maxG = per_cu * ctx.num_cus;
    W.slot_val.ensure(2 * (size_t)W.maxG);
    W.slot_row.ensure(2 * (size_t)W.maxG);
    W.cand.ensure(2 * (size_t)W.maxG * kPW);
    W.diag.ensure(2 * kPW);
    W.bar.ensure(kBarWords);
    W.info.ensure(1);
    W.plan.ensure(4 * kMaxSwaps);
    W.plan_side.ensure(4 * kMaxSwaps);
    BH_CHECK_HIP(hipMemsetAsync(W.bar.p, 0, kBarWords * sizeof(unsigned), s));
    BH_CHECK_HIP(hipMemsetAsync(W.info.p, 0, sizeof(int), s));
    // block column of the right-looking LU (trailing-update depth); the FP32
    // path's deeper blocks halve the C traffic per flop of the trailing update
    const int NB = sizeof(T) == 4 ? kNBFP32 : 256;   // deepest block (workspace sizes)
    int planes_of = -1;   // block whose L21 planes PA holds
    auto trailing = [&](hipStream_t st, int k, int kb, int c0, int c1) {
        // A[k+kb:, c0:c1) -= L21 U12[:, c0:c1)
        if (c1 <= c0 || k + kb >= n) return;
        if constexpr (sizeof(T) == 4) {
            if (split) {
                gemm_split(ctx, st, A, lda, k + kb, k, k, c0, k + kb, c0, n - (k + kb), c1 - c0, kb, PA, PB,
                           planes_of == k);
                planes_of = k;
                return;
            }
        }
        gemm<T>(ctx, st, A, lda, k + kb, k, k, c0, k + kb, c0, n - (k + kb), c1 - c0, kb);
    };
    if constexpr (sizeof(T) == 8) {
        for (int k = 0; k < n; k += NB) {
            const int kb = std::min(NB, n - k);
            factor_block<T>(ctx, s, W, A, lda, n, k, kb, ipiv);
            laswp<T>(ctx, s, A, lda, k, k + kb, ipiv, 0, k, k + kb, n, W.plan.p);
            if (k + kb < n) {
                block_row_u12<T>(ctx, s, W, A, lda, n, k, kb, k + kb, n);
                trailing(s, k, kb, k + kb, n);
            }
        }
    } else {
        // FP32 factors only precondition GMRES-IR: their panels use tournament pivoting
        // candidate sets of all levels: n_0 + n_1 + ... < 2 n_0 + levels
        const size_t sets = 2 * ((size_t)(n + kTsluT - 1) / kTsluT) + 8;
        W.tcnt.ensure(sets);
        W.tidx.ensure(sets * kPW);
        W.tval.ensure(sets * kPW * kPW);
        BH_CHECK_HIP(hipMemsetAsync(W.tcnt.p, 0, sets * sizeof(int), s));
        W.top.ensure((size_t)kPW * kPW);
        W.oldtop.ensure((size_t)kPW * kPW);
        W.tplan.ensure(1 + 2 * kPlanMax);
        W.eblk.ensure((size_t)(kNBInnerMax / kIB) * kIB * kIB);
        W.eblk_side.ensure((size_t)(kNBInnerMax / kIB) * kIB * kIB);
        hipStream_t s2 = nullptr;
        hipEvent_t ready = nullptr, done = nullptr;
        struct Cleanup {   // a stream and events of this call only (no cache)
            hipStream_t &s2;
            hipEvent_t &a, &b;
            bool own;
            ~Cleanup() {
                if (!own) return;
                if (s2) { (void)hipStreamSynchronize(s2); (void)hipStreamDestroy(s2); }
                if (a) (void)hipEventDestroy(a);
                if (b) (void)hipEventDestroy(b);
            }
        } cleanup{s2, ready, done, cache == nullptr};
        if (cache) {
            if (!cache->s2) BH_CHECK_HIP(hipStreamCreateWithFlags(&cache->s2, hipStreamNonBlocking));
            if (!cache->ready) BH_CHECK_HIP(hipEventCreateWithFlags(&cache->ready, hipEventDisableTiming));
            if (!cache->done) BH_CHECK_HIP(hipEventCreateWithFlags(&cache->done, hipEventDisableTiming));
            s2 = cache->s2;
            ready = cache->ready;
            done = cache->done;
        } else {
            BH_CHECK_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
            BH_CHECK_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
            BH_CHECK_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        }
        factor_block<T>(ctx, s, W, A, lda, n, 0, std::min(fp32_block_depth(n, 0, split), n), ipiv);
        // block k is factored; its interchanges reach the other columns (in pieces
        // of at most kMaxSwaps interchanges, applied in order) over [c0, c1) and [c2, c3)
        auto swaps = [&](int k, int kb, int c0, int c1, int c2, int c3) {
            for (int q = k; q < k + kb; q += kMaxSwaps)
                laswp<T>(ctx, s, A, lda, q, std::min(q + kMaxSwaps, k + kb), ipiv, c0, c1, c2, c3, W.plan.p);
        };
        for (int k = 0, kb = 0; k < n; k += kb) {
            kb = std::min(fp32_block_depth(n, k, split), n - k);
            if (k + kb >= n) {
                swaps(k, kb, 0, k, 0, 0);
                break;
            }
            const int nk = k + kb, nkb = std::min(fp32_block_depth(n, nk, split), n - nk);
            // near the end the block-column chain is the critical path: there the
            // next block column goes first on its own (interchanges, block row,
            // update) and the rest is brought up to date while it is factored;
            // earlier, one pass over all columns costs the context stream less
            const bool ahead_first = n - nk <= kAheadFirstBelow;
            if (ahead_first) {
                swaps(k, kb, nk, nk + nkb, 0, 0);
                block_row_u12<T>(ctx, s, W, A, lda, n, k, kb, nk, nk + nkb, split ? &PB : nullptr);
            } else {
                swaps(k, kb, 0, k, nk, n);
                block_row_u12<T>(ctx, s, W, A, lda, n, k, kb, nk, n, split ? &PB : nullptr);
            }
            trailing(s, k, kb, nk, nk + nkb);
            BH_CHECK_HIP(hipEventRecord(ready, s));
            if (!ahead_first) {
                // the rest of the update first in host order: the next block column's
                // many short launches are then queued while the long product runs
                trailing(s, k, kb, nk + nkb, n);
            }
            BH_CHECK_HIP(hipStreamWaitEvent(s2, ready, 0));
            factor_block<T>(ctx, s2, W, A, lda, n, nk, nkb, ipiv);
            BH_CHECK_HIP(hipEventRecord(done, s2));
            // the rest, concurrently
            if (ahead_first) {
                swaps(k, kb, 0, k, nk + nkb, n);
                block_row_u12<T>(ctx, s, W, A, lda, n, k, kb, nk + nkb, n, split ? &PB : nullptr, false);
                trailing(s, k, kb, nk + nkb, n);
            }
            BH_CHECK_HIP(hipStreamWaitEvent(s, done, 0));
        }
    }
    if constexpr (sizeof(T) == 4) {
        if (dinv) invert_diagonal_blocks(s, (const float2 *)A, lda, n, dinv);
    }
    BH_CHECK_HIP(hipMemcpyAsync(info_host, W.info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    BH_CHECK_HIP(hipStreamSynchronize(s));
}

// forward (unit lower) then backward (upper) substitution on X (n x nrhs, work precision)
template <typename T>
void lu_solve_inplace(hipStream_t s, const typename CT<T>::t *A, long long lda, int n, typename CT<T>::t *X, int nrhs)
{
    const int nb = (n + kTB - 1) / kTB;
    k_trsv_first<T, true><<<1, 256, 0, s>>>(A, lda, n, X, nrhs, 0);
    check("k_trsv_first");
    for (int b = 0; b + 1 < nb; ++b) {
        k_trsv_step<T, true><<<nb - 1 - b, 256, 0, s>>>(A, lda, n, X, nrhs, b, nb);
        check("k_trsv_step");
    }
    k_trsv_first<T, false><<<1, 256, 0, s>>>(A, lda, n, X, nrhs, nb - 1);
    check("k_trsv_first");
    for (int b = nb - 1; b > 0; --b) {
        k_trsv_step<T, false><<<b, 256, 0, s>>>(A, lda, n, X, nrhs, b, nb);
        check("k_trsv_step");
    }
}

// Inverses of the diagonal blocks of both FP32 factors: Dinv[0 .. nb) lower, [nb .. 2nb) upper.
void invert_diagonal_blocks(hipStream_t s, const float2 *LU, long long lda, int n, float2 *Dinv)
{
    const int nb = (n + kIB - 1) / kIB;
    const size_t bsz = (size_t)nb * kIB * kIB;
    k_trinv<true><<<nb, kIB * kTrinvParts, 0, s>>>(LU, lda, n, Dinv);
    check("k_trinv");
    k_trinv<false><<<nb, kIB * kTrinvParts, 0, s>>>(LU, lda, n, Dinv + bsz);
    check("k_trinv");
    k_blk_coupling<true><<<dim3(nb, kIB / 16), 256, 0, s>>>(LU, lda, n, Dinv, Dinv + 2 * bsz);
    check("k_blk_coupling");
    k_blk_coupling<false><<<dim3(nb, kIB / 16), 256, 0, s>>>(LU, lda, n, Dinv + bsz, Dinv + 3 * bsz);
    check("k_blk_coupling");
}

// Synchronisation words of the one-launch triangular solves of one matrix.
// words[0 .. cnt) = 0 (a kernel rather than a memset node, for callers that capture graphs)
__global__ void k_fill_int(int *words, long long cnt)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) words[i] = 0;
}

struct TrsvSync {
    int *words;           // [0] ticket counter, [1] error flag, [2 ..] ready flags
    long long &tickets;   // tickets handed out so far
    int &epoch;           // ready-flag value of the last triangle
    int cus;              // workgroups of a launch at most (one per CU)
};

template <int NRHS>
void lu_solve_inv_n(hipStream_t s, const float2 *A, long long lda, int n, const float2 *Dinv, float2 *X, TrsvSync &ts)
{
    const int nb = (n + kIB - 1) / kIB;
    const size_t bsz = (size_t)nb * kIB * kIB;
    const float2 *DL = Dinv, *DU = Dinv + bsz, *GL = Dinv + 2 * bsz, *GU = Dinv + 3 * bsz;
    for (int tri = 0; tri < 2; ++tri) {
        if (ts.tickets + 2LL * nb >= (1LL << 30) || ts.epoch >= (1 << 30)) {   // restart the counters
            k_fill_int<<<(unsigned)((nb + 2 + 255) / 256), 256, 0, s>>>(ts.words, nb + 2);
            check("k_fill_int");
            ts.tickets = 0;
            ts.epoch = 0;
        }
        const int grid = std::max(1, std::min(nb, ts.cus));
        const int ticket0 = (int)ts.tickets, epoch = ++ts.epoch;
        ts.tickets += nb + grid;
        if (tri == 0) k_trsv_flag<true, NRHS><<<grid, kTrsvT, 0, s>>>(A, lda, n, DL, GL, X, ts.words, ticket0, epoch);
        else k_trsv_flag<false, NRHS><<<grid, kTrsvT, 0, s>>>(A, lda, n, DU, GU, X, ts.words, ticket0, epoch);
        check("k_trsv_flag");
    }
}

void lu_solve_inv(hipStream_t s, const float2 *A, long long lda, int n, const float2 *Dinv, float2 *X, int nrhs,
                  TrsvSync &ts)
{
    switch (nrhs) {
    case 1: lu_solve_inv_n<1>(s, A, lda, n, Dinv, X, ts); break;
    case 2: lu_solve_inv_n<2>(s, A, lda, n, Dinv, X, ts); break;
    case 3: lu_solve_inv_n<3>(s, A, lda, n, Dinv, X, ts); break;
    case 4: lu_solve_inv_n<4>(s, A, lda, n, Dinv, X, ts); break;
    case 5: lu_solve_inv_n<5>(s, A, lda, n, Dinv, X, ts); break;
    case 6: lu_solve_inv_n<6>(s, A, lda, n, Dinv, X, ts); break;
    case 7: lu_solve_inv_n<7>(s, A, lda, n, Dinv, X, ts); break;
    case 8: lu_solve_inv_n<8>(s, A, lda, n, Dinv, X, ts); break;
    default: BH_FAIL(BH_ERR_INVALID, "nrhs %d out of range", nrhs);
    }
}

long long padded_ld(int64_t n) { return (n + 63) / 64 * 64; }

// ---------------------------------------------------------------- batched FP32 factorisation
// Several matrices factored in shared launches: every launch of the single-matrix
// schedule (512-deep block columns: tournament panels, their apply, interchanges, block
// rows, split-bf16 trailing updates; then the diagonal-block inverses of the solves) is
// issued once for all matrices that have it at that step.  A workgroup finds its matrix in
// the table of each matrix's first workgroup and runs the single-matrix kernel's body on
// that matrix's arguments, so each matrix gets the factors (and pivots) it gets alone.
struct PanelBA {
    float2 *A;
    long long lda;
    int n, c, pw, nsets0, prev, upd_n;
    TsluWork tw;
    TsluFinal fin;
};
struct ApplyBA {
    float2 *A;
    long long lda;
    int n, c, pw;
    const float2 *top, *oldtop;
    const int *plan;
    int a0, a1, b0, b1, nsw;
};
struct SwapBA {
    const long long *ipiv;
    int k1, k2;
    int *plan;
};
struct LaswpBA {
    float2 *A;
    long long lda;
    const int *plan;
    int len, c0, c1, c2, c3;
};
struct TrinvBA {
    const float2 *A;
    long long lda;
    int n;
    float2 *Dinv;
};
struct ToEBA {
    float2 *E;
    int nblk;
};
struct CgemmBA {
    const float2 *A;
    long long lda;
    const float2 *B;
    long long ldb;
    float2 *C;
    long long ldc;
    int M, N, K;
};
struct SplitRowsBA {
    const float2 *A;
    long long lda;
    int rows, K;
    __bf16 *P;
};
struct SplitColsBA {
    const float2 *B;
    long long ldb;
    int K, cols, gx;
    __bf16 *P;
};
struct Bf16BA {
    const __bf16 *PA, *PB;
    float2 *C;
    long long ldc;
    int M, N, K;
};
struct CouplingBA {
    const float2 *A;
    long long lda;
    int n, gx;
    const float2 *Dinv;
    float2 *G;
};

#define BH_BATCH_ITEM(ARGS)                              \
    const int r_ = batch_find(first, count, blockIdx.x); \
    const auto &a = ARGS[r_];                            \
    const int bid = (int)blockIdx.x - first[r_];

__global__ __launch_bounds__(kTsluT, 2) void k_tslu_panel_batch(const PanelBA *args, const int *first, int count)
{
    BH_BATCH_ITEM(args)
    tslu_panel_body(a.A, a.lda, a.n, a.c, a.pw, a.nsets0, a.prev, a.upd_n, a.tw, a.fin, bid);
}

__global__ __launch_bounds__(256) void k_tslu_apply_batch(const ApplyBA *args, const int *first, int count)
{
    BH_BATCH_ITEM(args)
    tslu_apply_body(a.A, a.lda, a.n, a.c, a.pw, a.top, a.oldtop, a.plan, a.a0, a.a1, a.b0, a.b1, a.nsw, bid);
}

__global__ __launch_bounds__(256) void k_swap_plan_batch(const SwapBA *args, const int *first, int count)
{
    BH_BATCH_ITEM(args)
    (void)bid;
    swap_plan_body(a.ipiv, a.k1, a.k2, a.plan);
}

__global__ __launch_bounds__(256) void k_laswp_apply_batch(const LaswpBA *args, const int *first, int count)
{
    BH_BATCH_ITEM(args)
    laswp_apply_body<float>(a.A, a.lda, a.plan, a.len, a.c0, a.c1, a.c2, a.c3, bid);
}

template <bool LOWER>
__global__ __launch_bounds__(kIB * kTrinvParts) void k_trinv_batch(const TrinvBA *args, const int *first, int count)
{
    BH_BATCH_ITEM(args)
    trinv_body<LOWER>(a.A, a.lda, a.n, a.Dinv, bid);
}

__global__ __launch_bounds__(256) void k_inv_to_e_batch(const ToEBA *args, const int *first, int count)
{
    BH_BATCH_ITEM(args)
    const long long idx = (long long)bid * 256 + threadIdx.x;
    if (idx >= (long long)a.nblk * kIB * kIB) return;
    const int i = (int)(idx / kIB % kIB), j = (int)(idx % kIB);
    const float2 v = a.E[idx];
    a.E[idx] = i == j ? float2{0, 0} : float2{-v.x, -v.y};
}

__global__ __launch_bounds__(256, 2) void k_cgemm_batch(const CgemmBA *args, const int *first, int count)
{
    BH_BATCH_ITEM(args)
    cgemm_tile<float, 4, 2, true>(bid, a.A, a.lda, a.B, a.ldb, a.C, a.ldc, a.M, a.N, a.K);
}

__global__ __launch_bounds__(256) void k_split_rows_batch(const SplitRowsBA *args, const int *first, int count)
{
    BH_BATCH_ITEM(args)
    split_rows_body(a.A, a.lda, a.rows, a.K, a.P, (long long)bid * 256 + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_split_cols_batch(const SplitColsBA *args, const int *first, int count)
{
    BH_BATCH_ITEM(args)
    split_cols_body(a.B, a.ldb, a.K, a.cols, a.P, bid % a.gx, bid / a.gx);
}

__global__ __launch_bounds__(512, 2) void k_cgemm_bf16x3_batch(const Bf16BA *args, const int *first, int count)
{
    BH_BATCH_ITEM(args)
    cgemm_bf16x3_body<2, 4, 4, 2>(a.PA, a.PB, a.C, a.ldc, a.M, a.N, a.K, bid);
}

template <bool LOWER>
__global__ __launch_bounds__(256) void k_blk_coupling_batch(const CouplingBA *args, const int *first, int count)
{
    BH_BATCH_ITEM(args)
    blk_coupling_body<LOWER>(a.A, a.lda, a.n, a.Dinv, a.G, bid % a.gx, bid / a.gx);
}

#undef BH_BATCH_ITEM

// Largest matrix the batched schedule takes: every block column 512 deep (the single
// schedule's deeper blocks start above this trailing size).
constexpr int kBatchMaxN = kNBDeepAbove;

namespace {

enum BatchKind : int {
    kBkPanel, kBkApply, kBkSwap, kBkLaswp, kBkTrinvL, kBkTrinvU, kBkToE, kBkCgemm, kBkSplitRows, kBkSplitCols,
    kBkBf16, kBkCoupleL, kBkCoupleU
};

struct BatchLaunch {
    int kind;
    int count;            // matrices in this launch
    int blocks;           // total workgroups
    size_t args, first;   // offsets into the uploaded table
    size_t shm;
    double work;          // for the profiler: the single-matrix kernels' work units, summed
};

// Host-side schedule: launches in order and one table of their arguments.
struct BatchPlan {
    std::vector<char> blob;
    std::vector<BatchLaunch> launches;
    template <typename S>
    size_t put(const std::vector<S> &v) {
        size_t off = (blob.size() + 15) & ~(size_t)15;
        blob.resize(off + v.size() * sizeof(S));
        std::memcpy(blob.data() + off, v.data(), v.size() * sizeof(S));
        return off;
    }
    // items: (args, blocks) per matrix; matrices with no blocks are left out
    template <typename S>
    void add(int kind, const std::vector<std::pair<S, long long>> &items, size_t shm = 0, double work = 0) {
        std::vector<S> a;
        std::vector<int> first{0};
        for (const auto &it : items) {
            if (it.second <= 0) continue;
            BH_REQUIRE((long long)first.back() + it.second < (1LL << 31), "batched launch too large");
            a.push_back(it.first);
            first.push_back(first.back() + (int)it.second);
        }
        if (a.empty()) return;
        BatchLaunch L{kind, (int)a.size(), first.back(), put(a), 0, shm, work};
        L.first = put(first);
        launches.push_back(L);
    }
};

struct BatchMat {
    bh_matrix *M;
    int n;
    long long lda;
    float2 *A;
    long long *ipiv;
    // workspace (device)
    int *tcnt, *tidx, *tplan, *plan, *info;
    float2 *tval, *top, *oldtop, *eblk, *dinv;
    __bf16 *pa, *pb;
};

void plan_factor(BatchPlan &P, std::vector<BatchMat> &mats)
{
    int nmax = 0;
    for (auto &m : mats) nmax = std::max(nmax, m.n);
    for (int k = 0; k < nmax; k += kNBInner) {
        // the block column [k, k + kb) of every matrix that has it: its panels
        for (int sp = 0; sp < kNBInner; sp += kPW) {
            std::vector<std::pair<PanelBA, long long>> pan;
            std::vector<std::pair<ApplyBA, long long>> app;
            double wp = 0, wa = 0;
            for (auto &m : mats) {
                if (m.n <= k) continue;
                const int kb = std::min(kNBInner, m.n - k);
                if (sp >= kb) continue;
                const int c = k + sp, pw = std::min(kPW, kb - sp), prev = sp > 0;
                const int nsets0 = (m.n - c + kTsluT - 1) / kTsluT;
                const int upd_n = prev ? std::max(0, k + kb - (c + kPW)) : 0;
                const long long tiles = upd_n > 0 ? (long long)((m.n - c + 127) / 128) * ((upd_n + 63) / 64) : 0;
                PanelBA a{m.A, m.lda, m.n, c, pw, nsets0, prev, upd_n, TsluWork{m.tcnt, m.tidx, m.tval},
                          TsluFinal{m.ipiv, m.top, m.oldtop, m.tplan, m.info}};
                pan.push_back({a, nsets0 + tiles});
                wp += (double)(m.n - c) * pw * sizeof(float2) * (prev ? 3 : 1);
                const int w = (c - k) + (k + kb - (c + pw));
                const int nsw = (w + 15) / 16;
                ApplyBA b{m.A, m.lda, m.n, c, pw, m.top, m.oldtop, m.tplan, k, c, c + pw, k + kb, nsw};
                app.push_back({b, nsw + (m.n - c + 255) / 256});
                wa += (double)(m.n - c) * pw * sizeof(float2) * 2;
            }
            if (pan.empty()) break;
            P.add(kBkPanel, pan, 0, wp);
            P.add(kBkApply, app, 0, wa);
        }
        // interchanges of the block reach the columns left of it, and right of it when
        // it is not the last block (pieces of kMaxSwaps, in order)
        for (int q0 = 0; q0 < kNBInner; q0 += kMaxSwaps) {
            std::vector<std::pair<SwapBA, long long>> sw;
            std::vector<std::pair<LaswpBA, long long>> la;
            double wl = 0;
            for (auto &m : mats) {
                if (m.n <= k) continue;
                const int kb = std::min(kNBInner, m.n - k), nk = k + kb;
                const int q = k + q0;
                if (q >= nk) continue;
                const int len = std::min(kMaxSwaps, nk - q);
                const int c2 = nk < m.n ? nk : 0, c3 = nk < m.n ? m.n : 0;
                const int w = k + (c3 - c2);
                if (w <= 0) continue;
                sw.push_back({SwapBA{m.ipiv, q, q + len, m.plan}, 1});
                constexpr int cw = laswp_cols<float>();
                la.push_back({LaswpBA{m.A, m.lda, m.plan, len, 0, k, c2, c3}, (w + cw - 1) / cw});
                wl += 2.0 * 2 * len * w * sizeof(float2);
            }
            P.add(kBkSwap, sw);
            P.add(kBkLaswp, la, 0, wl);
        }
        // block row U12 = L11^-1 A12 and the trailing update, for matrices with columns right
        // of the block: inverses of L11's 128-row diagonal blocks, then per 128-row block
        // one product with I - inv(L_bb) and one K = 128 update of the rows below
        std::vector<BatchMat *> rest;
        for (auto &m : mats)
            if (m.n > k + kNBInner) rest.push_back(&m);
        if (rest.empty()) continue;
        const int nk = k + kNBInner, nbb = kNBInner / kIB;
        {
            std::vector<std::pair<TrinvBA, long long>> ti;
            std::vector<std::pair<ToEBA, long long>> te;
            for (auto *m : rest) {
                ti.push_back({TrinvBA{m->A + (long long)k * m->lda + k, m->lda, kNBInner, m->eblk}, nbb});
                te.push_back({ToEBA{m->eblk, nbb}, ((long long)nbb * kIB * kIB + 255) / 256});
            }
            P.add(kBkTrinvL, ti, 0, 4.0 * kIB * (double)kIB * kIB * nbb / 3.0 * rest.size());
            P.add(kBkToE, te);
        }
        for (int b = 0; b < nbb; ++b) {
            const int r = k + b * kIB;
            std::vector<std::pair<CgemmBA, long long>> ap, up;
            double wap = 0, wup = 0;
            for (auto *m : rest) {
                const int N = m->n - nk;
                float2 *C = m->A + (long long)r * m->lda + nk;
                ap.push_back({CgemmBA{m->eblk + (size_t)b * kIB * kIB, kIB, C, m->lda, C, m->lda, kIB, N, kIB},
                              (N + 63) / 64});
                wap += 8.0 * kIB * (double)N * kIB;
                if (b + 1 < nbb) {
                    const int Mr = nk - (r + kIB);
                    up.push_back({CgemmBA{m->A + (long long)(r + kIB) * m->lda + r, m->lda, C, m->lda,
                                          m->A + (long long)(r + kIB) * m->lda + nk, m->lda, Mr, N, kIB},
                                  (long long)((N + 63) / 64) * ((Mr + 127) / 128)});
                    wup += 8.0 * Mr * (double)N * kIB;
                }
            }
            P.add(kBkCgemm, ap, 0, wap);
            P.add(kBkCgemm, up, 0, wup);
        }
        {
            std::vector<std::pair<SplitRowsBA, long long>> sr;
            std::vector<std::pair<SplitColsBA, long long>> sc;
            std::vector<std::pair<Bf16BA, long long>> gm;
            double wsplit = 0, wg = 0;
            for (auto *m : rest) {
                const int Mt = m->n - nk, Nt = m->n - nk, K = kNBInner;
                sr.push_back({SplitRowsBA{m->A + (long long)nk * m->lda + k, m->lda, Mt, K, m->pa},
                              ((long long)Mt * K + 255) / 256});
                const int gx = (Nt + 31) / 32;
                sc.push_back({SplitColsBA{m->A + (long long)k * m->lda + nk, m->lda, K, Nt, gx, m->pb},
                              (long long)gx * ((K + 31) / 32)});
                gm.push_back({Bf16BA{m->pa, m->pb, m->A + (long long)nk * m->lda + nk, m->lda, Mt, Nt, K},
                              (long long)((Nt + 127) / 128) * ((Mt + 127) / 128)});
                wsplit += 8.0 * ((double)Mt + Nt) * K * 2.0;
                wg += 24.0 * Mt * (double)Nt * K;
            }
            P.add(kBkSplitRows, sr, 0, wsplit);
            P.add(kBkSplitCols, sc);
            P.add(kBkBf16, gm, 0, wg);
        }
    }
    // the solves' diagonal-block inverses and couplings
    std::vector<std::pair<TrinvBA, long long>> tl, tu;
    std::vector<std::pair<CouplingBA, long long>> cl, cu;
    for (auto &m : mats) {
        const int nb = (m.n + kIB - 1) / kIB;
        const size_t bsz = (size_t)nb * kIB * kIB;
        tl.push_back({TrinvBA{m.A, m.lda, m.n, m.dinv}, nb});
        tu.push_back({TrinvBA{m.A, m.lda, m.n, m.dinv + bsz}, nb});
        cl.push_back({CouplingBA{m.A, m.lda, m.n, nb, m.dinv, m.dinv + 2 * bsz}, (long long)nb * (kIB / 16)});
        cu.push_back({CouplingBA{m.A, m.lda, m.n, nb, m.dinv + bsz, m.dinv + 3 * bsz}, (long long)nb * (kIB / 16)});
    }
    P.add(kBkTrinvL, tl);
    P.add(kBkTrinvU, tu);
    P.add(kBkCoupleL, cl);
    P.add(kBkCoupleU, cu);
}

const char *batch_kernel_name(int kind)
{
    switch (kind) {
    case kBkPanel: return "k_tslu_panel";
    case kBkApply: return "k_tslu_apply";
    case kBkSwap: return "k_swap_plan";
    case kBkLaswp: return "k_laswp";
    case kBkTrinvL:
    case kBkTrinvU: return "k_trinv";
    case kBkToE: return "k_inv_to_e";
    case kBkCgemm: return "k_cgemm";
    case kBkSplitRows:
    case kBkSplitCols: return "k_split";
    case kBkBf16: return "k_cgemm_bf16x3";
    default: return "k_blk_coupling";
    }
}

void run_plan(Ctx &ctx, const BatchPlan &P, const char *dtab)
{
    hipStream_t s = ctx.stream;
    for (const auto &L : P.launches) {
        const unsigned g = (unsigned)L.blocks;
        const int *first = (const int *)(dtab + L.first);
        const void *a = dtab + L.args;
        ProfScope prof(ctx, batch_kernel_name(L.kind), L.work, s);
        switch (L.kind) {
        case kBkPanel: k_tslu_panel_batch<<<g, kTsluT, 0, s>>>((const PanelBA *)a, first, L.count); break;
        case kBkApply: k_tslu_apply_batch<<<g, 256, 0, s>>>((const ApplyBA *)a, first, L.count); break;
        case kBkSwap: k_swap_plan_batch<<<g, 256, 0, s>>>((const SwapBA *)a, first, L.count); break;
        case kBkLaswp: k_laswp_apply_batch<<<g, 256, L.shm, s>>>((const LaswpBA *)a, first, L.count); break;
        case kBkTrinvL:
            k_trinv_batch<true><<<g, kIB * kTrinvParts, 0, s>>>((const TrinvBA *)a, first, L.count);
            break;
        case kBkTrinvU:
            k_trinv_batch<false><<<g, kIB * kTrinvParts, 0, s>>>((const TrinvBA *)a, first, L.count);
            break;
        case kBkToE: k_inv_to_e_batch<<<g, 256, 0, s>>>((const ToEBA *)a, first, L.count); break;
        case kBkCgemm: k_cgemm_batch<<<g, 256, 0, s>>>((const CgemmBA *)a, first, L.count); break;
        case kBkSplitRows: k_split_rows_batch<<<g, 256, 0, s>>>((const SplitRowsBA *)a, first, L.count); break;
        case kBkSplitCols: k_split_cols_batch<<<g, 256, 0, s>>>((const SplitColsBA *)a, first, L.count); break;
        case kBkBf16: k_cgemm_bf16x3_batch<<<g, 512, 0, s>>>((const Bf16BA *)a, first, L.count); break;
        case kBkCoupleL:
            k_blk_coupling_batch<true><<<g, 256, 0, s>>>((const CouplingBA *)a, first, L.count);
            break;
        case kBkCoupleU:
            k_blk_coupling_batch<false><<<g, 256, 0, s>>>((const CouplingBA *)a, first, L.count);
            break;
        default: BH_FAIL(BH_ERR_INVALID, "unknown batched launch %d", L.kind);
        }
        check(batch_kernel_name(L.kind));
    }
}

}  // namespace

}  // namespace bh

using namespace bh;

namespace {

// FP32 factorisation inputs: the matrix rounded to FP32 in lu32 (unless the assembly wrote
// it there) and the buffer of the solves' diagonal-block inverses.
void fp32_prepare(Ctx &ctx, bh_matrix *M, long long lda)
{
    const int64_t n = M->n;
    pool_release(ctx, M->lu64);
    if (M->lu32.n != (size_t)lda * n * 2) {
        pool_alloc(ctx, M->lu32, (size_t)lda * n * 2);
        M->lu32_fresh = false;
    }
    if (!M->lu32_fresh) {
        const unsigned cgrid = (unsigned)(((long long)n * n + 255) / 256);
        k_to_work<float><<<cgrid, 256, 0, ctx.stream>>>((const double2 *)M->a64.p, (int)n, (float2 *)M->lu32.p, lda);
        check("k_to_work");
    }
    M->lu32_fresh = false;   // it now holds the factors
    const size_t nb = (size_t)(n + kIB - 1) / kIB;
    pool_alloc(ctx, M->dinv32, 4 * nb * kIB * kIB * 2);   // L and U diagonal inverses, then their couplings
}

// After a factorisation: the FP32 solves' vector and block flags (so that solving
// allocates nothing), the permutation (P b)[i] = b[perm[i]] from the sequential
// interchanges, and the matrix state.
void factor_finish(Ctx &ctx, bh_matrix *M, long long lda, int precision, int info)
{
    const int64_t n = M->n;
    if (precision != BH_FP64) {
        if (M->x32.n < (size_t)n * kMaxRhs * 2) M->x32.alloc((size_t)n * kMaxRhs * 2);
        const size_t nflags = (size_t)((n + kIB - 1) / kIB) + 2;
        if (M->tsync.n < nflags) {
            M->tsync.alloc(nflags);
            k_fill_int<<<(unsigned)((nflags + 255) / 256), 256, 0, ctx.stream>>>(M->tsync.p, (long long)nflags);
            check("k_fill_int");
            M->tsync_tickets = 0;
            M->tsync_epoch = 0;
        }
    }
    std::vector<long long> ip(n);
    BH_CHECK_HIP(hipMemcpy(ip.data(), M->ipiv.p, sizeof(long long) * n, hipMemcpyDeviceToHost));
    std::vector<int> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = (int)i;
    for (int64_t i = 0; i < n; ++i) std::swap(perm[i], perm[ip[i]]);
    M->perm.upload(perm.data(), n, ctx.stream);
    BH_CHECK_HIP(hipStreamSynchronize(ctx.stream));
    M->lda = lda;
    M->factored = info ? -1 : (precision == BH_FP64 ? BH_FP64 : BH_FP32);
}

}  // namespace

extern "C" int bh_matrix_factor(bh_matrix *M, int precision)
{
    BH_API_BEGIN
    BH_REQUIRE(M != nullptr, "NULL matrix");
    BH_REQUIRE(precision == BH_FP64 || precision == BH_FP32 || precision == BH_FP32_SPLIT,
               "precision must be BH_FP64, BH_FP32 or BH_FP32_SPLIT");
    std::lock_guard<std::mutex> lock(M->mu);
    Ctx &ctx = *M->ctx;
    const int64_t n = M->n;
    BH_REQUIRE(n > 0 && n < (int64_t(1) << 31), "bad matrix size");
    const long long lda = padded_ld(n);
    M->ipiv.alloc(n);
    hipEvent_t e0, e1;
    BH_CHECK_HIP(hipEventCreate(&e0));
    BH_CHECK_HIP(hipEventCreate(&e1));
    int info = 0;
    BH_CHECK_HIP(hipEventRecord(e0, ctx.stream));
    if (precision == BH_FP64) {
        pool_release(ctx, M->lu32);
        M->lu32_fresh = false;
        pool_release(ctx, M->dinv32);
        pool_alloc(ctx, M->lu64, (size_t)lda * n * 2);
        const unsigned cgrid = (unsigned)(((long long)n * n + 255) / 256);
        k_to_work<double><<<cgrid, 256, 0, ctx.stream>>>((const double2 *)M->a64.p, (int)n, (double2 *)M->lu64.p, lda);
        check("k_to_work");
        lu_factor<double>(ctx, (double2 *)M->lu64.p, lda, (int)n, M->ipiv.p, &info);
    } else {
        fp32_prepare(ctx, M, lda);
        lu_factor<float>(ctx, (float2 *)M->lu32.p, lda, (int)n, M->ipiv.p, &info, precision == BH_FP32_SPLIT,
                         (float2 *)M->dinv32.p);
    }
    BH_CHECK_HIP(hipEventRecord(e1, ctx.stream));
    BH_CHECK_HIP(hipEventSynchronize(e1));
    float ms = 0;
    BH_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    M->seconds_factor = ms * 1e-3;
    factor_finish(ctx, M, lda, precision, info);
    if (info) BH_FAIL(BH_ERR_SINGULAR, "SingularException(%d): exactly zero pivot", info);
    BH_API_END
}

extern "C" int bh_matrix_factor_batch(bh_ctx *ctxh, int64_t count, bh_matrix *const *mats, int precision)
{
    BH_API_BEGIN
    BH_REQUIRE(ctxh != nullptr && (count == 0 || mats != nullptr), "NULL argument");
    BH_REQUIRE(count >= 0 && count <= (1 << 20), "bad batch size %lld", (long long)count);
    BH_REQUIRE(precision == BH_FP64 || precision == BH_FP32 || precision == BH_FP32_SPLIT,
               "precision must be BH_FP64, BH_FP32 or BH_FP32_SPLIT");
    Ctx &ctx = *ctxh;
    for (int64_t i = 0; i < count; ++i) {
        BH_REQUIRE(mats[i] != nullptr, "NULL matrix %lld", (long long)i);
        BH_REQUIRE(mats[i]->ctx == &ctx, "matrix %lld belongs to another context", (long long)i);
        BH_REQUIRE(mats[i]->n > 0 && mats[i]->n < (int64_t(1) << 31), "bad matrix size");
        for (int64_t j = 0; j < i; ++j) BH_REQUIRE(mats[j] != mats[i], "matrix %lld appears twice", (long long)i);
    }
    // the shared schedule takes FP32 split-bf16 factorisations of up to kBatchMaxN rows;
    // the others are factored one by one, as bh_matrix_factor does
    std::vector<bh_matrix *> shared, alone;
    for (int64_t i = 0; i < count; ++i)
        (precision == BH_FP32_SPLIT && mats[i]->n <= kBatchMaxN ? shared : alone).push_back(mats[i]);
    int first_error = BH_OK;
    std::string first_msg;
    for (bh_matrix *M : alone) {
        const int rc = bh_matrix_factor(M, precision);
        if (rc != BH_OK && first_error == BH_OK) {
            first_error = rc;
            first_msg = bh_last_error();
        }
    }
    if (!shared.empty()) {
        std::vector<std::unique_lock<std::mutex>> locks;
        for (bh_matrix *M : shared) locks.emplace_back(M->mu);
        // workspace of each matrix, one allocation: counters and info first (cleared together)
        std::vector<BatchMat> bm(shared.size());
        size_t ints = 0, rest = 0;
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        std::vector<size_t> sets(shared.size());
        for (size_t i = 0; i < shared.size(); ++i) {
            const int n = (int)shared[i]->n;
            sets[i] = 2 * ((size_t)(n + kTsluT - 1) / kTsluT) + 8;
            ints += al((sets[i] + 1) * sizeof(int));
            const size_t trail = n > kNBInner ? (size_t)(n - kNBInner) : 0;
            rest += al(sets[i] * kPW * sizeof(int)) + al((1 + 2 * kPlanMax) * sizeof(int)) +
                    al(4 * kMaxSwaps * sizeof(int)) + al(sets[i] * kPW * kPW * sizeof(float2)) +
                    2 * al(kPW * kPW * sizeof(float2)) + al((size_t)(kNBInner / kIB) * kIB * kIB * sizeof(float2)) +
                    2 * al(4 * trail * kNBInner * sizeof(__bf16));
        }
        DevBuf<char> ws;
        ws.alloc(ints + rest);
        k_fill_int<<<(unsigned)((ints / sizeof(int) + 255) / 256), 256, 0, ctx.stream>>>((int *)ws.p,
                                                                                        (long long)(ints / sizeof(int)));
        check("k_fill_int");
        size_t oi = 0, orest = ints;
        auto take = [&](size_t &off, size_t bytes) {
            char *p = ws.p + off;
            off += al(bytes);
            return p;
        };
        hipEvent_t e0, e1;
        BH_CHECK_HIP(hipEventCreate(&e0));
        BH_CHECK_HIP(hipEventCreate(&e1));
        BH_CHECK_HIP(hipEventRecord(e0, ctx.stream));
        for (size_t i = 0; i < shared.size(); ++i) {
            bh_matrix *M = shared[i];
            const int n = (int)M->n;
            const long long lda = padded_ld(n);
            M->ipiv.alloc(n);
            fp32_prepare(ctx, M, lda);
            BatchMat &b = bm[i];
            b.M = M;
            b.n = n;
            b.lda = lda;
            b.A = (float2 *)M->lu32.p;
            b.ipiv = M->ipiv.p;
            b.dinv = (float2 *)M->dinv32.p;
            b.tcnt = (int *)take(oi, (sets[i] + 1) * sizeof(int));
            b.info = b.tcnt + sets[i];
            const size_t trail = n > kNBInner ? (size_t)(n - kNBInner) : 0;
            b.tidx = (int *)take(orest, sets[i] * kPW * sizeof(int));
            b.tplan = (int *)take(orest, (1 + 2 * kPlanMax) * sizeof(int));
            b.plan = (int *)take(orest, 4 * kMaxSwaps * sizeof(int));
            b.tval = (float2 *)take(orest, sets[i] * kPW * kPW * sizeof(float2));
            b.top = (float2 *)take(orest, kPW * kPW * sizeof(float2));
            b.oldtop = (float2 *)take(orest, kPW * kPW * sizeof(float2));
            b.eblk = (float2 *)take(orest, (size_t)(kNBInner / kIB) * kIB * kIB * sizeof(float2));
            b.pa = (__bf16 *)take(orest, 4 * trail * kNBInner * sizeof(__bf16));
            b.pb = (__bf16 *)take(orest, 4 * trail * kNBInner * sizeof(__bf16));
        }
        BatchPlan plan;
        plan_factor(plan, bm);
        DevBuf<char> tab;
        tab.alloc(std::max<size_t>(plan.blob.size(), 16));
        struct SyncOnExit {   // no workspace is freed under running kernels, also on errors
            hipStream_t s;
            ~SyncOnExit() { (void)hipStreamSynchronize(s); }
        } sync_on_exit{ctx.stream};
        BH_CHECK_HIP(hipMemcpyAsync(tab.p, plan.blob.data(), plan.blob.size(), hipMemcpyHostToDevice, ctx.stream));
        run_plan(ctx, plan, tab.p);
        BH_CHECK_HIP(hipEventRecord(e1, ctx.stream));
        std::vector<int> infos(shared.size());
        for (size_t i = 0; i < shared.size(); ++i)
            BH_CHECK_HIP(hipMemcpyAsync(&infos[i], bm[i].info, sizeof(int), hipMemcpyDeviceToHost, ctx.stream));
        BH_CHECK_HIP(hipStreamSynchronize(ctx.stream));
        float ms = 0;
        BH_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        for (size_t i = 0; i < shared.size(); ++i) {
            bh_matrix *M = shared[i];
            M->seconds_factor = ms * 1e-3;   // the whole batch's factorisation time
            factor_finish(ctx, M, bm[i].lda, precision, infos[i]);
            if (infos[i] && first_error == BH_OK) {
                first_error = BH_ERR_SINGULAR;
                char buf[160];
                std::snprintf(buf, sizeof buf, "SingularException(%d): exactly zero pivot (batch matrix of n = %lld)",
                              infos[i], (long long)M->n);
                first_msg = buf;
            }
        }
    }
    if (first_error != BH_OK) BH_FAIL(first_error, "%s", first_msg.c_str());
    BH_API_END
}

namespace {

// Z x = b with the factors: X (device, n x nrhs row-major, FP64) from B (same layout)
void apply_factors(bh_matrix *M, const double2 *B, double2 *Xd, int nrhs)
{
    hipStream_t s = M->ctx->stream;
    const int n = (int)M->n;
    const long long cnt = (long long)n * nrhs;
    const unsigned g = (unsigned)((cnt + 255) / 256);
    if (M->factored == BH_FP64) {
        k_permute_in<double><<<g, 256, 0, s>>>(B, n, nrhs, M->perm.p, Xd);
        check("k_permute_in");
        lu_solve_inplace<double>(s, (const double2 *)M->lu64.p, M->lda, n, Xd, nrhs);
    } else {
        if (M->x32.n < (size_t)cnt * 2) M->x32.alloc((size_t)n * kMaxRhs * 2);
        float2 *x32 = (float2 *)M->x32.p;
        k_permute_in<float><<<g, 256, 0, s>>>(B, n, nrhs, M->perm.p, x32);
        check("k_permute_in");
        {
            const int nb = (n + kIB - 1) / kIB;
            if (M->tsync.n < (size_t)nb + 2) {
                M->tsync.alloc((size_t)nb + 2);
                k_fill_int<<<(unsigned)((nb + 2 + 255) / 256), 256, 0, s>>>(M->tsync.p, (long long)nb + 2);
                check("k_fill_int");
                M->tsync_tickets = 0;
                M->tsync_epoch = 0;
            }
            TrsvSync ts{M->tsync.p, M->tsync_tickets, M->tsync_epoch, M->ctx->num_cus};
            ProfScope prof(*M->ctx, "k_trsv_inv", 8.0 * n * (double)n);   // one read of both FP32 factors
            lu_solve_inv(s, (const float2 *)M->lu32.p, M->lda, n, (const float2 *)M->dinv32.p, x32, nrhs, ts);
        }
        k_widen<float><<<g, 256, 0, s>>>(x32, cnt, Xd);
        check("k_widen");
    }
}

double host_norm(const std::vector<double2> &v, int n, int nrhs, int q)
{
    double s = 0;
    for (int i = 0; i < n; ++i) s += v[(size_t)i * nrhs + q].x * v[(size_t)i * nrhs + q].x +
                                    v[(size_t)i * nrhs + q].y * v[(size_t)i * nrhs + q].y;
    return std::sqrt(s);
}

struct GmresWork {
    double2 *Vb, *Zb, *w, *r, *part, *hd;
};

__global__ void k_zero(double2 *x, long long cnt)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) x[i] = double2{0.0, 0.0};
}

// GMRES on A D = R0 for the q columns of R0 (row-major n x q) at once: one right-
// preconditioned flexible GMRES per column (A M^-1 u = r0, d = M^-1 u, with the
// preconditioned basis Z_k = M^-1 V_k kept, so d = Z y needs no further solve), FP64
// vectors and matrix products, classical Gram-Schmidt applied twice, Givens rotations,
// restarts from the recomputed true residual.  The columns share every launch (the
// factors and A are read once per step for all of them); each column's arithmetic is that
// of its own GMRES.  Column c stops at ||r0_c - A d_c|| <= atol[c] (a column starting
// there does no step) or after maxinner steps.  Returns the steps summed over the columns.
// Device workspace (restart r): Vb (r + 1) n q, Zb r n q, w n q, r n q, part 256 (r + 1) q,
// hd (r + 1) q complex doubles.
int gmres_block(bh_matrix *M, const double2 *R0, int q, const double *r0norm, double2 *D, const double *atol,
                int restart, int maxinner, const GmresWork &G)
{
    using C = std::complex<double>;
    hipStream_t s = M->ctx->stream;
    const int n = (int)M->n;
    const long long nq = (long long)n * q;
    const unsigned gq = (unsigned)((nq + 255) / 256);
    double2 *const Vb = G.Vb, *const Zb = G.Zb, *const w = G.w, *const r = G.r, *const part = G.part,
                   *const hd = G.hd;
    std::vector<double2> hp((size_t)256 * (restart + 1) * q);
    auto sum_cols = [&](int m, int j, int c) {   // sum of the 256 block partials of (j, c)
        double ax = 0, ay = 0;
        for (int b = 0; b < 256; ++b) {
            const double2 v = hp[((size_t)b * m + j) * q + c];
            ax += v.x;
            ay += v.y;
        }
        return C(ax, ay);
    };
    k_zero<<<gq, 256, 0, s>>>(D, nq);
    check("k_zero");
    std::vector<double> beta(r0norm, r0norm + q);
    std::vector<int> total(q, 0);
    std::vector<char> live(q);
    for (int c = 0; c < q; ++c) live[c] = beta[c] > atol[c];
    const double2 *rcur = R0;
    for (;;) {
        bool any = false;
        for (int c = 0; c < q; ++c) {
            if (live[c] && total[c] >= maxinner) live[c] = 0;
            any = any || live[c];
        }
        if (!any) break;
        ColScale sc{};
        for (int c = 0; c < q; ++c) sc.s[c] = live[c] ? 1.0 / beta[c] : 0.0;
        k_scale_cols<<<gq, 256, 0, s>>>(rcur, n, q, sc, Vb);
        check("k_scale_cols");
        std::vector<std::vector<C>> H(q, std::vector<C>((size_t)(restart + 1) * restart, 0.0)), cs(q, std::vector<C>(restart)),
            sn(q, std::vector<C>(restart)), gvec(q, std::vector<C>(restart + 1, 0.0));
        std::vector<int> kc(q, 0);          // steps of this cycle per column
        std::vector<char> run(live);        // columns still stepping in this cycle
        std::vector<char> conv(q, 0);
        for (int c = 0; c < q; ++c) gvec[c][0] = beta[c];
        for (int k = 0; k < restart; ++k) {
            bool stepping = false;
            for (int c = 0; c < q; ++c) stepping = stepping || run[c];
            if (!stepping) break;
            double2 *zk = Zb + (size_t)k * nq;
            apply_factors(M, Vb + (size_t)k * nq, zk, q);
            {
                ProfScope prof(*M->ctx, "k_matvec", 16.0 * n * (double)n);
                k_matvec_cols<<<(n + 3) / 4, 256, 0, s>>>((const double2 *)M->a64.p, n, zk, w, q);
            }
            check("k_matvec");
            std::vector<std::vector<C>> h(q, std::vector<C>(k + 1, 0.0));
            for (int pass = 0; pass < 2; ++pass) {
                k_dots_cols<<<256, 256, 0, s>>>(Vb, n, k + 1, q, w, part);
                check("k_dots");
                BH_CHECK_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double2) * 256 * (k + 1) * q,
                                            hipMemcpyDeviceToHost, s));
                BH_CHECK_HIP(hipStreamSynchronize(s));
                std::vector<double2> hh((size_t)(k + 1) * q, double2{0, 0});
                for (int c = 0; c < q; ++c) {
                    if (!run[c]) continue;
                    for (int j = 0; j <= k; ++j) {
                        const C v = sum_cols(k + 1, j, c);
                        hh[(size_t)j * q + c] = {v.real(), v.imag()};
                        h[c][j] += v;
                    }
                }
                BH_CHECK_HIP(hipMemcpyAsync(hd, hh.data(), sizeof(double2) * (k + 1) * q, hipMemcpyHostToDevice, s));
                k_axpys_cols<<<gq, 256, 0, s>>>(Vb, n, k + 1, q, hd, w);
                check("k_axpys");
            }
            k_dots_cols<<<256, 256, 0, s>>>(w, n, 1, q, w, part);
            check("k_dots");
            BH_CHECK_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double2) * 256 * q, hipMemcpyDeviceToHost, s));
            BH_CHECK_HIP(hipStreamSynchronize(s));
            ColScale nx{};
            for (int c = 0; c < q; ++c) {
                if (!run[c]) continue;
                const double hn = std::sqrt(std::max(0.0, sum_cols(1, 0, c).real()));
                auto &Hc = H[c];
                for (int j = 0; j <= k; ++j) Hc[(size_t)j * restart + k] = h[c][j];
                Hc[(size_t)(k + 1) * restart + k] = hn;
                for (int j = 0; j < k; ++j) {
                    const C x = Hc[(size_t)j * restart + k], y = Hc[(size_t)(j + 1) * restart + k];
                    Hc[(size_t)j * restart + k] = std::conj(cs[c][j]) * x + std::conj(sn[c][j]) * y;
                    Hc[(size_t)(j + 1) * restart + k] = -sn[c][j] * x + cs[c][j] * y;
                }
                const C x = Hc[(size_t)k * restart + k], y = Hc[(size_t)(k + 1) * restart + k];
                const double den = std::sqrt(std::norm(x) + std::norm(y));
                cs[c][k] = den == 0 ? C(1.0) : x / den;
                sn[c][k] = den == 0 ? C(0.0) : y / den;
                Hc[(size_t)k * restart + k] = den;
                Hc[(size_t)(k + 1) * restart + k] = 0.0;
                gvec[c][k + 1] = -sn[c][k] * gvec[c][k];
                gvec[c][k] = std::conj(cs[c][k]) * gvec[c][k];
                kc[c] = k + 1;
                ++total[c];
                if (std::abs(gvec[c][k + 1]) <= atol[c] || hn == 0) {
                    conv[c] = 1;
                    run[c] = 0;
                } else if (total[c] >= maxinner) {
                    run[c] = 0;
                } else {
                    nx.s[c] = 1.0 / hn;
                }
            }
            if (k + 1 < restart) {
                k_scale_cols<<<gq, 256, 0, s>>>(w, n, q, nx, Vb + (size_t)(k + 1) * nq);
                check("k_scale_cols");
            }
        }
        // D[:, c] += Z y_c with H_c y_c = g_c (upper triangular kc x kc)
        int kmax = 0;
        for (int c = 0; c < q; ++c) kmax = std::max(kmax, kc[c]);
        if (kmax > 0) {
            std::vector<double2> yn((size_t)kmax * q, double2{0, 0});
            for (int c = 0; c < q; ++c) {
                const int k = kc[c];
                const auto &Hc = H[c];
                std::vector<C> y(k, 0.0);
                for (int i = k - 1; i >= 0; --i) {
                    C t = gvec[c][i];
                    for (int j = i + 1; j < k; ++j) t -= Hc[(size_t)i * restart + j] * y[j];
                    y[i] = Hc[(size_t)i * restart + i] == 0.0 ? C(0.0) : t / Hc[(size_t)i * restart + i];
                }
                for (int j = 0; j < k; ++j) yn[(size_t)j * q + c] = {-y[j].real(), -y[j].imag()};
            }
            BH_CHECK_HIP(hipMemcpyAsync(hd, yn.data(), sizeof(double2) * kmax * q, hipMemcpyHostToDevice, s));
            k_axpys_cols<<<gq, 256, 0, s>>>(Zb, n, kmax, q, hd, D);
            check("k_axpys");
        }
        bool again = false;
        for (int c = 0; c < q; ++c) {
            if (conv[c]) live[c] = 0;
            if (live[c] && total[c] >= maxinner) live[c] = 0;
            again = again || live[c];
        }
        if (!again) break;
        // restart from the true residuals R0 - A D
        {
            ProfScope prof(*M->ctx, "k_residual", 16.0 * n * (double)n);
            k_residual<<<(n + 3) / 4, 256, 0, s>>>((const double2 *)M->a64.p, n, D, R0, r, q);
        }
        check("k_residual");
        k_dots_cols<<<256, 256, 0, s>>>(r, n, 1, q, r, part);
        check("k_dots");
        BH_CHECK_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double2) * 256 * q, hipMemcpyDeviceToHost, s));
        BH_CHECK_HIP(hipStreamSynchronize(s));
        for (int c = 0; c < q; ++c) {
            beta[c] = std::sqrt(std::max(0.0, sum_cols(1, 0, c).real()));
            if (live[c] && beta[c] <= atol[c]) live[c] = 0;
        }
        rcur = r;
    }
    BH_CHECK_HIP(hipStreamSynchronize(s));
    int sum = 0;
    for (int c = 0; c < q; ++c) sum += total[c];
    return sum;
}

}  // namespace

namespace {

constexpr int kGmresRestart = 30;

// Device workspace of one solve: db, dx, dr (n x nrhs each), rcol, dcol (n each) and the
// GMRES workspace, each piece 256-byte aligned inside one buffer.
// x += d elementwise
__global__ void k_add_to(double2 *__restrict__ x, long long cnt, const double2 *__restrict__ d)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) {
        x[i].x += d[i].x;
        x[i].y += d[i].y;
    }
}

struct SolveWork {
    double2 *db, *dx, *dr, *dd;
    GmresWork g;
};

size_t solve_work_layout(int n, int nrhs, char *base, SolveWork *W)
{
    size_t off = 0;
    auto take = [&](size_t count) {
        double2 *p = base ? (double2 *)(base + off) : nullptr;
        off += (count * sizeof(double2) + 255) & ~(size_t)255;
        return p;
    };
    const size_t nq = (size_t)n * nrhs, r = kGmresRestart;
    SolveWork w{};
    w.db = take(nq);
    w.dx = take(nq);
    w.dr = take(nq);
    w.dd = take(nq);
    w.g.Vb = take((r + 1) * nq);
    w.g.Zb = take(r * nq);
    w.g.w = take(nq);
    w.g.r = take(nq);
    w.g.part = take(256 * (r + 1) * (size_t)nrhs);
    w.g.hd = take((r + 1) * (size_t)nrhs);
    if (W) *W = w;
    return off;
}

// X = Z \ B with the factors of M (see bh_matrix_solve); W the device workspace.  With
// timed = false no HIP events are created (stats->seconds_solve is then host wall time).
void solve_impl(bh_matrix *M, int64_t nrhs, const double *B, double *X, double tol, int maxit,
                bh_solve_stats *stats, const SolveWork &W, bool timed)
{
    hipStream_t s = M->ctx->stream;
    const int n = (int)M->n, q = (int)nrhs;
    const long long cnt = (long long)n * q;
    // host column-major -> device row-major n x nrhs
    std::vector<double2> hb(cnt), hx(cnt);
    for (int r = 0; r < q; ++r)
        for (int i = 0; i < n; ++i) hb[(size_t)i * q + r] = {B[2 * ((size_t)r * n + i)], B[2 * ((size_t)r * n + i) + 1]};
    BH_CHECK_HIP(hipMemcpyAsync(W.db, hb.data(), sizeof(double2) * cnt, hipMemcpyHostToDevice, s));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const auto h0 = std::chrono::steady_clock::now();
    if (timed) {
        BH_CHECK_HIP(hipEventCreate(&e0));
        BH_CHECK_HIP(hipEventCreate(&e1));
        BH_CHECK_HIP(hipEventRecord(e0, s));
    }
    apply_factors(M, W.db, W.dx, q);
    int iters = 0, inner = 0;
    double res = 0;
    std::vector<double> bn(q);
    for (int r = 0; r < q; ++r) bn[r] = host_norm(hb, n, q, r);
    // the FP32 triangular solves report a block wait that gave up in tsync[1]; it is
    // read (and cleared) after every synchronisation that ends a use of the solves,
    // before anything derived from their results is judged
    auto check_trsv = [&]() {
        if (M->factored != BH_FP32 || !M->tsync.p) return;
        int trsv_err = 0;
        BH_CHECK_HIP(hipMemcpyAsync(&trsv_err, M->tsync.p + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        BH_CHECK_HIP(hipStreamSynchronize(s));
        if (!trsv_err) return;
        // results are void; the counters restart
        k_fill_int<<<(unsigned)((M->tsync.n + 255) / 256), 256, 0, s>>>(M->tsync.p, (long long)M->tsync.n);
        check("k_fill_int");
        BH_CHECK_HIP(hipStreamSynchronize(s));
        M->tsync_tickets = 0;
        M->tsync_epoch = 0;
        BH_FAIL(BH_ERR_HIP, "FP32 triangular solve: a block wait exceeded its limit");
    };
    const bool refine = M->factored == BH_FP32 || tol > 0;
    if (refine) {
        const int maxit_ = maxit > 0 ? maxit : 20;
        for (;;) {
            {
                ProfScope prof(*M->ctx, "k_residual", 16.0 * n * (double)n);   // one FP64 read of A
                k_residual<<<(n + 3) / 4, 256, 0, s>>>((const double2 *)M->a64.p, n, W.dx, W.db, W.dr, q);
            }
            check("k_residual");
            std::vector<double2> hr(cnt);
            BH_CHECK_HIP(hipMemcpyAsync(hr.data(), W.dr, sizeof(double2) * cnt, hipMemcpyDeviceToHost, s));
            BH_CHECK_HIP(hipStreamSynchronize(s));
            check_trsv();
            res = 0;
            std::vector<double> rel(q);
            for (int r = 0; r < q; ++r) {
                rel[r] = bn[r] > 0 ? host_norm(hr, n, q, r) / bn[r] : 0.0;
                if (!std::isfinite(rel[r])) BH_FAIL(BH_ERR_NOCONV, "refinement produced a non-finite residual");
                res = std::max(res, rel[r]);
            }
            const double target = tol > 0 ? tol : 1e-12;
            if (res <= target || iters >= maxit_) break;
            ++iters;
            // solve A D = R for the columns above the target, each to 1/2 of its remaining
            // target, then X += D (columns already there have D = 0)
            std::vector<double> rn(q), atol(q);
            for (int r = 0; r < q; ++r) {
                rn[r] = rel[r] * bn[r];
                atol[r] = rel[r] <= target ? std::numeric_limits<double>::infinity() : 0.5 * target * bn[r];
            }
            inner += gmres_block(M, W.dr, q, rn.data(), W.dd, atol.data(), kGmresRestart, 200, W.g);
            k_add_to<<<(unsigned)((cnt + 255) / 256), 256, 0, s>>>(W.dx, cnt, W.dd);
            check("k_add_to");
        }
    }
    float ms = 0;
    if (timed) {
        BH_CHECK_HIP(hipEventRecord(e1, s));
        BH_CHECK_HIP(hipEventSynchronize(e1));
        BH_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    BH_CHECK_HIP(hipMemcpyAsync(hx.data(), W.dx, sizeof(double2) * cnt, hipMemcpyDeviceToHost, s));
    BH_CHECK_HIP(hipStreamSynchronize(s));
    check_trsv();
    if (!timed) ms = (float)(1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - h0).count());
    for (int r = 0; r < q; ++r)
        for (int i = 0; i < n; ++i) {
            X[2 * ((size_t)r * n + i)] = hx[(size_t)i * q + r].x;
            X[2 * ((size_t)r * n + i) + 1] = hx[(size_t)i * q + r].y;
        }
    if (stats) {
        stats->iterations = iters;
        stats->inner = inner;
        stats->residual = res;
        stats->seconds_factor = M->seconds_factor;
        stats->seconds_solve = ms * 1e-3;
    }
    if (refine && tol > 0 && res > tol)
        BH_FAIL(BH_ERR_NOCONV, "refinement stopped at relative residual %.3e > %.3e after %d steps", res, tol, iters);
}

}  // namespace

extern "C" int bh_matrix_solve(bh_matrix *M, int64_t nrhs, const double *B, double *X, double tol, int maxit,
                               bh_solve_stats *stats)
{
    BH_API_BEGIN
    BH_REQUIRE(M && B && X, "NULL argument");
    BH_REQUIRE(M->factored == BH_FP64 || M->factored == BH_FP32, "matrix is not factored");
    BH_REQUIRE(nrhs >= 1 && nrhs <= 8, "nrhs must be 1..8");
    std::lock_guard<std::mutex> lock(M->mu);
    DevBuf<char> work(solve_work_layout((int)M->n, (int)nrhs, nullptr, nullptr));
    SolveWork W;
    solve_work_layout((int)M->n, (int)nrhs, work.p, &W);
    solve_impl(M, nrhs, B, X, tol, maxit, stats, W, true);
    BH_API_END
}

// ---------------------------------------------------------------- allocation-free solve
namespace {

// Scratch memory of bh_matrix_solve_noalloc: one copy per device, held until the process
// ends.  Host code never uses the array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device, when a context is created there.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

namespace bh {
// Called by bh_ctx_create with the context's device current.
void scratch_register(int device)
{
    BH_REQUIRE(device >= 0 && device < kMaxDevices, "device %d beyond the scratch table", device);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    BH_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BH_REQUIRE(p != nullptr, "no device address for the solve scratch on device %d", device);
    g_scratch_ptr[device] = p;
}
}  // namespace bh

extern "C" int bh_matrix_solve_noalloc(bh_matrix *M, int64_t nrhs, const double *B, double *X, double tol,
                                       int maxit, bh_solve_stats *stats)
{
    BH_API_BEGIN
    BH_REQUIRE(M && B && X, "NULL argument");
    BH_REQUIRE(M->factored == BH_FP64 || M->factored == BH_FP32, "matrix is not factored");
    BH_REQUIRE(nrhs >= 1 && nrhs <= 8, "nrhs must be 1..8");
    // apply_factors allocates nothing when bh_matrix_factor has sized these buffers
    BH_REQUIRE(M->factored == BH_FP64 ||
                   (M->x32.n >= (size_t)M->n * nrhs * 2 && M->tsync.n >= (size_t)((M->n + kIB - 1) / kIB) + 2),
               "FP32 factors without their solve buffers");
    const size_t need = solve_work_layout((int)M->n, (int)nrhs, nullptr, nullptr);
    BH_REQUIRE(need <= kScratchBytes,
               "bh_matrix_solve_noalloc: n = %lld with %lld right-hand sides needs %zu bytes of scratch, more than "
               "its %zu (use bh_matrix_solve)", (long long)M->n, (long long)nrhs, need, kScratchBytes);
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        base = (M->ctx->device >= 0 && M->ctx->device < kMaxDevices) ? g_scratch_ptr[M->ctx->device] : nullptr;
    }
    BH_REQUIRE(base != nullptr, "no solve scratch registered for device %d", M->ctx->device);
    std::lock_guard<std::mutex> lock(M->mu);
    SolveWork W;
    solve_work_layout((int)M->n, (int)nrhs, (char *)base, &W);
    solve_impl(M, nrhs, B, X, tol, maxit, stats, W, false);
    BH_API_END
}

extern "C" int bh_matrix_pivots(bh_matrix *M, int64_t *ipiv)
{
    BH_API_BEGIN
    BH_REQUIRE(M && ipiv, "NULL argument");
    BH_REQUIRE(M->factored >= 0, "matrix is not factored");
    std::vector<long long> ip(M->n);
    BH_CHECK_HIP(hipMemcpy(ip.data(), M->ipiv.p, sizeof(long long) * M->n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < M->n; ++i) ipiv[i] = ip[i] + 1;
    BH_API_END
}
