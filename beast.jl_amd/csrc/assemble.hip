// Galerkin assembly of MWSingleLayer3D / MWDoubleLayer3D on RT x RT on gfx950.
//
// Reference path (src/integralop.jl:69-184): for every test x trial cell pair,
// quadrule() picks a rule (doublenumwiltonsauterqstrat.jl:32-69), momintegrals!
// integrates the 3x3 local block, and the block is scattered with the signed
// element-to-dof map into Z.  Here that becomes:
//
//   K0  k_classify_count / k_classify_fill: FP64 pair classification, bit-exact
//       with the reference, compacting the non-far pairs of every test cell
//       into a CSR list in ascending trial order (deterministic).
//   K2  k_sauter_schwab: Sauter-Schwab blocks of touching pairs, one wavefront
//       per pair, launched per class so no wavefront diverges on the class.
//   K3  k_wilton: near pairs by singularity extraction (regular part by double
//       Gauss, singular part by analytic potential integrals), one wavefront per pair.
//   K1+K4 k_far_tiles: one workgroup per 64x64 tile of Z.  It computes the
//       double-Gauss blocks of every (row cell, column cell) pair the tile needs,
//       chunk by chunk into LDS (non-far pairs read their K2/K3 block), and
//       gathers them through the signed dof map into registers; each Z entry is
//       written once, coalesced, without atomics.  The accumulation order of an
//       entry depends only on its own mesh and space, so a request gives the
//       same bits alone or batched with others.
//
// Local blocks use coordinates relative to cell centroids, so the expanded
// moment formulas below do not cancel catastrophically (and the FP32 variant
// keeps ~1e-7 relative accuracy).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "bh_device.h"
#include "bh_internal.h"
#include "bh_matrix.h"

namespace bh {

struct OpParams {
    int kind;
    double gr, gi;   // gamma
    double ar, ai;   // alpha
    double br, bi;   // beta
};

// ============================================================ K0 classification
// Bodies of the per-request kernels take the block's (or wave's) item explicitly, so that the
// batched kernels (one launch for the items of several requests) run the same arithmetic.
__device__ void classify_count_dev(const double *__restrict__ tgeo, const double *__restrict__ sgeo, int nfs,
                                   double k2, int mode, int p, int *__restrict__ counts, int *__restrict__ err)
{
    __shared__ int wsum[4];
    double tv[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) tv[k] = tgeo[(size_t)p * kGeo + k];
    int cnt = 0, bad = 0;
    for (int q = threadIdx.x; q < nfs; q += blockDim.x) {
        const double *s = sgeo + (size_t)q * kGeo;
        double sv[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) sv[k] = s[k];
        int pk;
        const int c = classify_pair(tv, sv, s[kGeoA], k2, mode, &pk);
        if (c < 0) bad = 1;
        else cnt += (c != kFar);
    }
    if (bad) atomicOr(err, 1);
    for (int off = 32; off >= 1; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) counts[p] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(256) void k_classify_count(const double *__restrict__ tgeo, int nft,
                                                        const double *__restrict__ sgeo, int nfs, double k2,
                                                        int mode, int *__restrict__ counts, int *__restrict__ err)
{
    classify_count_dev(tgeo, sgeo, nfs, k2, mode, blockIdx.x, counts, err);
}

__device__ void classify_fill_dev(const double *__restrict__ tgeo, const double *__restrict__ sgeo, int nfs,
                                  double k2, int mode, const int *__restrict__ offsets, int4 *__restrict__ entries,
                                  int p)
{
    __shared__ int wtot[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double tv[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) tv[k] = tgeo[(size_t)p * kGeo + k];
    int base = offsets[p];
    for (int q0 = 0; q0 < nfs; q0 += 256) {
        const int q = q0 + threadIdx.x;
        int c = kFar, pk = 0;
        if (q < nfs) {
            const double *s = sgeo + (size_t)q * kGeo;
            double sv[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) sv[k] = s[k];
            c = classify_pair(tv, sv, s[kGeoA], k2, mode, &pk);
            if (c < 0) c = kFar;   // reported by the count pass
        }
        const bool flag = (c != kFar);
        const unsigned long long m = __ballot(flag);
        const int pre = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wtot[wid] = __popcll(m);
        __syncthreads();
        int off = 0;
        for (int w = 0; w < wid; ++w) off += wtot[w];
        const int tot = wtot[0] + wtot[1] + wtot[2] + wtot[3];
        if (flag) entries[base + off + pre] = make_int4(p, q, c, pk);
        base += tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_classify_fill(const double *__restrict__ tgeo, int nft,
                                                       const double *__restrict__ sgeo, int nfs, double k2, int mode,
                                                       const int *__restrict__ offsets, int4 *__restrict__ entries)
{
    classify_fill_dev(tgeo, sgeo, nfs, k2, mode, offsets, entries, blockIdx.x);
}

// Slot of this lane in list l (active lanes only): one atomic per distinct list of the
// wave, the lanes of a list ranked by lane (the lists' order is not significant: each
// entry's block is computed on its own and written to the entry's place)
__device__ __forceinline__ int list_slot(int *counters, int l, bool active)
{
    const int lane = threadIdx.x & 63;
    unsigned long long todo = __ballot(active);
    int slot = -1;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int lv = __shfl(l, leader, 64);
        const unsigned long long m = __ballot(active && l == lv);
        int base = 0;
        if (lane == leader) base = atomicAdd(&counters[lv], __popcll(m));
        base = __shfl(base, leader, 64);
        if (active && l == lv) slot = base + __popcll(m & ((1ull << lane) - 1ull));
        todo &= ~m;
    }
    return slot;
}

__global__ void k_class_lists(const int4 *__restrict__ entries, int n, int *__restrict__ counters,
                              int *__restrict__ lists, int cap)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = e < n;
    const int c = active ? entries[e].z : 0;   // 1..4
    const int slot = list_slot(counters, c, active);
    if (active) lists[(size_t)c * cap + slot] = e;
}

// Grid-restricted classification.  A pair without touching vertices is near
// only if dmin^2 < 0.64 area_s (max(dmin^2 |gamma|^2, dmin^2 / (16 area_s)) <
// 0.2^2 needs the second term below it), and touching needs dmin < dtol; so a
// pair whose centroids are at least D = sqrt(0.64 max area_s) + rho_t + rho_s
// apart (rho: the largest vertex-to-centroid distance of either mesh, D widened
// by a relative margin far above rounding) is far.  Trial cells are binned by
// centroid into a uniform grid of cell size >= D; a test cell classifies only
// the trial cells of the 27 grid cells around its own, with the same
// classify_pair as the full scan, and writes its non-far pairs in ascending
// trial order, so the compacted list is the full scan's exactly.
struct PairGridView {
    const int *start;   // [ncells + 1] into cells
    const int *cells;   // trial cell ids, ascending within each grid cell
    int nx, ny, nz;
    double ox, oy, oz, inv_h;
};
constexpr int kGridFillCap = 1024;   // non-far pairs of one test cell held in LDS

// The <= 27 neighbouring grid cells of centroid c: their ranges in cells[] and
// prefix counts (thread 0 fills, the caller synchronises).
__device__ void grid_ranges(const PairGridView &g, const double *c, int *b0, int *pre)
{
    const int ix = min(max((int)floor((c[0] - g.ox) * g.inv_h), 0), g.nx - 1);
    const int iy = min(max((int)floor((c[1] - g.oy) * g.inv_h), 0), g.ny - 1);
    const int iz = min(max((int)floor((c[2] - g.oz) * g.inv_h), 0), g.nz - 1);
    int r = 0, tot = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = ix + dx, y = iy + dy, z = iz + dz;
                int len = 0, b = 0;
                if (x >= 0 && x < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz) {
                    const int cell = (z * g.ny + y) * g.nx + x;
                    b = g.start[cell];
                    len = g.start[cell + 1] - b;
                }
                b0[r] = b;
                pre[r] = tot;
                tot += len;
                ++r;
            }
    pre[27] = tot;
}

__device__ __forceinline__ int grid_candidate(const PairGridView &g, const int *b0, const int *pre, int idx)
{
    int r = 0;
    while (pre[r + 1] <= idx) ++r;
    return g.cells[b0[r] + idx - pre[r]];
}

__device__ void classify_count_grid_dev(const double *__restrict__ tgeo, const double *__restrict__ sgeo,
                                        const PairGridView &g, double k2, int mode, int p, int *__restrict__ counts,
                                        int *__restrict__ err)
{
    __shared__ int wsum[4], b0[27], pre[28];
    double tv[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) tv[k] = tgeo[(size_t)p * kGeo + k];
    if (threadIdx.x == 0) grid_ranges(g, tgeo + (size_t)p * kGeo + kGeoC, b0, pre);
    __syncthreads();
    int cnt = 0, bad = 0;
    for (int idx = threadIdx.x; idx < pre[27]; idx += blockDim.x) {
        const int q = grid_candidate(g, b0, pre, idx);
        const double *sp = sgeo + (size_t)q * kGeo;
        double sv[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) sv[k] = sp[k];
        int pk;
        const int c = classify_pair(tv, sv, sp[kGeoA], k2, mode, &pk);
        if (c < 0) bad = 1;
        else cnt += (c != kFar);
    }
    if (bad) atomicOr(err, 1);
    for (int off = 32; off >= 1; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) counts[p] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(256) void k_classify_count_grid(const double *__restrict__ tgeo, int nft,
                                                             const double *__restrict__ sgeo, PairGridView g,
                                                             double k2, int mode, int *__restrict__ counts,
                                                             int *__restrict__ err)
{
    classify_count_grid_dev(tgeo, sgeo, g, k2, mode, blockIdx.x, counts, err);
}

__device__ void classify_fill_grid_dev(const double *__restrict__ tgeo, const double *__restrict__ sgeo,
                                       const PairGridView &g, double k2, int mode, const int *__restrict__ offsets,
                                       int4 *__restrict__ entries, int p)
{
    __shared__ int b0[27], pre[28], s_n;
    __shared__ int s_q[kGridFillCap], s_c[kGridFillCap], s_pk[kGridFillCap];
    double tv[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) tv[k] = tgeo[(size_t)p * kGeo + k];
    if (threadIdx.x == 0) {
        grid_ranges(g, tgeo + (size_t)p * kGeo + kGeoC, b0, pre);
        s_n = 0;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < pre[27]; idx += blockDim.x) {
        const int q = grid_candidate(g, b0, pre, idx);
        const double *sp = sgeo + (size_t)q * kGeo;
        double sv[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) sv[k] = sp[k];
        int pk;
        const int c = classify_pair(tv, sv, sp[kGeoA], k2, mode, &pk);
        if (c > kFar) {   // c < 0 is reported by the count pass
            const int slot = atomicAdd(&s_n, 1);
            if (slot < kGridFillCap) {
                s_q[slot] = q;
                s_c[slot] = c;
                s_pk[slot] = pk;
            }
        }
    }
    __syncthreads();
    const int m = min(s_n, kGridFillCap), base = offsets[p];   // s_n <= kGridFillCap: the host checks the counts
    for (int e = threadIdx.x; e < m; e += blockDim.x) {
        const int q = s_q[e];
        int rank = 0;
        for (int f = 0; f < m; ++f) rank += s_q[f] < q;
        entries[base + rank] = make_int4(p, q, s_c[e], s_pk[e]);
    }
}

__global__ __launch_bounds__(256) void k_classify_fill_grid(const double *__restrict__ tgeo, int nft,
                                                            const double *__restrict__ sgeo, PairGridView g,
                                                            double k2, int mode, const int *__restrict__ offsets,
                                                            int4 *__restrict__ entries)
{
    classify_fill_grid_dev(tgeo, sgeo, g, k2, mode, offsets, entries, blockIdx.x);
}

// Full class matrix (debug / bit-exact test entry).
__global__ void k_classify_matrix(const double *__restrict__ tgeo, int nft, const double *__restrict__ sgeo, int nfs,
                                  double k2, int mode, int8_t *__restrict__ out)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)nft * nfs) return;
    const int p = (int)(idx / nfs), q = (int)(idx % nfs);
    double tv[9], sv[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        tv[k] = tgeo[(size_t)p * kGeo + k];
        sv[k] = sgeo[(size_t)q * kGeo + k];
    }
    int pk;
    out[idx] = (int8_t)classify_pair(tv, sv, sgeo[(size_t)q * kGeo + kGeoA], k2, mode, &pk);
}

// ============================================================ shared helpers
// Complex Green's function G = exp(-gamma R) / (4 pi R).
template <typename R>
__device__ __forceinline__ cx<R> green(R gr, R gi, R dist, R idist)
{
    R s, c;
    if constexpr (sizeof(R) == 8) sincos_f64(gi * dist, &s, &c);
    else sincosf(gi * dist, &s, &c);
    R amp = idist * (R)kInv4Pi;
    if (gr != (R)0) {
        if constexpr (sizeof(R) == 8) amp *= exp(-gr * dist);
        else amp *= expf(-gr * dist);
    }
    return {c * amp, -s * amp};
}

// Block from EFIE moments: z_ij = ph s1 [alpha (Mxy - a_i.My - b_j.Mx + (a_i.b_j) M0) + 4 beta M0],
// with a_i.My and b_j.Mx formed once per i and per j, and s1 and the common phase ph of
// all moments folded into alpha and beta
template <typename R>
__device__ __forceinline__ void efie_block(const OpParams &op, R s1, const R *a, const R *b, cx<R> M0, const cx<R> *Mx,
                                          const cx<R> *My, cx<R> Mxy, cx<R> *z, int stride, cx<R> ph = {1, 0})
{
    const cx<R> sa = cmul(ph, cx<R>{s1 * (R)op.ar, s1 * (R)op.ai});
    const cx<R> bM0 = cmul(cmul(cx<R>{(R)op.br, (R)op.bi}, M0), ph);
    const cx<R> c4{(R)4 * s1 * bM0.r, (R)4 * s1 * bM0.i};
    cx<R> AY[3], BX[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const R *ai = a + 3 * i, *bi = b + 3 * i;
        AY[i] = {ai[0] * My[0].r + ai[1] * My[1].r + ai[2] * My[2].r, ai[0] * My[0].i + ai[1] * My[1].i + ai[2] * My[2].i};
        BX[i] = {Mxy.r - (bi[0] * Mx[0].r + bi[1] * Mx[1].r + bi[2] * Mx[2].r),
                 Mxy.i - (bi[0] * Mx[0].i + bi[1] * Mx[1].i + bi[2] * Mx[2].i)};
    }
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const R ab = dot3(a + 3 * i, b + 3 * j);
            const cx<R> t{BX[j].r - AY[i].r + ab * M0.r, BX[j].i - AY[i].i + ab * M0.i};
            z[(i + 3 * j) * stride] = {sa.r * t.r - sa.i * t.i + c4.r, sa.r * t.i + sa.i * t.r + c4.i};
        }
}

// Block from MFIE-K moments (common origin): z_ij = s1 [(b_j - a_i).Kc + (b_j x a_i).Kr]
template <typename R>
__device__ __forceinline__ void mfie_block(R s1, const R *a, const R *b, const cx<R> *Kc, const cx<R> *Kr, cx<R> *z,
                                          int stride)
{
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const R *ai = a + 3 * i, *bj = b + 3 * j;
            R d[3] = {bj[0] - ai[0], bj[1] - ai[1], bj[2] - ai[2]};
            R c[3];
            cross3(c, bj, ai);
            cx<R> t;
            t.r = d[0] * Kc[0].r + d[1] * Kc[1].r + d[2] * Kc[2].r + c[0] * Kr[0].r + c[1] * Kr[1].r + c[2] * Kr[2].r;
            t.i = d[0] * Kc[0].i + d[1] * Kc[1].i + d[2] * Kc[2].i + c[0] * Kr[0].i + c[1] * Kr[1].i + c[2] * Kr[2].i;
            z[(i + 3 * j) * stride] = {s1 * t.r, s1 * t.i};
        }
}

// ============================================================ K2 Sauter-Schwab
struct GLRule {
    int n;
    double x[kMaxGL];
    double w[kMaxGL];
};

// Region parameterisations of Sauter & Schwab, "Boundary Element Methods"
// (2011) Sec. 5.2, on the reference element {0 <= x2 <= x1 <= 1}.
__device__ __forceinline__ double ss_region(int cls, int r, double xi, double e1, double e2, double e3, double *xh,
                                            double *yh)
{
    if (cls == kCF) {
        switch (r) {
        case 0: xh[0] = xi; xh[1] = xi * (1 - e1 + e1 * e2); yh[0] = xi * (1 - e1 * e2 * e3); yh[1] = xi * (1 - e1); break;
        case 1: xh[0] = xi * (1 - e1 * e2 * e3); xh[1] = xi * (1 - e1); yh[0] = xi; yh[1] = xi * (1 - e1 + e1 * e2); break;
        case 2: xh[0] = xi; xh[1] = xi * e1 * (1 - e2 + e2 * e3); yh[0] = xi * (1 - e1 * e2); yh[1] = xi * e1 * (1 - e2); break;
        case 3: xh[0] = xi * (1 - e1 * e2); xh[1] = xi * e1 * (1 - e2); yh[0] = xi; yh[1] = xi * e1 * (1 - e2 + e2 * e3); break;
        case 4: xh[0] = xi * (1 - e1 * e2 * e3); xh[1] = xi * e1 * (1 - e2 * e3); yh[0] = xi; yh[1] = xi * e1 * (1 - e2); break;
        default: xh[0] = xi; xh[1] = xi * e1 * (1 - e2); yh[0] = xi * (1 - e1 * e2 * e3); yh[1] = xi * e1 * (1 - e2 * e3); break;
        }
        return xi * xi * xi * e1 * e1 * e2;
    }
    if (cls == kCE) {
        switch (r) {
        case 0: xh[0] = xi; xh[1] = xi * e1 * e3; yh[0] = xi * (1 - e1 * e2); yh[1] = xi * e1 * (1 - e2);
            return xi * xi * xi * e1 * e1;
        case 1: xh[0] = xi; xh[1] = xi * e1; yh[0] = xi * (1 - e1 * e2 * e3); yh[1] = xi * e1 * e2 * (1 - e3); break;
        case 2: xh[0] = xi * (1 - e1 * e2); xh[1] = xi * e1 * (1 - e2); yh[0] = xi; yh[1] = xi * e1 * e2 * e3; break;
        case 3: xh[0] = xi * (1 - e1 * e2 * e3); xh[1] = xi * e1 * e2 * (1 - e3); yh[0] = xi; yh[1] = xi * e1; break;
        default: xh[0] = xi * (1 - e1 * e2 * e3); xh[1] = xi * e1 * (1 - e2 * e3); yh[0] = xi; yh[1] = xi * e1 * e2; break;
        }
        return xi * xi * xi * e1 * e1 * e2;
    }
    if (r == 0) { xh[0] = xi; xh[1] = xi * e1; yh[0] = xi * e2; yh[1] = xi * e2 * e3; }
    else { xh[0] = xi * e2; xh[1] = xi * e2 * e3; yh[0] = xi; yh[1] = xi * e1; }
    return xi * xi * xi * e2;
}

__device__ __forceinline__ int ss_nregions(int cls) { return cls == kCF ? 6 : (cls == kCE ? 5 : 2); }

// Vertex order for the Sauter-Schwab parameterisation: shared vertices first
// (matched, in the order found with the test vertex outermost), then the
// remaining vertices in ascending local order.
__device__ __forceinline__ void ss_order(int cls, int pk, int *pi, int *pj)
{
    const int nh = cls == kCF ? 3 : (cls == kCE ? 2 : 1);
    int ui = 0, uj = 0;
    for (int h = 0; h < nh; ++h) {
        pi[h] = (pk >> (2 * h)) & 3;
        pj[h] = (pk >> (6 + 2 * h)) & 3;
        ui |= 1 << pi[h];
        uj |= 1 << pj[h];
    }
    int a = nh, b = nh;
    for (int k = 0; k < 3; ++k) {
        if (!(ui & (1 << k))) pi[a++] = k;
        if (!(uj & (1 << k))) pj[b++] = k;
    }
}

// one wave (lane = 0..63) computes the block of entry e
template <int OPK>
__device__ void ss_dev(const double *__restrict__ tgeo, const double *__restrict__ sgeo,
                       const int4 *__restrict__ entries, int e, int cls, const GLRule &gl, const OpParams &op,
                       double *__restrict__ blocks, int lane)
{
    const int4 en = entries[e];
    const double *tg = tgeo + (size_t)en.x * kGeo, *sg = sgeo + (size_t)en.y * kGeo;
    const double c[3] = {tg[kGeoC], tg[kGeoC + 1], tg[kGeoC + 2]};
    double a[9], b[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        a[k] = tg[k] - c[k % 3];
        b[k] = sg[k] - c[k % 3];
    }
    int pi[3], pj[3];
    ss_order(cls, en.w, pi, pj);
    double P0[3], dP1[3], dP2[3], Q0[3], dQ1[3], dQ2[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        P0[k] = a[3 * pi[0] + k];
        dP1[k] = a[3 * pi[1] + k] - P0[k];
        dP2[k] = a[3 * pi[2] + k] - a[3 * pi[1] + k];
        Q0[k] = b[3 * pj[0] + k];
        dQ1[k] = b[3 * pj[1] + k] - Q0[k];
        dQ2[k] = b[3 * pj[2] + k] - b[3 * pj[1] + k];
    }
    const int n = gl.n, n2 = n * n, n4 = n2 * n2;
    const int nreg = ss_nregions(cls);
    // moments: EFIE (M0, Mx, My, Mxy) / MFIE-K (Kc, Kr) / n x K (m0..m7, see below)
    constexpr int NM = OPK == BH_OP_MW_DOUBLELAYER_ROTATED ? 32 : 16;
    const double nrm[3] = {tg[kGeoN], tg[kGeoN + 1], tg[kGeoN + 2]};
    double m[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) m[k] = 0.0;
    for (int r = 0; r < nreg; ++r) {
        for (int idx = lane; idx < n4; idx += 64) {
            const int i0 = idx % n, i1 = (idx / n) % n, i2 = (idx / n2) % n, i3 = idx / (n2 * n);
            const double w4 = gl.w[i0] * gl.w[i1] * gl.w[i2] * gl.w[i3];
            double xh[2], yh[2];
            const double J = ss_region(cls, r, gl.x[i0], gl.x[i1], gl.x[i2], gl.x[i3], xh, yh);
            double x[3], y[3], rr[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                x[k] = P0[k] + xh[0] * dP1[k] + xh[1] * dP2[k];
                y[k] = Q0[k] + yh[0] * dQ1[k] + yh[1] * dQ2[k];
                rr[k] = x[k] - y[k];
            }
            const double R = sqrt(dot3(rr, rr));
            const double iR = 1.0 / R;
            const cx<double> G = green<double>(op.gr, op.gi, R, iR);
            const double wj = w4 * J;
            if constexpr (OPK == BH_OP_MW_SINGLELAYER) {
                const cx<double> wg{wj * G.r, wj * G.i};
                const double xy = dot3(x, y);
                m[0] += wg.r; m[1] += wg.i;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    m[2 + 2 * k] += wg.r * x[k]; m[3 + 2 * k] += wg.i * x[k];
                    m[8 + 2 * k] += wg.r * y[k]; m[9 + 2 * k] += wg.i * y[k];
                }
                m[14] += wg.r * xy; m[15] += wg.i * xy;
            } else if constexpr (OPK == BH_OP_MW_DOUBLELAYER_ROTATED) {
                // (x - a).r (n.(y - b)) - (x - a).(y - b) (n.r), h = -(gamma + 1/R) G / R:
                // m0 = h (x.r)(n.y), m1 = h x.r, m2 = h r (n.y), m3 = h r,
                // m4 = h (x.y)(n.r), m5 = h y (n.r), m6 = h x (n.r), m7 = h n.r
                const cx<double> gp{op.gr + iR, op.gi};
                cx<double> h = cmul(gp, G);
                h.r *= -iR * wj;
                h.i *= -iR * wj;
                const double ny = dot3(nrm, y), nr = dot3(nrm, rr), xr = dot3(x, rr), xy = dot3(x, y);
                m[0] += h.r * xr * ny; m[1] += h.i * xr * ny;
                m[2] += h.r * xr; m[3] += h.i * xr;
                m[16] += h.r * xy * nr; m[17] += h.i * xy * nr;
                m[30] += h.r * nr; m[31] += h.i * nr;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    m[4 + 2 * k] += h.r * rr[k] * ny; m[5 + 2 * k] += h.i * rr[k] * ny;
                    m[10 + 2 * k] += h.r * rr[k]; m[11 + 2 * k] += h.i * rr[k];
                    m[18 + 2 * k] += h.r * y[k] * nr; m[19 + 2 * k] += h.i * y[k] * nr;
                    m[24 + 2 * k] += h.r * x[k] * nr; m[25 + 2 * k] += h.i * x[k] * nr;
                }
            } else {
                // h = -(gamma + 1/R) G / R
                const cx<double> gp{op.gr + iR, op.gi};
                cx<double> h = cmul(gp, G);
                h.r *= -iR * wj;
                h.i *= -iR * wj;
                double xy[3];
                cross3(xy, x, y);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    m[2 * k] += h.r * xy[k]; m[2 * k + 1] += h.i * xy[k];
                    m[6 + 2 * k] += h.r * rr[k]; m[7 + 2 * k] += h.i * rr[k];
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) m[k] = wave_sum(m[k]);
    if (lane != 0) return;
    cx<double> z[9];
    if constexpr (OPK == BH_OP_MW_DOUBLELAYER_ROTATED) {
        // z_ij = m0 - nb m1 - a.m2 + nb a.m3 - m4 + a.m5 + b.m6 - (a.b) m7, nb = n.b_j
        const cx<double> m0{m[0], m[1]}, m1{m[2], m[3]}, m4{m[16], m[17]}, m7{m[30], m[31]};
        const cx<double> m2[3] = {{m[4], m[5]}, {m[6], m[7]}, {m[8], m[9]}};
        const cx<double> m3[3] = {{m[10], m[11]}, {m[12], m[13]}, {m[14], m[15]}};
        const cx<double> m5[3] = {{m[18], m[19]}, {m[20], m[21]}, {m[22], m[23]}};
        const cx<double> m6[3] = {{m[24], m[25]}, {m[26], m[27]}, {m[28], m[29]}};
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double *ai = a + 3 * i, *bj = b + 3 * j;
                const double nb = dot3(nrm, bj), ab = dot3(ai, bj);
                z[i + 3 * j] = {m0.r - nb * m1.r - dot3c_r(ai, m2) + nb * dot3c_r(ai, m3) - m4.r + dot3c_r(ai, m5) +
                                    dot3c_r(bj, m6) - ab * m7.r,
                                m0.i - nb * m1.i - dot3c_i(ai, m2) + nb * dot3c_i(ai, m3) - m4.i + dot3c_i(ai, m5) +
                                    dot3c_i(bj, m6) - ab * m7.i};
            }
    } else if constexpr (OPK == BH_OP_MW_SINGLELAYER) {
        const cx<double> M0{m[0], m[1]};
        const cx<double> Mx[3] = {{m[2], m[3]}, {m[4], m[5]}, {m[6], m[7]}};
        const cx<double> My[3] = {{m[8], m[9]}, {m[10], m[11]}, {m[12], m[13]}};
        const cx<double> Mxy{m[14], m[15]};
        efie_block<double>(op, 1.0, a, b, M0, Mx, My, Mxy, z, 1);
    } else {
        const cx<double> Kc[3] = {{m[0], m[1]}, {m[2], m[3]}, {m[4], m[5]}};
        const cx<double> Kr[3] = {{m[6], m[7]}, {m[8], m[9]}, {m[10], m[11]}};
        mfie_block<double>(1.0, a, b, Kc, Kr, z, 1);
    }
    double *out = blocks + (size_t)e * 18;
#pragma unroll
    for (int k = 0; k < 9; ++k) { out[2 * k] = z[k].r; out[2 * k + 1] = z[k].i; }
}

template <int OPK>
__global__ __launch_bounds__(256) void k_sauter_schwab(const double *__restrict__ tgeo, const double *__restrict__ sgeo,
                                                       const int4 *__restrict__ entries, const int *__restrict__ list,
                                                       int nlist, int cls, GLRule gl, OpParams op,
                                                       double *__restrict__ blocks)
{
    const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wv >= nlist) return;
    ss_dev<OPK>(tgeo, sgeo, entries, list[wv], cls, gl, op, blocks, threadIdx.x & 63);
}

// ============================================================ K3 Wilton near
// Analytic integrals over the triangle (b0, b1, b2) seen from x (all in one
// frame): S[-1,0,1] = int R^k, V[-1,0,1] = int (y - rho) R^k, G[-1] = int grad_y 1/R,
// G[1] = int grad_y R; rho is the projection of x on the triangle plane.
// Formulation: D. R. Wilton et al., IEEE TAP 32 (1984) 276-281 and R. D. Graglia,
// IEEE TAP 41 (1993) 1448-1455: edge integrals I_e^k = int_e R^k dl and the
// recursions sum_e p_e I_e^q = (q+2) int R^q - q d^2 int R^(q-2),
// int (y-rho) R^q = 1/(q+2) sum_e m_e I_e^(q+2).
struct WiltonOut {
    double Sm1, S0, S1;
    double Vm1[3], V0[3], V1[3];
    double Gm1[3], G1[3];
    double rho[3];
};

__device__ void wilton_integrals(const double *b, const double *x, double area, WiltonOut &o)
{
    double e13[3] = {b[0] - b[6], b[1] - b[7], b[2] - b[8]};
    double e23[3] = {b[3] - b[6], b[4] - b[7], b[5] - b[8]};
    double n[3];
    cross3(n, e13, e23);
    const double nn = sqrt(dot3(n, n));
    n[0] /= nn; n[1] /= nn; n[2] /= nn;
    const double xb[3] = {x[0] - b[0], x[1] - b[1], x[2] - b[2]};
    const double d = dot3(xb, n), D = fabs(d);
    for (int k = 0; k < 3; ++k) o.rho[k] = x[k] - d * n[k];
    double Sm1 = 0, S1 = 0, Om = 0;
    for (int k = 0; k < 3; ++k) { o.Vm1[k] = 0; o.V0[k] = 0; o.V1[k] = 0; o.Gm1[k] = 0; }
    for (int e = 0; e < 3; ++e) {
        const double *pa = b + 3 * e, *pb = b + 3 * ((e + 1) % 3);
        double t[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
        const double l = sqrt(dot3(t, t));
        t[0] /= l; t[1] /= l; t[2] /= l;
        double m[3];
        cross3(m, t, n);
        const double ar[3] = {pa[0] - o.rho[0], pa[1] - o.rho[1], pa[2] - o.rho[2]};
        const double sm = dot3(ar, t), sp = sm + l;
        const double p = dot3(ar, m);
        const double R02 = p * p + d * d;
        const double Rm = sqrt(sm * sm + R02), Rp = sqrt(sp * sp + R02);
        double Im1;
        if (sm >= 0) Im1 = log((sp + Rp) / (sm + Rm));
        else if (sp <= 0) Im1 = log((Rm - sm) / (Rp - sp));
        else Im1 = log((sp + Rp) * (Rm - sm) / R02);
        const double I1 = 0.5 * (sp * Rp - sm * Rm) + 0.5 * R02 * Im1;
        const double I2 = (sp * sp * sp - sm * sm * sm) / 3.0 + R02 * (sp - sm);
        const double I3 = 0.25 * (sp * Rp * Rp * Rp - sm * Rm * Rm * Rm) + 0.375 * R02 * (sp * Rp - sm * Rm) +
                          0.375 * R02 * R02 * Im1;
        const double np = p * sp, nm = p * sm;
        const double bp = np == 0.0 ? 0.0 : atan(np / (R02 + D * Rp));
        const double bm = nm == 0.0 ? 0.0 : atan(nm / (R02 + D * Rm));
        Om += bp - bm;
        Sm1 += p * Im1;
        S1 += p * I1;
        for (int k = 0; k < 3; ++k) {
            o.Vm1[k] += m[k] * I1;
            o.V0[k] += m[k] * (0.5 * I2);
            o.V1[k] += m[k] * (I3 / 3.0);
            o.Gm1[k] += m[k] * Im1;
        }
    }
    Sm1 -= D * Om;
    o.Sm1 = Sm1;
    o.S0 = area;
    o.S1 = (S1 + d * d * Sm1) / 3.0;
    const double sgnOm = d > 0 ? Om : (d < 0 ? -Om : 0.0);
    for (int k = 0; k < 3; ++k) {
        o.Gm1[k] += n[k] * sgnOm;
        o.G1[k] = o.Vm1[k] - d * n[k] * Sm1;
    }
}

// exp(z) - 1 for complex z, accurate for small |z|.
__device__ __forceinline__ cx<double> cexpm1(double a, double b)
{
    const double sb2 = sin(0.5 * b);
    return {expm1(a) * cos(b) - 2.0 * sb2 * sb2, exp(a) * sin(b)};
}

template <int OPK>
__device__ void wilton_dev(const double *__restrict__ tgeo, const double *__restrict__ sgeo,
                           const int4 *__restrict__ entries, int e, const double *__restrict__ tq, int Qo,
                           const double *__restrict__ sq, int Qi, const OpParams &op, double *__restrict__ blocks,
                           int lane)
{
    const int4 en = entries[e];
    const double *tg = tgeo + (size_t)en.x * kGeo, *sg = sgeo + (size_t)en.y * kGeo;
    const double c[3] = {tg[kGeoC], tg[kGeoC + 1], tg[kGeoC + 2]};
    double a[9], b[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        a[k] = tg[k] - c[k % 3];
        b[k] = sg[k] - c[k % 3];
    }
    const double dcs[3] = {sg[kGeoC] - c[0], sg[kGeoC + 1] - c[1], sg[kGeoC + 2] - c[2]};
    const double At = tg[kGeoA], As = sg[kGeoA];
    const double *xq = tq + (size_t)en.x * Qo * 4, *yq = sq + (size_t)en.y * Qi * 4;
    double m[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = 0.0;
    // regular part: double Gauss with G - (1/R - gamma + gamma^2 R / 2) / (4 pi)
    for (int idx = lane; idx < Qo * Qi; idx += 64) {
        const int p = idx / Qi, q = idx % Qi;
        const double x[3] = {xq[4 * p], xq[4 * p + 1], xq[4 * p + 2]};
        const double y[3] = {yq[4 * q] + dcs[0], yq[4 * q + 1] + dcs[1], yq[4 * q + 2] + dcs[2]};
        const double w = xq[4 * p + 3] * yq[4 * q + 3];
        const double rr[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
        const double R = sqrt(dot3(rr, rr));
        const double gRr = op.gr * R, gRi = op.gi * R;
        if constexpr (OPK == BH_OP_MW_SINGLELAYER) {
            // (expm1(-gR) + gR - (gR)^2/2) / (4 pi R)
            const cx<double> em = cexpm1(-gRr, -gRi);
            const double sq2r = gRr * gRr - gRi * gRi, sq2i = 2.0 * gRr * gRi;
            const double f = w / (4.0 * kPi * R);
            const cx<double> wg{(em.r + gRr - 0.5 * sq2r) * f, (em.i + gRi - 0.5 * sq2i) * f};
            const double xy = dot3(x, y);
            m[0] += wg.r; m[1] += wg.i;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                m[2 + 2 * k] += wg.r * x[k]; m[3 + 2 * k] += wg.i * x[k];
                m[8 + 2 * k] += wg.r * y[k]; m[9 + 2 * k] += wg.i * y[k];
            }
            m[14] += wg.r * xy; m[15] += wg.i * xy;
        } else {
            // (-(gR + 1) exp(-gR) + 1 - (gR)^2/2) / (4 pi R^3)
            double s, co;
            sincos(gRi, &s, &co);
            const double ea = exp(-gRr);
            const cx<double> ex{ea * co, -ea * s};
            const cx<double> t = cmul(cx<double>{gRr + 1.0, gRi}, ex);
            const double sq2r = gRr * gRr - gRi * gRi, sq2i = 2.0 * gRr * gRi;
            const double f = w / (4.0 * kPi * R * R * R);
            const cx<double> h{(-t.r + 1.0 - 0.5 * sq2r) * f, (-t.i - 0.5 * sq2i) * f};
            double xy[3];
            cross3(xy, x, y);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                m[2 * k] += h.r * xy[k]; m[2 * k + 1] += h.i * xy[k];
                m[6 + 2 * k] += h.r * rr[k]; m[7 + 2 * k] += h.i * rr[k];
            }
        }
    }
    // singular part: analytic inner integrals at each outer point
    if (lane < Qo) {
        const double x[3] = {xq[4 * lane], xq[4 * lane + 1], xq[4 * lane + 2]};
        const double w = xq[4 * lane + 3];
        WiltonOut o;
        wilton_integrals(b, x, As, o);
        const cx<double> g{op.gr, op.gi};
        const cx<double> g2 = cmul(g, g);
        if constexpr (OPK == BH_OP_MW_SINGLELAYER) {
            const double f = w / (4.0 * kPi);
            const cx<double> IG{(o.Sm1 - g.r * o.S0 + 0.5 * g2.r * o.S1) * f, (-g.i * o.S0 + 0.5 * g2.i * o.S1) * f};
            cx<double> Y[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Y[k].r = (o.Vm1[k] - g.r * o.V0[k] + 0.5 * g2.r * o.V1[k]) * f + o.rho[k] * IG.r;
                Y[k].i = (-g.i * o.V0[k] + 0.5 * g2.i * o.V1[k]) * f + o.rho[k] * IG.i;
            }
            m[0] += IG.r; m[1] += IG.i;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                m[2 + 2 * k] += IG.r * x[k]; m[3 + 2 * k] += IG.i * x[k];
                m[8 + 2 * k] += Y[k].r; m[9 + 2 * k] += Y[k].i;
            }
            m[14] += x[0] * Y[0].r + x[1] * Y[1].r + x[2] * Y[2].r;
            m[15] += x[0] * Y[0].i + x[1] * Y[1].i + x[2] * Y[2].i;
        } else {
            const double f = w / (4.0 * kPi);
            cx<double> IdG[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                IdG[k].r = (-o.Gm1[k] - 0.5 * g2.r * o.G1[k]) * f;
                IdG[k].i = (-0.5 * g2.i * o.G1[k]) * f;
            }
            // Kc += IdG x x, Kr += IdG
            const double cr[3] = {IdG[1].r * x[2] - IdG[2].r * x[1], IdG[2].r * x[0] - IdG[0].r * x[2],
                                  IdG[0].r * x[1] - IdG[1].r * x[0]};
            const double ci[3] = {IdG[1].i * x[2] - IdG[2].i * x[1], IdG[2].i * x[0] - IdG[0].i * x[2],
                                  IdG[0].i * x[1] - IdG[1].i * x[0]};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                m[2 * k] += cr[k]; m[2 * k + 1] += ci[k];
                m[6 + 2 * k] += IdG[k].r; m[7 + 2 * k] += IdG[k].i;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = wave_sum(m[k]);
    if (lane != 0) return;
    const double s1 = 1.0 / (4.0 * At * As);
    cx<double> z[9];
    if constexpr (OPK == BH_OP_MW_SINGLELAYER) {
        const cx<double> M0{m[0], m[1]};
        const cx<double> Mx[3] = {{m[2], m[3]}, {m[4], m[5]}, {m[6], m[7]}};
        const cx<double> My[3] = {{m[8], m[9]}, {m[10], m[11]}, {m[12], m[13]}};
        const cx<double> Mxy{m[14], m[15]};
        efie_block<double>(op, s1, a, b, M0, Mx, My, Mxy, z, 1);
    } else {
        const cx<double> Kc[3] = {{m[0], m[1]}, {m[2], m[3]}, {m[4], m[5]}};
        const cx<double> Kr[3] = {{m[6], m[7]}, {m[8], m[9]}, {m[10], m[11]}};
        mfie_block<double>(s1, a, b, Kc, Kr, z, 1);
    }
    double *out = blocks + (size_t)e * 18;
#pragma unroll
    for (int k = 0; k < 9; ++k) { out[2 * k] = z[k].r; out[2 * k + 1] = z[k].i; }
}

template <int OPK>
__global__ __launch_bounds__(256) void k_wilton(const double *__restrict__ tgeo, const double *__restrict__ sgeo,
                                                const int4 *__restrict__ entries, const int *__restrict__ list,
                                                int nlist, const double *__restrict__ tq, int Qo,
                                                const double *__restrict__ sq, int Qi, OpParams op,
                                                double *__restrict__ blocks)
{
    const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wv >= nlist) return;
    wilton_dev<OPK>(tgeo, sgeo, entries, list[wv], tq, Qo, sq, Qi, op, blocks, threadIdx.x & 63);
}

// ============================================================ K1 far field
struct FarArgs {
    const double *tgeo, *sgeo;
    const void *tq, *sq;          // [nf][Q][4] in R: x', y', z' (rel. centroid), weight
    int Qo, Qi;
    int mode;
    double k2;
    int shifted;                  // FP64, lossless: phases taken relative to the centroid distance
    const int *nptr;              // CSR of non-far pairs per test cell
    const int4 *nent;
    const double *nblk;           // 18 doubles per entry
    OpParams op;
};

constexpr int kTileCells = 128;    // max cells per dof-tile side

// The test cells of a dof tile staged in LDS field by field (kTileCells apart), so that
// the lanes of a wave, which hold consecutive test cells, read consecutive words.
template <typename R>
struct RowStage {
    const double *cen;    // [3][kTileCells] centroid
    const R *av;          // [9][kTileCells] vertices relative to the centroid
    const double *area;   // [kTileCells]
    const R *pts;         // [Qo * 4][kTileCells] points relative to the centroid, weight
    int rc;               // the test cell's place in the tile
};

// Far-field 3x3 block of (test p, trial q) by double Gauss, written to z[k*stride].
// QO, QI: the point counts when fixed at compile time (0: read from A).
// maybe_near = false: the caller knows the pair is not in p's non-far list.
// STAGED: the test cell's data come from st (LDS) instead of A's arrays.
template <typename R, int OPK, int QO = 0, int QI = 0, bool SHIFT = false, bool STAGED = false>
__device__ __forceinline__ void far_block(const FarArgs &A, int p, int q, cx<R> *z, int stride, bool maybe_near = true,
                                          RowStage<R> st = {})
{
    const int Qo = QO ? QO : A.Qo, Qi = QI ? QI : A.Qi;
    // the test-point loop of the small default rules is unrolled too (independent point
    // pairs of several test points in flight: measured 10% faster at 3 x 4 points)
    constexpr int kUnrollO = (QO > 0 && QO <= 4) ? QO : 1;
    const double *tg = A.tgeo + (size_t)p * kGeo, *sg = A.sgeo + (size_t)q * kGeo;
    if (A.mode != kModeDoubleNum && maybe_near) {
        // the classification pass listed p's non-far trial cells in ascending
        // order with their blocks: a pair found there is not far
        const int end = A.nptr[p + 1];
        int lo = A.nptr[p], hi = end;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (A.nent[mid].y < q) lo = mid + 1;
            else hi = mid;
        }
        if (lo < end && A.nent[lo].y == q) {
            const double *blk = A.nblk + (size_t)lo * 18;
#pragma unroll
            for (int k = 0; k < 9; ++k) z[k * stride] = {(R)blk[2 * k], (R)blk[2 * k + 1]};
            return;
        }
    }
    R dc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) dc[k] = (R)((STAGED ? st.cen[k * kTileCells + st.rc] : tg[kGeoC + k]) - sg[kGeoC + k]);
    // vertices relative to the centroids are only needed once the moments are
    // summed; loading them late keeps them out of the point loops' registers
    auto load_verts = [&](R *a, R *b) {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            a[k] = STAGED ? st.av[k * kTileCells + st.rc] : (R)(tg[k] - tg[kGeoC + k % 3]);
            b[k] = (R)(sg[k] - sg[kGeoC + k % 3]);
        }
    };
    const R gr = (R)A.op.gr, gi = (R)A.op.gi;
    // shifted: exp(-i gi R) = exp(-i gi Rc) exp(-i gi (R - Rc)), Rc the centroid distance,
    // the second factor by its Taylor series (|gi (R - Rc)| <= |gi| (r_p + r_q), bounded on
    // the host so that the series' remainder stays below 1e-17)
    R cc = (R)1, sc = (R)0, rc = (R)0;
    constexpr bool shifted = SHIFT && sizeof(R) == 8;
    if constexpr (shifted) {
        rc = sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
        double s_, c_;
        sincos_f64((double)(gi * rc), &s_, &c_);
        sc = (R)s_;
        cc = (R)c_;
    }
    auto kernel_g = [&](R dist, R idist) -> cx<R> {
        if constexpr (!shifted) return green<R>(gr, gi, dist, idist);
        const R t = gi * (dist - rc), u = t * t;
        // cos t = sum (-1)^k u^k / (2k)!, sin t = t sum (-1)^k u^k / (2k+1)!, k <= 6
        R pc = (R)(1.0 / 479001600.0), ps = (R)(1.0 / 6227020800.0);
        pc = pc * u - (R)(1.0 / 3628800.0);
        ps = ps * u - (R)(1.0 / 39916800.0);
        pc = pc * u + (R)(1.0 / 40320.0);
        ps = ps * u + (R)(1.0 / 362880.0);
        pc = pc * u - (R)(1.0 / 720.0);
        ps = ps * u - (R)(1.0 / 5040.0);
        pc = pc * u + (R)(1.0 / 24.0);
        ps = ps * u + (R)(1.0 / 120.0);
        pc = pc * u - (R)0.5;
        ps = ps * u - (R)(1.0 / 6.0);
        const R ct = pc * u + (R)1, st = t * (ps * u + (R)1);
        const R amp = idist * (R)kInv4Pi;
        // exp(-i gi (R - Rc)) / (4 pi R); the pair's common factor exp(-i gi Rc) = cc - i sc
        // multiplies its moments once (rot below), every quantity summed being linear in G
        return {ct * amp, -st * amp};
    };
    // moments (or block entries) times exp(-i gi Rc) when the phases are shifted
    auto rot = [&](cx<R> &v) {
        if constexpr (shifted) v = {cc * v.r + sc * v.i, cc * v.i - sc * v.r};
    };
    // test point v (coordinates 4 ip .. 4 ip + 2, weight 4 ip + 3)
    const R *xq = STAGED ? st.pts + st.rc : (const R *)A.tq + (size_t)p * Qo * 4;
    auto xpt = [&](int v) -> R { return xq[STAGED ? v * kTileCells : v]; };
    auto t_area = [&]() -> double { return STAGED ? st.area[st.rc] : tg[kGeoA]; };   // read where used
    const R *yq = (const R *)A.sq + (size_t)q * Qi * 4;
    if constexpr (OPK == BH_OP_MW_SINGLELAYER) {
        cx<R> T0{0, 0}, T1{0, 0}, Tx[3] = {{0, 0}, {0, 0}, {0, 0}}, Ty[3] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll kUnrollO
        for (int ip = 0; ip < Qo; ++ip) {
            const R x[3] = {xpt(4 * ip), xpt(4 * ip + 1), xpt(4 * ip + 2)};
            const R wp = xpt(4 * ip + 3);
            const R xd[3] = {x[0] + dc[0], x[1] + dc[1], x[2] + dc[2]};
            cx<R> S0{0, 0}, S1[3] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll
            for (int iq = 0; iq < Qi; ++iq) {
                const R y[3] = {yq[4 * iq], yq[4 * iq + 1], yq[4 * iq + 2]};
                const R wq = yq[4 * iq + 3];
                const R r[3] = {xd[0] - y[0], xd[1] - y[1], xd[2] - y[2]};
                const R R2 = dot3(r, r);
                R dist, idist;
                if constexpr (sizeof(R) == 8) { idist = rsqrt_f64(R2); dist = R2 * idist; }
                else { idist = rsqrtf(R2); dist = R2 * idist; }
                const cx<R> G = kernel_g(dist, idist);
                const cx<R> wg{wq * G.r, wq * G.i};
                S0.r += wg.r; S0.i += wg.i;
#pragma unroll
                for (int k = 0; k < 3; ++k) { S1[k].r += wg.r * y[k]; S1[k].i += wg.i * y[k]; }
            }
            T0.r += wp * S0.r; T0.i += wp * S0.i;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Tx[k].r += wp * x[k] * S0.r; Tx[k].i += wp * x[k] * S0.i;
                Ty[k].r += wp * S1[k].r; Ty[k].i += wp * S1[k].i;
            }
            T1.r += wp * (x[0] * S1[0].r + x[1] * S1[1].r + x[2] * S1[2].r);
            T1.i += wp * (x[0] * S1[0].i + x[1] * S1[1].i + x[2] * S1[2].i);
        }
        R a[9], b[9];
        load_verts(a, b);
        const R s1 = (R)(1.0 / (4.0 * t_area() * sg[kGeoA]));
        // the moments' common factor exp(-i gi Rc) (shifted phases) enters with alpha and beta
        efie_block<R>(A.op, s1, a, b, T0, Tx, Ty, T1, z, stride, cx<R>{cc, -sc});
    } else if constexpr (OPK == BH_OP_MW_DOUBLELAYER_ROTATED) {
        // On a flat cell paired with itself n . r = n . g_j = 0 for every point pair, so the
        // block is exactly zero; in FP32 the rounding of those dot products times the 1/R^2
        // kernel would leave noise far above the FP32 far-field error, so it is written as zero
        // (FP64 keeps the computed value, below 1e-15 of the neighbouring entries).
        if constexpr (sizeof(R) == 4) {
            if (p == q && A.tgeo == A.sgeo) {
#pragma unroll
                for (int k = 0; k < 9; ++k) z[k * stride] = {(R)0, (R)0};
                return;
            }
        }
        // f_i . (n x (K x g_j)) = (f_i . K)(n . g_j) - (f_i . g_j)(n . K), K = h r, r = xd - y,
        // n the test normal, f_i = u_i / (2 A_t), g_j = (y - b_j) / (2 A_s).  With u_i = x - a_i
        // and v_j = xd - b_j (so y - b_j = v_j - r, and the r (n . r) terms cancel):
        // 4 A_t A_s f_i . (n x (K x g_j)) = h [(u_i . r)(n . v_j) - (u_i . v_j)(n . r)].
        // Per test point only Rs = sum w h r and Ss = sum w h (n . r) over the trial points
        // are needed; r itself enters, so close point pairs do not cancel.
        const R nrm[3] = {(R)tg[kGeoN], (R)tg[kGeoN + 1], (R)tg[kGeoN + 2]};
        R a[9], b[9];
        load_verts(a, b);
        cx<R> zz[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) zz[k] = {0, 0};
#pragma unroll kUnrollO
        for (int ip = 0; ip < Qo; ++ip) {
            const R x[3] = {xpt(4 * ip), xpt(4 * ip + 1), xpt(4 * ip + 2)};
            const R wp = xpt(4 * ip + 3);
            const R xd[3] = {x[0] + dc[0], x[1] + dc[1], x[2] + dc[2]};
            cx<R> Rs[3] = {{0, 0}, {0, 0}, {0, 0}}, Ss{0, 0};
            // the trial points are read again for every test point (L1 hits) instead of being
            // held in registers across the test-point loop (13 points: 104 VGPRs), which
            // lets the 12 x 13-point tiles run two waves per SIMD
            const R *yqi = yq;
            asm volatile("" : "+v"(yqi));
#pragma unroll
            for (int iq = 0; iq < Qi; ++iq) {
                const R y[3] = {yqi[4 * iq], yqi[4 * iq + 1], yqi[4 * iq + 2]};
                const R wq = yqi[4 * iq + 3];
                const R r[3] = {xd[0] - y[0], xd[1] - y[1], xd[2] - y[2]};
                const R R2 = dot3(r, r);
                R dist, idist;
                if constexpr (sizeof(R) == 8) { idist = rsqrt_f64(R2); dist = R2 * idist; }
                else { idist = rsqrtf(R2); dist = R2 * idist; }
                const cx<R> G = kernel_g(dist, idist);
                cx<R> h = cmul(cx<R>{gr + idist, gi}, G);
                const R f = -idist * wq;
                h.r *= f; h.i *= f;
                if constexpr (sizeof(R) == 4) {   // FP64 forms Ss = n . Rs after the loop
                    const R nr = dot3(nrm, r);
                    Ss.r += h.r * nr; Ss.i += h.i * nr;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) { Rs[k].r += h.r * r[k]; Rs[k].i += h.i * r[k]; }
            }
            // sum w h (n . r) = n . sum w h r: in FP64 the dot product with the normal is taken once
            // per test point (in FP32 per point pair, where the rounding of n . r on flat cells
            // stays at the level the coincident-cell case above zeroes)
            if constexpr (sizeof(R) == 8) Ss = {dot3c_r(nrm, Rs), dot3c_i(nrm, Rs)};
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const R v[3] = {xd[0] - b[3 * j], xd[1] - b[3 * j + 1], xd[2] - b[3 * j + 2]};
                const R nv = wp * dot3(nrm, v);
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const R u[3] = {x[0] - a[3 * i], x[1] - a[3 * i + 1], x[2] - a[3 * i + 2]};
                    const R uv = wp * dot3(u, v);
                    zz[i + 3 * j].r += dot3c_r(u, Rs) * nv - uv * Ss.r;
                    zz[i + 3 * j].i += dot3c_i(u, Rs) * nv - uv * Ss.i;
                }
            }
        }
        const R s1 = (R)(1.0 / (4.0 * t_area() * sg[kGeoA]));
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            rot(zz[k]);
            z[k * stride] = {s1 * zz[k].r, s1 * zz[k].i};
        }
    } else {
        // M1 = sum w x.U, MU = sum w U, MX = sum w x x Rv, MR = sum w Rv with
        // U = xd x H1, Rv = xd H0 - H1, H0 = sum w h, H1 = sum w h y, h = -(gamma + 1/R) G / R
        cx<R> M1{0, 0}, MU[3] = {{0, 0}, {0, 0}, {0, 0}}, MX[3] = {{0, 0}, {0, 0}, {0, 0}},
                        MR[3] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll kUnrollO
        for (int ip = 0; ip < Qo; ++ip) {
            const R x[3] = {xpt(4 * ip), xpt(4 * ip + 1), xpt(4 * ip + 2)};
            const R wp = xpt(4 * ip + 3);
            const R xd[3] = {x[0] + dc[0], x[1] + dc[1], x[2] + dc[2]};
            cx<R> H0{0, 0}, H1[3] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll
            for (int iq = 0; iq < Qi; ++iq) {
                const R y[3] = {yq[4 * iq], yq[4 * iq + 1], yq[4 * iq + 2]};
                const R wq = yq[4 * iq + 3];
                const R r[3] = {xd[0] - y[0], xd[1] - y[1], xd[2] - y[2]};
                const R R2 = dot3(r, r);
                R dist, idist;
                if constexpr (sizeof(R) == 8) { idist = rsqrt_f64(R2); dist = R2 * idist; }
                else { idist = rsqrtf(R2); dist = R2 * idist; }
                const cx<R> G = kernel_g(dist, idist);
                cx<R> h = cmul(cx<R>{gr + idist, gi}, G);
                const R f = -idist * wq;
                h.r *= f; h.i *= f;
                H0.r += h.r; H0.i += h.i;
#pragma unroll
                for (int k = 0; k < 3; ++k) { H1[k].r += h.r * y[k]; H1[k].i += h.i * y[k]; }
            }
            cx<R> U[3], Rv[3];
            U[0] = {xd[1] * H1[2].r - xd[2] * H1[1].r, xd[1] * H1[2].i - xd[2] * H1[1].i};
            U[1] = {xd[2] * H1[0].r - xd[0] * H1[2].r, xd[2] * H1[0].i - xd[0] * H1[2].i};
            U[2] = {xd[0] * H1[1].r - xd[1] * H1[0].r, xd[0] * H1[1].i - xd[1] * H1[0].i};
#pragma unroll
            for (int k = 0; k < 3; ++k) Rv[k] = {xd[k] * H0.r - H1[k].r, xd[k] * H0.i - H1[k].i};
            M1.r += wp * (x[0] * U[0].r + x[1] * U[1].r + x[2] * U[2].r);
            M1.i += wp * (x[0] * U[0].i + x[1] * U[1].i + x[2] * U[2].i);
            const cx<R> XR[3] = {{x[1] * Rv[2].r - x[2] * Rv[1].r, x[1] * Rv[2].i - x[2] * Rv[1].i},
                                 {x[2] * Rv[0].r - x[0] * Rv[2].r, x[2] * Rv[0].i - x[0] * Rv[2].i},
                                 {x[0] * Rv[1].r - x[1] * Rv[0].r, x[0] * Rv[1].i - x[1] * Rv[0].i}};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                MU[k].r += wp * U[k].r; MU[k].i += wp * U[k].i;
                MX[k].r += wp * XR[k].r; MX[k].i += wp * XR[k].i;
                MR[k].r += wp * Rv[k].r; MR[k].i += wp * Rv[k].i;
            }
        }
        rot(M1);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            rot(MU[k]);
            rot(MX[k]);
            rot(MR[k]);
        }
        R a[9], b[9];
        load_verts(a, b);
        const R s1 = (R)(1.0 / (4.0 * t_area() * sg[kGeoA]));
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const R *ai = a + 3 * i, *bj = b + 3 * j;
                R c[3];
                cross3(c, bj, ai);
                cx<R> t;
                t.r = M1.r - (ai[0] * MU[0].r + ai[1] * MU[1].r + ai[2] * MU[2].r) -
                      (bj[0] * MX[0].r + bj[1] * MX[1].r + bj[2] * MX[2].r) +
                      (c[0] * MR[0].r + c[1] * MR[1].r + c[2] * MR[2].r);
                t.i = M1.i - (ai[0] * MU[0].i + ai[1] * MU[1].i + ai[2] * MU[2].i) -
                      (bj[0] * MX[0].i + bj[1] * MX[1].i + bj[2] * MX[2].i) +
                      (c[0] * MR[0].i + c[1] * MR[1].i + c[2] * MR[2].i);
                z[(i + 3 * j) * stride] = {s1 * t.r, s1 * t.i};
            }
    }
}

constexpr int kTile = 64;          // dofs per tile side
constexpr int kPairChunk = 512;    // pair blocks staged in LDS at once (two per thread)
constexpr int kPairChunkStaged = 256;   // the same when the test cells are staged too (one per thread)

struct TileSide {
    const int *ptr;       // [ntiles+1] into cells
    const int *cells;     // global cell ids
    const int *slot;      // [ndofs][2]: (local cell << 2 | refid), -1 none
    const double *coef;   // [ndofs][2]
    int ndofs;
};

struct TileArgs {
    FarArgs far;
    TileSide row, col;
    void *Z;              // output, complex R
    long long ld;
    void *Z32;            // optional row-major complex-float copy (FP64 row-major launches)
    long long ld32;
    int rowmajor;
    int accumulate;
    double scale_r, scale_i;
    const void *rpack;    // the row tiles' test cells packed field by field (single layer, default rule)
};

// v of lane l (l wave-uniform)
__device__ __forceinline__ double readlane_r(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ float readlane_r(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// LDS of a far tile: the pair blocks of a chunk and, for the double layer's small default
// rule, the tile's test cells (whose per-lane reads would otherwise touch one cache line
// per lane), later the row-major transpose, in one buffer within 80 KB (two workgroups
// per CU).  Staging halves the chunk, and so doubles the chunk barriers; measured on the
// 26,664-cell sphere it pays for the double layer (78.4 -> 75.7 ms) and not for the
// single layer (66.2 -> 68.5 ms), whose test cells are read through the cache.
template <typename R, int OPK, int QO, bool ROWMAJOR>
struct FarTileLds {
    static constexpr bool kStage = OPK == BH_OP_MW_DOUBLELAYER && QO > 0 && QO <= 4;
    // the single layer reads the same field-major layout from a packed copy in global memory
    // (A.rpack: consecutive lanes read consecutive words), keeping its 512-pair chunks
    static constexpr bool kPacked = OPK == BH_OP_MW_SINGLELAYER && QO > 0 && QO <= 4;
    static constexpr size_t kRowBytes = (size_t)kTileCells * (4 * sizeof(double) + (9 + 4 * (QO > 0 ? QO : 1)) * sizeof(R));
    static constexpr int kChunk = kStage ? kPairChunkStaged : kPairChunk;
    static constexpr size_t kBlk = (size_t)kChunk * 9 * sizeof(cx<R>);
    static constexpr size_t kStageBytes = kStage ? kRowBytes : 0;
    static constexpr size_t kT = ROWMAJOR ? (size_t)kTile * (kTile + 1) * sizeof(cx<R>) : 0;
    static constexpr size_t kBytes = kBlk + kStageBytes > kT ? kBlk + kStageBytes : kT;
};

template <typename R, int OPK, bool ROWMAJOR, int QO, int QI, bool SHIFT = false>
__device__ void far_tiles_dev(const TileArgs &A, int tr, int tc)
{
    using L = FarTileLds<R, OPK, QO, ROWMAJOR>;
    constexpr bool kStage = L::kStage;
    constexpr int kChunk = L::kChunk;
    __shared__ int s_rc[kTileCells], s_cc[kTileCells];
    __shared__ int s_rslot[kTile][2], s_cslot[kTile][2];
    __shared__ R s_rcoef[kTile][2], s_ccoef[kTile][2];
    __shared__ __align__(16) unsigned char s_raw[L::kBytes];
    __shared__ unsigned s_near[kTileCells * kTileCells / 32];   // bit rc * kTileCells + cc: non-far pair
    cx<R> *s_blk = reinterpret_cast<cx<R> *>(s_raw);
    double *s_cen = reinterpret_cast<double *>(s_raw + L::kBlk);   // [3][kTileCells], then area
    double *s_area = s_cen + 3 * kTileCells;
    R *s_av = reinterpret_cast<R *>(s_area + kTileCells);        // [9][kTileCells], then points
    R *s_pts = s_av + 9 * kTileCells;
    const int tid = threadIdx.x;
    const int r0 = A.row.ptr[tr], nr = A.row.ptr[tr + 1] - r0;
    const int c0 = A.col.ptr[tc], nc = A.col.ptr[tc + 1] - c0;
    const int mbase = tr * kTile, nbase = tc * kTile;
    for (int k = tid; k < nr; k += 256) {
        const int p = A.row.cells[r0 + k];
        s_rc[k] = p;
        if constexpr (kStage) {
            const double *tg = A.far.tgeo + (size_t)p * kGeo;
#pragma unroll
            for (int c = 0; c < 3; ++c) s_cen[c * kTileCells + k] = tg[kGeoC + c];
            s_area[k] = tg[kGeoA];
#pragma unroll
            for (int v = 0; v < 9; ++v) s_av[v * kTileCells + k] = (R)(tg[v] - tg[kGeoC + v % 3]);
            const R *tq = (const R *)A.far.tq + (size_t)p * QO * 4;
#pragma unroll
            for (int v = 0; v < 4 * QO; ++v) s_pts[v * kTileCells + k] = tq[v];
        }
    }
    for (int k = tid; k < nc; k += 256) s_cc[k] = A.col.cells[c0 + k];
    if (tid < 2 * kTile) {
        const int l = tid >> 1, s = tid & 1;
        const int m = mbase + l, n = nbase + l;
        s_rslot[l][s] = m < A.row.ndofs ? A.row.slot[2 * m + s] : -1;
        s_rcoef[l][s] = m < A.row.ndofs ? (R)A.row.coef[2 * m + s] : (R)0;
        s_cslot[l][s] = n < A.col.ndofs ? A.col.slot[2 * n + s] : -1;
        s_ccoef[l][s] = n < A.col.ndofs ? (R)A.col.coef[2 * n + s] : (R)0;
    }
    for (int k = tid; k < kTileCells * kTileCells / 32; k += 256) s_near[k] = 0u;
    // ownership: a lane owns one row dof and 16 column dofs (its wave's), so that in the
    // gather a column dof's cell lies in the current chunk or not for the whole wave
    const int own = tid & 63, grp = __builtin_amdgcn_readfirstlane(tid >> 6);
    cx<R> acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = {(R)0, (R)0};
    __syncthreads();
    // mark the tile's non-far pairs: each row cell's non-far list (ascending trial
    // cells) looked up in the sorted column cells, so that far pairs skip the
    // search of far_block
    const bool lists = A.far.mode != kModeDoubleNum;
    if (lists) {
        for (int rc = tid; rc < nr; rc += 256) {
            const int p = s_rc[rc];
            const int e1 = A.far.nptr[p + 1];
            for (int e = A.far.nptr[p]; e < e1; ++e) {
                const int q = A.far.nent[e].y;
                int lo = 0, hi = nc;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (s_cc[mid] < q) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < nc && s_cc[lo] == q) {
                    const int bit = rc * kTileCells + lo;
                    atomicOr(&s_near[bit >> 5], 1u << (bit & 31));
                }
            }
        }
        __syncthreads();
    }

    const int CC = nr > 0 ? max(1, kChunk / nr) : 1;
    for (int cb = 0; cb < nc; cb += CC) {
        const int cn = min(CC, nc - cb);
        // phase 1: pair blocks of (all row cells) x (column cells cb .. cb+cn)
#pragma unroll 1
        for (int t = tid; t < nr * cn; t += 256) {
            const int rc = t % nr, cc = t / nr;
            const int bit = rc * kTileCells + cb + cc;
            const bool near = lists && ((s_near[bit >> 5] >> (bit & 31)) & 1u);
            if constexpr (L::kPacked) {
                const unsigned char *pk = (const unsigned char *)A.rpack + (size_t)tr * L::kRowBytes;
                const double *pc = (const double *)pk;
                const R *pv = (const R *)(pc + 4 * kTileCells);
                far_block<R, OPK, QO, QI, SHIFT, true>(A.far, s_rc[rc], s_cc[cb + cc], s_blk + t, kChunk, near,
                                                       RowStage<R>{pc, pv, pc + 3 * kTileCells, pv + 9 * kTileCells, rc});
            } else {
                far_block<R, OPK, QO, QI, SHIFT, kStage>(A.far, s_rc[rc], s_cc[cb + cc], s_blk + t, kChunk, near,
                                                         RowStage<R>{s_cen, s_av, s_area, s_pts, rc});
            }
        }
        __syncthreads();
        // phase 2: signed gather, store(a * (z * b), m, n) order of integralop.jl:173-184.
        // The wave's 32 column slots (its 16 column dofs, two cells each) are read at once,
        // lane l < 32 holding slot l, and the slots whose cell lies in this chunk form a
        // wave-uniform mask, so no slot waits on its own LDS read
        {
            const int sl = own & 31, nls = grp * 16 + (sl >> 1);
            const int cs_l = s_cslot[nls][sl & 1];
            const R bc_l = s_ccoef[nls][sl & 1];
            const int lc_l = (cs_l >> 2) - cb;
            const unsigned hit = (unsigned)__ballot(cs_l >= 0 && lc_l >= 0 && lc_l < cn);   // lanes 0..31
            const int rs[2] = {s_rslot[own][0], s_rslot[own][1]};
            const R ac[2] = {s_rcoef[own][0], s_rcoef[own][1]};
#pragma unroll
            for (int e = 0; e < 16; ++e) {
#pragma unroll
                for (int sb = 0; sb < 2; ++sb) {
                    const int slot = 2 * e + sb;
                    if (!((hit >> slot) & 1u)) continue;
                    const int cslt = __builtin_amdgcn_readlane(cs_l, slot);
                    const R bc = readlane_r(bc_l, slot);
                    const int lc = (cslt >> 2) - cb, j = cslt & 3;
#pragma unroll
                    for (int sa = 0; sa < 2; ++sa) {
                        if (rs[sa] < 0) continue;
                        const cx<R> zz = s_blk[(lc * nr + (rs[sa] >> 2)) + ((rs[sa] & 3) + 3 * j) * kChunk];
                        acc[e].r += ac[sa] * (zz.r * bc);
                        acc[e].i += ac[sa] * (zz.i * bc);
                    }
                }
            }
        }
        __syncthreads();
    }
    // write once; a row-major Z is written along rows through an LDS transpose (the pair
    // blocks' space, free after the last chunk's barrier)
    cx<R> *Z = (cx<R> *)A.Z;
    const R sr = (R)A.scale_r, si = (R)A.scale_i;
    if constexpr (ROWMAJOR) {
        static_assert(kTile * (kTile + 1) * sizeof(cx<R>) <= L::kBytes, "transpose does not fit the LDS buffer");
        cx<R> *s_t = s_blk;
#pragma unroll
        for (int e = 0; e < 16; ++e) s_t[own * (kTile + 1) + grp * 16 + e] = acc[e];
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[e] = s_t[(grp * 16 + e) * (kTile + 1) + own];
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        int m, n;
        if constexpr (ROWMAJOR) { m = mbase + grp * 16 + e; n = nbase + own; }
        else { m = mbase + own; n = nbase + grp * 16 + e; }
        if (m >= A.row.ndofs || n >= A.col.ndofs) continue;
        const size_t idx = ROWMAJOR ? (size_t)m * A.ld + n : (size_t)m + (size_t)n * A.ld;
        cx<R> v{sr * acc[e].r - si * acc[e].i, sr * acc[e].i + si * acc[e].r};
        if (A.accumulate) { v.r += Z[idx].r; v.i += Z[idx].i; }
        Z[idx] = v;
        if constexpr (ROWMAJOR && sizeof(R) == 8) {
            if (A.Z32) ((float2 *)A.Z32)[(size_t)m * A.ld32 + n] = float2{(float)v.r, (float)v.i};
        }
    }
}

// The test cells of every row tile, packed in the layout of RowStage: per tile, centroid
// [3][kTileCells], area [kTileCells] (doubles), then vertices relative to the centroid
// [9][kTileCells] and points [4 Qo][kTileCells] in R.  Block = one row tile.
template <typename R>
__global__ __launch_bounds__(kTileCells) void k_pack_rows(const int *__restrict__ ptr, const int *__restrict__ cells,
                                                          const double *__restrict__ geo, const R *__restrict__ tq,
                                                          int Qo, unsigned char *pack, size_t tile_bytes)
{
    const int tile = blockIdx.x, k = threadIdx.x;
    const int r0 = ptr[tile], nr = ptr[tile + 1] - r0;
    if (k >= nr) return;
    const int p = cells[r0 + k];
    double *pc = (double *)(pack + (size_t)tile * tile_bytes);
    R *pv = (R *)(pc + 4 * kTileCells);
    const double *g = geo + (size_t)p * kGeo;
    for (int c = 0; c < 3; ++c) pc[c * kTileCells + k] = g[kGeoC + c];
    pc[3 * kTileCells + k] = g[kGeoA];
    for (int v = 0; v < 9; ++v) pv[v * kTileCells + k] = (R)(g[v] - g[kGeoC + v % 3]);
    for (int v = 0; v < 4 * Qo; ++v) pv[(9 + v) * kTileCells + k] = tq[(size_t)p * Qo * 4 + v];
}

// workgroups per CU the compiler must fit: two for n x K (see far_block), free otherwise
template <int OPK>
constexpr int far_min_blocks() { return OPK == BH_OP_MW_DOUBLELAYER_ROTATED ? 2 : 1; }

template <typename R, int OPK, bool ROWMAJOR, int QO, int QI, bool SHIFT = false>
__global__ __launch_bounds__(256, far_min_blocks<OPK>()) void k_far_tiles(TileArgs A)
{
    far_tiles_dev<R, OPK, ROWMAJOR, QO, QI, SHIFT>(A, blockIdx.y, blockIdx.x);
}

// Far blocks for an explicit pair list (inspection entry).
template <typename R, int OPK>
__global__ void k_far_pairs(FarArgs A, const int *tc, const int *sc, int n, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    cx<R> z[9];
    far_block<R, OPK>(A, tc[k], sc[k], z, 1);
    for (int t = 0; t < 9; ++t) { out[18 * (size_t)k + 2 * t] = (double)z[t].r; out[18 * (size_t)k + 2 * t + 1] = (double)z[t].i; }
}

// ============================================================ host side
namespace {

struct Strategy {
    int mode;
    int far_o, far_i, near_o, near_i;
    int ss_face, ss_edge, ss_vert;
};

// Triangle-rule degrees and Gauss-Legendre orders: below 1 is an invalid argument; above
// the library's tables a valid strategy the GPU path does not cover (the caller falls back).
void check_degree(int d, const char *what)
{
    if (d < 1) BH_FAIL(BH_ERR_INVALID, "%s: triangle rule degree %d must be at least 1", what, d);
    if (d > kMaxTriDegree)
        BH_FAIL(BH_ERR_UNSUPPORTED, "%s: no triangle rule of degree %d on the GPU path (1..%d)", what, d,
                kMaxTriDegree);
}

void check_order(int n, const char *what)
{
    if (n < 1) BH_FAIL(BH_ERR_INVALID, "%s: Sauter-Schwab order %d must be at least 1", what, n);
    if (n > kMaxGL)
        BH_FAIL(BH_ERR_UNSUPPORTED, "%s: Sauter-Schwab order %d above the GPU path's %d", what, n, kMaxGL);
}

Strategy parse_qs(const bh_quadstrat *qs)
{
    BH_REQUIRE(qs != nullptr, "quadstrat is NULL");
    Strategy s{};
    const int *p = qs->params;
    switch (qs->kind) {
    case BH_QS_WILTON_SAUTER:
        s = {kModeWiltonSauter, p[0], p[1], p[2], p[3], p[5], p[6], p[7]};
        for (int k = 0; k < 4; ++k) check_degree(p[k], "DoubleNumWiltonSauterQStrat");
        for (int k = 5; k < 8; ++k) check_order(p[k], "DoubleNumWiltonSauterQStrat");
        break;
    case BH_QS_SAUTER:
        s = {kModeSauter, p[0], p[1], 0, 0, p[3], p[4], p[5]};
        for (int k = 0; k < 2; ++k) check_degree(p[k], "DoubleNumSauterQstrat");
        for (int k = 3; k < 6; ++k) check_order(p[k], "DoubleNumSauterQstrat");
        break;
    case BH_QS_DOUBLENUM:
        s = {kModeDoubleNum, p[0], p[1], 0, 0, 0, 0, 0};
        for (int k = 0; k < 2; ++k) check_degree(p[k], "DoubleNumQStrat");
        break;
    case BH_QS_SINGLENUM:
        BH_FAIL(BH_ERR_UNSUPPORTED, "SingleNumQStrat applies to local operators only");
    default:
        BH_FAIL(BH_ERR_UNSUPPORTED, "quadrature strategy kind %d is not supported on the GPU path", qs->kind);
    }
    return s;
}

OpParams parse_op(const bh_operator *op)
{
    BH_REQUIRE(op != nullptr, "operator is NULL");
    BH_REQUIRE(op->kind == BH_OP_MW_SINGLELAYER || op->kind == BH_OP_MW_DOUBLELAYER || op->kind == BH_OP_NCROSS ||
                   op->kind == BH_OP_MW_DOUBLELAYER_ROTATED || op->kind == BH_OP_IDENTITY,
               "operator kind %d is not supported", op->kind);
    OpParams p{op->kind, op->gamma[0], op->gamma[1], op->alpha[0], op->alpha[1], op->beta[0], op->beta[1]};
    for (double v : {p.gr, p.gi, p.ar, p.ai, p.br, p.bi})
        BH_REQUIRE(std::isfinite(v), "operator parameters must be finite");
    return p;
}

}  // namespace

// Quadrature points of a rule on every cell of a mesh (cached per degree):
// [nf][Q][4] = x - centroid (3), weight * 2 * area.  Points follow
// quadpoints(refspace, charts, rule): x = u v1 + v v2 + (1 - u - v) v3.
const void *quad_points(Mesh &m, int degree, int *Q, bool fp32, hipStream_t s)
{
    double u[kMaxTriPts], v[kMaxTriPts], w[kMaxTriPts];
    if (degree > kMaxTriDegree)
        BH_FAIL(BH_ERR_UNSUPPORTED, "no triangle rule of degree %d on the GPU path (1..%d)", degree, kMaxTriDegree);
    const int nq = triangle_rule(degree, u, v, w);
    BH_REQUIRE(nq > 0, "no triangle rule of degree %d", degree);
    *Q = nq;
    auto &cache64 = m.quad64[degree];
    if (cache64.n == 0 && m.nf > 0) {
        std::vector<double> h((size_t)m.nf * nq * 4);
        for (int64_t c = 0; c < m.nf; ++c) {
            const double *g = &m.geo[(size_t)c * kGeo];
            for (int q = 0; q < nq; ++q) {
                double *o = &h[((size_t)c * nq + q) * 4];
                for (int k = 0; k < 3; ++k) {
                    const double x = u[q] * g[k] + v[q] * g[3 + k] + (1.0 - u[q] - v[q]) * g[6 + k];
                    o[k] = x - g[kGeoC + k];
                }
                o[3] = w[q] * 2.0 * g[kGeoA];
            }
        }
        cache64.upload(h.data(), h.size(), s);
        std::vector<float> hf(h.begin(), h.end());
        m.quad32[degree].upload(hf.data(), hf.size(), s);
        BH_CHECK_HIP(hipStreamSynchronize(s));
    }
    return fp32 ? (const void *)m.quad32[degree].p : (const void *)cache64.p;
}

namespace {

struct NearData {
    DevBuf<int> nptr;
    DevBuf<int4> entries;
    DevBuf<double> blocks;
    int nent = 0;
    int count[5] = {0, 0, 0, 0, 0};
};

void check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) BH_FAIL(BH_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
}

// Trial-cell grid for the restricted classification (see k_classify_count_grid);
// empty (use == false) when the meshes are too small for it to pay.
struct PairGrid {
    bool use = false;
    DevBuf<int> start, cells;
    PairGridView view{};
};

void build_pair_grid(const Mesh &tm, const Mesh &sm, int mode, hipStream_t s, PairGrid &pg)
{
    const int nft = (int)tm.nf, nfs = (int)sm.nf;
    if ((long long)nft * nfs < (1LL << 22)) return;
    auto rho_of = [](const Mesh &m) {
        double r = 0;
        for (int64_t f = 0; f < m.nf; ++f) {
            const double *g = m.geo.data() + (size_t)f * kGeo;
            for (int i = 0; i < 3; ++i) {
                const double dx = g[3 * i] - g[kGeoC], dy = g[3 * i + 1] - g[kGeoC + 1], dz = g[3 * i + 2] - g[kGeoC + 2];
                r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz));
            }
        }
        return r;
    };
    double amax = 0, lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int q = 0; q < nfs; ++q) {
        const double *g = sm.geo.data() + (size_t)q * kGeo;
        amax = std::max(amax, g[kGeoA]);
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], g[kGeoC + d]);
            hi[d] = std::max(hi[d], g[kGeoC + d]);
        }
    }
    const double dtol = 1.0e3 * 2.220446049250313e-16;
    const double reach = std::max(mode == kModeWiltonSauter ? std::sqrt(0.64 * amax) : 0.0, dtol);
    double h = (reach + rho_of(tm) + rho_of(sm)) * (1.0 + 1e-6);
    if (!(h > 0) || !std::isfinite(h)) return;
    const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
    h = std::max(h, ext / 128.0);   // at most 128^3 grid cells
    int nd[3];
    for (int d = 0; d < 3; ++d) nd[d] = std::max(1, (int)std::floor((hi[d] - lo[d]) / h) + 1);
    const long long ncell = (long long)nd[0] * nd[1] * nd[2];
    std::vector<int> start(ncell + 1, 0), cells(nfs), key(nfs);
    const double inv_h = 1.0 / h;
    for (int q = 0; q < nfs; ++q) {
        const double *c = sm.geo.data() + (size_t)q * kGeo + kGeoC;
        int ix[3];
        for (int d = 0; d < 3; ++d) ix[d] = std::min(std::max((int)std::floor((c[d] - lo[d]) * inv_h), 0), nd[d] - 1);
        key[q] = (ix[2] * nd[1] + ix[1]) * nd[0] + ix[0];
        ++start[key[q] + 1];
    }
    for (long long i = 0; i < ncell; ++i) start[i + 1] += start[i];
    std::vector<int> fillp(start.begin(), start.end() - 1);
    for (int q = 0; q < nfs; ++q) cells[fillp[key[q]]++] = q;   // ascending q within a grid cell
    pg.start.alloc(ncell + 1);
    pg.cells.alloc(nfs);
    pg.start.upload(start.data(), start.size(), s);
    pg.cells.upload(cells.data(), cells.size(), s);
    pg.view = PairGridView{pg.start.p, pg.cells.p, nd[0], nd[1], nd[2], lo[0], lo[1], lo[2], inv_h};
    pg.use = true;
}

// K0 + K2 + K3: classification, compaction and the non-far blocks.
void build_near(Ctx &ctx, Mesh &tm, Mesh &sm, const Strategy &st, const OpParams &op, NearData &nd)
{
    hipStream_t s = ctx.stream;
    const int nft = (int)tm.nf, nfs = (int)sm.nf;
    nd.nptr.alloc(nft + 1);
    std::vector<int> off(nft + 1, 0);
    if (st.mode == kModeDoubleNum || nft == 0 || nfs == 0) {
        nd.nptr.upload(off.data(), off.size(), s);
        nd.nent = 0;
        return;
    }
    const double k2 = op.gr * op.gr + op.gi * op.gi;
    DevBuf<int> counts(nft), err(1);
    BH_CHECK_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
    PairGrid pg;
    build_pair_grid(tm, sm, st.mode, s, pg);
    ProfScope prof_c(ctx, "k_classify_count", (double)nft * nfs);
    if (pg.use)
        k_classify_count_grid<<<nft, 256, 0, s>>>(tm.d_geo.p, nft, sm.d_geo.p, pg.view, k2, st.mode, counts.p, err.p);
    else
        k_classify_count<<<nft, 256, 0, s>>>(tm.d_geo.p, nft, sm.d_geo.p, nfs, k2, st.mode, counts.p, err.p);
    check_launch("k_classify_count");
    std::vector<int> hc(nft);
    int herr = 0;
    BH_CHECK_HIP(hipMemcpyAsync(hc.data(), counts.p, sizeof(int) * nft, hipMemcpyDeviceToHost, s));
    BH_CHECK_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    BH_CHECK_HIP(hipStreamSynchronize(s));
    if (herr) BH_FAIL(BH_ERR_ASSERT, "AssertionError: hits <= 3 (two cells share more than 3 vertex positions)");
    int hmax = 0;
    for (int p = 0; p < nft; ++p) {
        off[p + 1] = off[p] + hc[p];
        hmax = std::max(hmax, hc[p]);
    }
    nd.nent = off[nft];
    nd.nptr.upload(off.data(), off.size(), s);
    if (nd.nent == 0) return;
    nd.entries.alloc(nd.nent);
    if (pg.use && hmax <= kGridFillCap)
        k_classify_fill_grid<<<nft, 256, 0, s>>>(tm.d_geo.p, nft, sm.d_geo.p, pg.view, k2, st.mode, nd.nptr.p,
                                                  nd.entries.p);
    else
        k_classify_fill<<<nft, 256, 0, s>>>(tm.d_geo.p, nft, sm.d_geo.p, nfs, k2, st.mode, nd.nptr.p, nd.entries.p);
    check_launch("k_classify_fill");
    DevBuf<int> counters(5), lists((size_t)5 * nd.nent);
    BH_CHECK_HIP(hipMemsetAsync(counters.p, 0, 5 * sizeof(int), s));
    k_class_lists<<<(nd.nent + 255) / 256, 256, 0, s>>>(nd.entries.p, nd.nent, counters.p, lists.p, nd.nent);
    check_launch("k_class_lists");
    BH_CHECK_HIP(hipMemcpyAsync(nd.count, counters.p, 5 * sizeof(int), hipMemcpyDeviceToHost, s));
    BH_CHECK_HIP(hipStreamSynchronize(s));
    nd.blocks.alloc((size_t)nd.nent * 18);
    // Sauter-Schwab, one launch per class
    const int cls_list[3] = {kCF, kCE, kCV};
    for (int ci = 0; ci < 3; ++ci) {
        const int cls = cls_list[ci];
        const int cnt = nd.count[cls];
        if (cnt == 0) continue;
        GLRule gl{};
        gl.n = cls == kCF ? st.ss_face : (cls == kCE ? st.ss_edge : st.ss_vert);
        legendre01(gl.n, gl.x, gl.w);
        const int grid = (cnt + 3) / 4;
        const int *lp = lists.p + (size_t)cls * nd.nent;
        const int nreg = cls == kCF ? 6 : (cls == kCE ? 5 : 2);
        ProfScope prof_s(ctx, "k_sauter_schwab", (double)cnt * nreg * gl.n * gl.n * gl.n * gl.n);
        if (op.kind == BH_OP_MW_SINGLELAYER)
            k_sauter_schwab<BH_OP_MW_SINGLELAYER><<<grid, 256, 0, s>>>(tm.d_geo.p, sm.d_geo.p, nd.entries.p, lp, cnt,
                                                                       cls, gl, op, nd.blocks.p);
        else if (op.kind == BH_OP_MW_DOUBLELAYER_ROTATED)
            k_sauter_schwab<BH_OP_MW_DOUBLELAYER_ROTATED><<<grid, 256, 0, s>>>(tm.d_geo.p, sm.d_geo.p, nd.entries.p,
                                                                               lp, cnt, cls, gl, op, nd.blocks.p);
        else
            k_sauter_schwab<BH_OP_MW_DOUBLELAYER><<<grid, 256, 0, s>>>(tm.d_geo.p, sm.d_geo.p, nd.entries.p, lp, cnt,
                                                                       cls, gl, op, nd.blocks.p);
        check_launch("k_sauter_schwab");
    }
    const int cnt = nd.count[kNear];
    if (cnt > 0) {
        int Qo = 0, Qi = 0;
        const double *tq = (const double *)quad_points(tm, st.near_o, &Qo, false, s);
        const double *sq = (const double *)quad_points(sm, st.near_i, &Qi, false, s);
        BH_REQUIRE(Qo <= 64, "near outer rule too large");
        const int grid = (cnt + 3) / 4;
        const int *lp = lists.p + (size_t)kNear * nd.nent;
        ProfScope prof_w(ctx, "k_wilton", (double)cnt * Qo * Qi);
        if (op.kind == BH_OP_MW_SINGLELAYER)
            k_wilton<BH_OP_MW_SINGLELAYER><<<grid, 256, 0, s>>>(tm.d_geo.p, sm.d_geo.p, nd.entries.p, lp, cnt, tq, Qo,
                                                                sq, Qi, op, nd.blocks.p);
        else
            k_wilton<BH_OP_MW_DOUBLELAYER><<<grid, 256, 0, s>>>(tm.d_geo.p, sm.d_geo.p, nd.entries.p, lp, cnt, tq, Qo,
                                                                sq, Qi, op, nd.blocks.p);
        check_launch("k_wilton");
    }
    BH_CHECK_HIP(hipStreamSynchronize(s));  // lists/counters go out of scope
}

// Dof tiles of a space: for each group of kTile consecutive dofs, the sorted
// cells their shapes live on, and per dof up to two (local cell, refid, coeff).
void build_tiles(Space &sp, hipStream_t s)
{
    if (sp.tiles_built) return;
    const int64_t n = sp.ndofs;
    const int64_t nt = (n + kTile - 1) / kTile;
    std::vector<int> ptr(nt + 1, 0), cells, slot(2 * n, -1);
    std::vector<double> coef(2 * n, 0.0);
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t lo = t * kTile, hi = std::min(n, lo + kTile);
        std::vector<int> tc;
        for (int64_t m = lo; m < hi; ++m) {
            const int ns = sp.ptr[m + 1] - sp.ptr[m];
            if (ns > 2) BH_FAIL(BH_ERR_UNSUPPORTED, "dof %lld has %d shapes; the GPU path supports at most 2",
                                (long long)m, ns);
            for (int k = sp.ptr[m]; k < sp.ptr[m + 1]; ++k) tc.push_back(sp.cell[k]);
        }
        std::sort(tc.begin(), tc.end());
        tc.erase(std::unique(tc.begin(), tc.end()), tc.end());
        if ((int)tc.size() > kTileCells) BH_FAIL(BH_ERR_UNSUPPORTED, "dof tile touches %d cells (max %d)",
                                                 (int)tc.size(), kTileCells);
        for (int64_t m = lo; m < hi; ++m) {
            int s2 = 0;
            for (int k = sp.ptr[m]; k < sp.ptr[m + 1]; ++k) {
                const int lc = (int)(std::lower_bound(tc.begin(), tc.end(), sp.cell[k]) - tc.begin());
                slot[2 * m + s2] = (lc << 2) | sp.ref[k];
                coef[2 * m + s2] = sp.coeff[k];
                ++s2;
            }
        }
        cells.insert(cells.end(), tc.begin(), tc.end());
        ptr[t + 1] = (int)cells.size();
    }
    sp.t_ptr.upload(ptr.data(), ptr.size(), s);
    sp.t_cells.upload(cells.data(), cells.size(), s);
    sp.t_slot.upload(slot.data(), slot.size(), s);
    sp.t_coef.upload(coef.data(), coef.size(), s);
    sp.ntiles = nt;
    BH_CHECK_HIP(hipStreamSynchronize(s));
    sp.tiles_built = true;
}

struct Plan {
    Strategy st;
    OpParams op;
    NearData nd;
    FarArgs far;
    long long nft = 0, nfs = 0;
};

// Lossless FP64: phases relative to the centroid distance when |gi| (r_p + r_q) stays below
// 0.369, where the degree-12/13 Taylor remainders of cos and sin are below 1e-17 (r: the
// largest vertex-to-centroid distance of each mesh).
bool shifted_phases(const OpParams &op, const Mesh &tm, const Mesh &sm, bool fp32)
{
    if (fp32 || op.gr != 0.0) return false;
    auto radius = [](const Mesh &m) {
        double r = 0.0;
        for (int64_t c = 0; c < m.nf; ++c) {
            const double *g = &m.geo[(size_t)c * kGeo];
            for (int a = 0; a < 3; ++a) {
                const double d0 = g[3 * a] - g[kGeoC], d1 = g[3 * a + 1] - g[kGeoC + 1], d2 = g[3 * a + 2] - g[kGeoC + 2];
                r = std::max(r, std::sqrt(d0 * d0 + d1 * d1 + d2 * d2));
            }
        }
        return r;
    };
    const double th = std::fabs(op.gi) * (radius(tm) + radius(sm)) * (1.0 + 1e-12);
    return th <= 0.369;
}

void make_plan(Ctx &ctx, const bh_operator *op, const bh_quadstrat *qs, Mesh &tm, Mesh &sm, bool fp32, Plan &P)
{
    P.st = parse_qs(qs);
    P.op = parse_op(op);
    if (P.op.kind == BH_OP_MW_DOUBLELAYER_ROTATED && P.st.mode == kModeWiltonSauter)
        BH_FAIL(BH_ERR_UNSUPPORTED, "DoubleLayerRotatedMW3D has no singularity extraction: "
                                    "DoubleNumWiltonSauterQStrat does not apply (use DoubleNumQStrat or "
                                    "DoubleNumSauterQstrat)");
    BH_REQUIRE(P.op.kind != BH_OP_NCROSS && P.op.kind != BH_OP_IDENTITY, "local operators have no pair plan");
    P.nft = tm.nf;
    P.nfs = sm.nf;
    build_near(ctx, tm, sm, P.st, P.op, P.nd);
    FarArgs &F = P.far;
    F.tgeo = tm.d_geo.p;
    F.sgeo = sm.d_geo.p;
    F.tq = quad_points(tm, P.st.far_o, &F.Qo, fp32, ctx.stream);
    F.sq = quad_points(sm, P.st.far_i, &F.Qi, fp32, ctx.stream);
    F.mode = P.st.mode;
    F.k2 = P.op.gr * P.op.gr + P.op.gi * P.op.gi;
    F.shifted = shifted_phases(P.op, tm, sm, fp32) ? 1 : 0;
    F.nptr = P.nd.nptr.p;
    F.nent = P.nd.entries.p;
    F.nblk = P.nd.blocks.p;
    F.op = P.op;
}

// The row tiles' test cells packed for the single layer's default-rule tiles (k_pack_rows).
template <typename R>
const void *pack_rows(hipStream_t s, const Space &X, const FarArgs &F, DevBuf<char> &buf)
{
    const size_t tb = (size_t)kTileCells * (4 * sizeof(double) + (9 + 4 * F.Qo) * sizeof(R));
    buf.ensure(tb * (size_t)X.ntiles);
    k_pack_rows<R><<<(unsigned)X.ntiles, kTileCells, 0, s>>>(X.t_ptr.p, X.t_cells.p, F.tgeo, (const R *)F.tq, F.Qo,
                                                             (unsigned char *)buf.p, tb);
    check_launch("k_pack_rows");
    return buf.p;
}

template <typename R>
void launch_tiles(Ctx &ctx, const Plan &P, Space &X, Space &Y, void *Z, long long ld, bool rowmajor, bool accumulate,
                  double sr, double si, DevBuf<char> &rpack, void *Z32 = nullptr, long long ld32 = 0)
{
    TileArgs A{};
    A.far = P.far;
    A.Z32 = Z32;
    A.ld32 = ld32;
    A.row = {X.t_ptr.p, X.t_cells.p, X.t_slot.p, X.t_coef.p, (int)X.ndofs};
    A.col = {Y.t_ptr.p, Y.t_cells.p, Y.t_slot.p, Y.t_coef.p, (int)Y.ndofs};
    A.Z = Z;
    A.ld = ld;
    A.rowmajor = rowmajor;
    A.accumulate = accumulate;
    A.scale_r = sr;
    A.scale_i = si;
    BH_REQUIRE(X.ntiles <= 65535 && Y.ntiles <= 65535, "too many dof tiles for one launch");
    dim3 grid((unsigned)Y.ntiles, (unsigned)X.ntiles);
    const bool sl = P.op.kind == BH_OP_MW_SINGLELAYER;
    // algorithmic work: far point pairs x flops per pair of the reference formulation
    // (EFIE 170, MFIE-K 210; mwops.jl:101-117,141-155 with doublenumints.jl:32-38)
    const double far_pp = ((double)P.nft * P.nfs - P.nd.nent) * P.far.Qo * P.far.Qi;
    ProfScope prof(ctx, "k_far_tiles", far_pp * (sl ? 170.0 : 210.0));
    // the default far rules (Dunavant degrees 2 and 3: 3 x 4 points) get fully
    // unrolled point loops
    const bool dflt = P.far.Qo == 3 && P.far.Qi == 4;
    // the centroid-shifted phases (FP64, lossless, small cells) for the default rules
    const bool shift = sizeof(R) == 8 && dflt && P.far.shifted;
    if (sl && dflt) A.rpack = pack_rows<R>(ctx.stream, X, P.far, rpack);
#define BH_TILES(OPK, RM)                                                                    \
    do {                                                                                     \
        if (shift) k_far_tiles<R, OPK, RM, 3, 4, true><<<grid, 256, 0, ctx.stream>>>(A);    \
        else if (dflt) k_far_tiles<R, OPK, RM, 3, 4><<<grid, 256, 0, ctx.stream>>>(A);      \
        else k_far_tiles<R, OPK, RM, 0, 0><<<grid, 256, 0, ctx.stream>>>(A);                \
    } while (0)
    if (P.op.kind == BH_OP_MW_DOUBLELAYER_ROTATED) {
        // its default rules (Dunavant degrees 6 and 7: 12 x 13 points) get unrolled inner
        // point loops, with centroid-shifted phases where they apply; other rules the
        // generic loops
        constexpr int DLR = BH_OP_MW_DOUBLELAYER_ROTATED;
        const bool r67 = P.far.Qo == 12 && P.far.Qi == 13;
        const bool s67 = sizeof(R) == 8 && r67 && P.far.shifted;
#define BH_TILES_DLR(RM)                                                                     \
    do {                                                                                     \
        if (s67) k_far_tiles<R, DLR, RM, 12, 13, true><<<grid, 256, 0, ctx.stream>>>(A);    \
        else if (r67) k_far_tiles<R, DLR, RM, 12, 13><<<grid, 256, 0, ctx.stream>>>(A);     \
        else k_far_tiles<R, DLR, RM, 0, 0><<<grid, 256, 0, ctx.stream>>>(A);                \
    } while (0)
        if (rowmajor) BH_TILES_DLR(true);
        else BH_TILES_DLR(false);
#undef BH_TILES_DLR
    } else if (rowmajor) {
        if (sl) BH_TILES(BH_OP_MW_SINGLELAYER, true);
        else BH_TILES(BH_OP_MW_DOUBLELAYER, true);
    } else {
        if (sl) BH_TILES(BH_OP_MW_SINGLELAYER, false);
        else BH_TILES(BH_OP_MW_DOUBLELAYER, false);
    }
#undef BH_TILES
    check_launch("k_far_tiles");
}

}  // namespace

// Local operators on RT x RT over one mesh (the matched-mesh loop of
// localop.jl:114-146): Z[m, n] += scale a (sum_q w_q f_i(x_q) . T f_j(x_q)) b over the
// cells m and n share, f_i = (x - v_i) / (2A); T f = n x f for NCross
// (identityop.jl:20-22) and T f = f for Identity (identityop.jl:9-10).  One thread per
// test dof, which alone writes its row, in shape order (deterministic).
template <typename R, bool ROWMAJOR, int KIND>
__global__ __launch_bounds__(256) void k_localop(const double *__restrict__ geo, const double *__restrict__ qp, int Q,
                                                 int ndofs, const int *__restrict__ ptr, const int *__restrict__ cell,
                                                 const int *__restrict__ ref, const double *__restrict__ coeff,
                                                 const int *__restrict__ e2d, const double *__restrict__ e2c, int kmax,
                                                 cx<R> *Z, long long ld, double sr, double si)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= ndofs) return;
    for (int k = ptr[m]; k < ptr[m + 1]; ++k) {
        const int c = cell[k], i = ref[k];
        const double a = coeff[k];
        if (a == 0.0) continue;
        const double *g = geo + (size_t)c * kGeo;
        const double *nrm = g + kGeoN;
        const double s2a = 1.0 / (2.0 * g[kGeoA]);
        double loc[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < Q; ++q) {
            const double *pq = qp + ((size_t)c * Q + q) * 4;
            const double x[3] = {pq[0] + g[kGeoC], pq[1] + g[kGeoC + 1], pq[2] + g[kGeoC + 2]};
            const double fi[3] = {(x[0] - g[3 * i]) * s2a, (x[1] - g[3 * i + 1]) * s2a, (x[2] - g[3 * i + 2]) * s2a};
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double fj[3] = {(x[0] - g[3 * j]) * s2a, (x[1] - g[3 * j + 1]) * s2a,
                                      (x[2] - g[3 * j + 2]) * s2a};
                if constexpr (KIND == BH_OP_NCROSS) {
                    double nf[3];
                    cross3(nf, nrm, fj);
                    loc[j] += pq[3] * dot3(fi, nf);
                } else {
                    loc[j] += pq[3] * dot3(fi, fj);
                }
            }
        }
        for (int j = 0; j < 3; ++j)
            for (int kb = 0; kb < kmax; ++kb) {
                const int n = e2d[((size_t)c * 3 + j) * kmax + kb];
                if (n < 0) continue;
                const double v = a * loc[j] * e2c[((size_t)c * 3 + j) * kmax + kb];
                cx<R> &z = Z[ROWMAJOR ? (long long)m * ld + n : (long long)n * ld + m];
                z.r += (R)(sr * v);
                z.i += (R)(si * v);
            }
    }
}

void assemble_local(Ctx &ctx, int kind, const bh_quadstrat *qs, bool fp32, Space &X, Space &Y, void *dZ, long long ld,
                    bool rowmajor, bool accumulate, double sr, double si)
{
    const char *name = kind == BH_OP_NCROSS ? "NCross" : "Identity";
    BH_REQUIRE(qs != nullptr, "quadstrat is NULL");
    if (qs->kind != BH_QS_SINGLENUM)
        BH_FAIL(BH_ERR_UNSUPPORTED, "%s needs SingleNumQStrat (quadstrat kind %d given)", name, qs->kind);
    if (X.mesh != Y.mesh) BH_FAIL(BH_ERR_UNSUPPORTED, "%s on different test and trial meshes has no GPU path", name);
    const int degree = qs->params[0];
    check_degree(degree, "SingleNumQStrat");
    const int64_t M = X.ndofs, N = Y.ndofs;
    const size_t es = fp32 ? 8 : 16;
    if (!accumulate) {
        if (rowmajor) BH_CHECK_HIP(hipMemset2DAsync(dZ, ld * es, 0, N * es, M, ctx.stream));
        else BH_CHECK_HIP(hipMemset2DAsync(dZ, ld * es, 0, M * es, N, ctx.stream));
    }
    int Q = 0;
    const double *qp = (const double *)quad_points(*X.mesh, degree, &Q, false, ctx.stream);
    const unsigned grid = (unsigned)((M + 255) / 256);
    ProfScope prof(ctx, "k_localop", (double)M * 2 * Q * 3 * 12.0);
#define BH_LOCAL(R, RM, K)                                                                                        \
    k_localop<R, RM, K><<<grid, 256, 0, ctx.stream>>>(X.mesh->d_geo.p, qp, Q, (int)M, X.d_ptr.p, X.d_cell.p,     \
                                                      X.d_ref.p, X.d_coeff.p, Y.d_e2d.p, Y.d_e2c.p, Y.kmax,       \
                                                      (cx<R> *)dZ, ld, sr, si)
#define BH_LOCAL_K(R, RM)                                                                                         \
    do {                                                                                                          \
        if (kind == BH_OP_NCROSS) BH_LOCAL(R, RM, BH_OP_NCROSS);                                                 \
        else BH_LOCAL(R, RM, BH_OP_IDENTITY);                                                                    \
    } while (0)
    if (fp32) {
        if (rowmajor) BH_LOCAL_K(float, true);
        else BH_LOCAL_K(float, false);
    } else {
        if (rowmajor) BH_LOCAL_K(double, true);
        else BH_LOCAL_K(double, false);
    }
#undef BH_LOCAL_K
#undef BH_LOCAL
    check_launch("k_localop");
    BH_CHECK_HIP(hipStreamSynchronize(ctx.stream));
}

bool assemble_device(Ctx &ctx, const bh_operator *op, const bh_quadstrat *qs, bool fp32, Space &X, Space &Y, void *dZ,
                     long long ld, bool rowmajor, bool accumulate, double sr, double si, void *dZ32, long long ld32)
{
    BH_REQUIRE(op != nullptr, "operator is NULL");
    if (op->kind == BH_OP_NCROSS || op->kind == BH_OP_IDENTITY) {
        if (X.ndofs == 0 || Y.ndofs == 0) return false;
        assemble_local(ctx, op->kind, qs, fp32, X, Y, dZ, ld, rowmajor, accumulate, sr, si);
        return false;
    }
    Plan P;
    make_plan(ctx, op, qs, *X.mesh, *Y.mesh, fp32, P);
    build_tiles(X, ctx.stream);
    build_tiles(Y, ctx.stream);
    if (X.ndofs == 0 || Y.ndofs == 0) return false;
    const bool copy32 = dZ32 != nullptr && !fp32 && rowmajor;
    DevBuf<char> rpack;   // held until the launch has ended (synchronised below)
    if (fp32) launch_tiles<float>(ctx, P, X, Y, dZ, ld, rowmajor, accumulate, sr, si, rpack);
    else launch_tiles<double>(ctx, P, X, Y, dZ, ld, rowmajor, accumulate, sr, si, rpack, copy32 ? dZ32 : nullptr, ld32);
    BH_CHECK_HIP(hipStreamSynchronize(ctx.stream));
    return copy32;
}

// ============================================================ batched assembly
// The items of a batch (test cells, non-far pairs, dof tiles) of all requests are numbered
// consecutively; a block or wave finds its request in the table of each request's first item
// and runs the same body as the single-request kernels, so every entry is bit-identical to
// the one a request assembled alone would get.

// One request's classification and non-far blocks.
struct BatchNear {
    const double *tgeo, *sgeo;
    int nfs;
    int mode;
    double k2;
    int use_grid, fill_grid;
    PairGridView grid;
    int *counts;          // [nft]
    const int *nptr;      // [nft + 1]
    int4 *entries;        // [nent]
    double *blocks;       // [nent][18]
    int *err;             // hits > 3 seen
    OpParams op;
    GLRule gl[3];         // common face, common edge, common vertex
    const double *tq, *sq;
    int Qo, Qi;           // near rules (Wilton)
};

__global__ __launch_bounds__(256) void k_classify_count_batch(const BatchNear *reqs, const int *first, int n)
{
    const int r = batch_find(first, n, blockIdx.x);
    const BatchNear &B = reqs[r];
    const int p = blockIdx.x - first[r];
    if (B.use_grid) classify_count_grid_dev(B.tgeo, B.sgeo, B.grid, B.k2, B.mode, p, B.counts, B.err);
    else classify_count_dev(B.tgeo, B.sgeo, B.nfs, B.k2, B.mode, p, B.counts, B.err);
}

__global__ __launch_bounds__(256) void k_classify_fill_batch(const BatchNear *reqs, const int *first, int n)
{
    const int r = batch_find(first, n, blockIdx.x);
    const BatchNear &B = reqs[r];
    const int p = blockIdx.x - first[r];
    if (B.fill_grid) classify_fill_grid_dev(B.tgeo, B.sgeo, B.grid, B.k2, B.mode, B.nptr, B.entries, p);
    else classify_fill_dev(B.tgeo, B.sgeo, B.nfs, B.k2, B.mode, B.nptr, B.entries, p);
}

// (request, entry) lists per (operator, class) over all entries of the batch: list
// 5 * op.kind + class, with op.kind in {0, 1, 3} (the integral operators)
constexpr int kBatchLists = 5 * (BH_OP_MW_DOUBLELAYER_ROTATED + 1);
__global__ void k_class_lists_batch(const BatchNear *reqs, const int *first, int n, int total, int *counters,
                                    int2 *lists)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = g < total;
    int r = 0, e = 0, l = 0;
    if (active) {
        r = batch_find(first, n, g);
        e = g - first[r];
        l = 5 * reqs[r].op.kind + reqs[r].entries[e].z;   // class 1..4
    }
    const int slot = list_slot(counters, l, active);
    if (active) lists[(size_t)l * total + slot] = make_int2(r, e);
}

template <int OPK>
__global__ __launch_bounds__(256) void k_sauter_schwab_batch(const BatchNear *reqs, const int2 *list, int nlist,
                                                             int cls)
{
    const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wv >= nlist) return;
    const int2 re = list[wv];
    const BatchNear &B = reqs[re.x];
    const GLRule &gl = B.gl[cls == kCF ? 0 : (cls == kCE ? 1 : 2)];
    ss_dev<OPK>(B.tgeo, B.sgeo, B.entries, re.y, cls, gl, B.op, B.blocks, threadIdx.x & 63);
}

template <int OPK>
__global__ __launch_bounds__(256) void k_wilton_batch(const BatchNear *reqs, const int2 *list, int nlist)
{
    const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wv >= nlist) return;
    const int2 re = list[wv];
    const BatchNear &B = reqs[re.x];
    wilton_dev<OPK>(B.tgeo, B.sgeo, B.entries, re.y, B.tq, B.Qo, B.sq, B.Qi, B.op, B.blocks, threadIdx.x & 63);
}

// dof tiles of several requests of one kernel configuration; tile t of a request is
// (row t / column tiles, column t mod column tiles)
template <typename R, int OPK, bool ROWMAJOR, int QO, int QI, bool SHIFT = false>
__global__ __launch_bounds__(256, far_min_blocks<OPK>()) void k_far_tiles_batch(const TileArgs *args, const int *first,
                                                                                 int n)
{
    const int r = batch_find(first, n, blockIdx.x);
    const TileArgs A = args[r];
    const int lt = blockIdx.x - first[r];
    const int ntc = (A.col.ndofs + kTile - 1) / kTile;
    far_tiles_dev<R, OPK, ROWMAJOR, QO, QI, SHIFT>(A, lt / ntc, lt % ntc);
}

namespace {

struct BatchItem {
    bh_matrix *M;
    Space *X;
    Strategy st;
    OpParams op;
    FarArgs far;
    long long nent = 0;   // non-far pairs (classified with their own blocks)
};

// Classification, compaction and the non-far blocks of all items with a classifying
// strategy, in shared launches.  Fills each item's far.nptr / nent / nblk.
void batch_near(Ctx &ctx, std::vector<BatchItem> &items, std::vector<PairGrid> &grids, DevBuf<int> &nptr_all,
                DevBuf<int4> &entries_all, DevBuf<double> &blocks_all)
{
    hipStream_t s = ctx.stream;
    std::vector<int> idx;
    for (int i = 0; i < (int)items.size(); ++i)
        if (items[i].st.mode != kModeDoubleNum && items[i].X->mesh->nf > 0) idx.push_back(i);
    const int m = (int)idx.size();
    // DoubleNumQStrat items (and empty meshes) need an empty non-far list: nptr all zero
    size_t zero_len = 1;
    for (auto &it : items) zero_len = std::max(zero_len, (size_t)it.X->mesh->nf + 1);
    std::vector<int> cell_first(m + 1, 0);
    for (int k = 0; k < m; ++k) cell_first[k + 1] = cell_first[k] + (int)items[idx[k]].X->mesh->nf;
    const int ncell = cell_first[m];
    nptr_all.alloc((size_t)ncell + m + zero_len);
    {
        std::vector<int> z(zero_len, 0);
        BH_CHECK_HIP(hipMemcpyAsync(nptr_all.p + ncell + m, z.data(), sizeof(int) * zero_len, hipMemcpyHostToDevice, s));
        BH_CHECK_HIP(hipStreamSynchronize(s));
    }
    for (auto &it : items) {
        it.far.nptr = nptr_all.p + ncell + m;
        it.far.nent = nullptr;
        it.far.nblk = nullptr;
    }
    if (m == 0) return;
    grids.resize(m);
    std::vector<BatchNear> tab(m);
    DevBuf<int> counts(ncell), errs(m), dfirst(m + 1);
    for (int k = 0; k < m; ++k) {
        BatchItem &it = items[idx[k]];
        Mesh &tm = *it.X->mesh;
        build_pair_grid(tm, tm, it.st.mode, s, grids[k]);
        BatchNear &B = tab[k];
        B.tgeo = tm.d_geo.p;
        B.sgeo = tm.d_geo.p;
        B.nfs = (int)tm.nf;
        B.mode = it.st.mode;
        B.k2 = it.op.gr * it.op.gr + it.op.gi * it.op.gi;
        B.use_grid = grids[k].use ? 1 : 0;
        B.grid = grids[k].view;
        B.counts = counts.p + cell_first[k];
        B.err = errs.p + k;
        B.op = it.op;
        const int ord[3] = {it.st.ss_face, it.st.ss_edge, it.st.ss_vert};
        for (int c = 0; c < 3; ++c) {
            B.gl[c].n = ord[c];
            legendre01(ord[c], B.gl[c].x, B.gl[c].w);
        }
        if (it.st.mode == kModeWiltonSauter) {
            B.tq = (const double *)quad_points(tm, it.st.near_o, &B.Qo, false, s);
            B.sq = (const double *)quad_points(tm, it.st.near_i, &B.Qi, false, s);
            BH_REQUIRE(B.Qo <= 64, "near outer rule too large");
        }
    }
    DevBuf<char> dtab(sizeof(BatchNear) * m);
    BH_CHECK_HIP(hipMemsetAsync(errs.p, 0, sizeof(int) * m, s));
    BH_CHECK_HIP(hipMemcpyAsync(dtab.p, tab.data(), sizeof(BatchNear) * m, hipMemcpyHostToDevice, s));
    BH_CHECK_HIP(hipMemcpyAsync(dfirst.p, cell_first.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, s));
    const BatchNear *dreq = (const BatchNear *)dtab.p;
    {
        ProfScope prof(ctx, "k_classify_count", (double)ncell);
        k_classify_count_batch<<<ncell, 256, 0, s>>>(dreq, dfirst.p, m);
    }
    check_launch("k_classify_count_batch");
    std::vector<int> hc(ncell), herr(m);
    BH_CHECK_HIP(hipMemcpyAsync(hc.data(), counts.p, sizeof(int) * ncell, hipMemcpyDeviceToHost, s));
    BH_CHECK_HIP(hipMemcpyAsync(herr.data(), errs.p, sizeof(int) * m, hipMemcpyDeviceToHost, s));
    BH_CHECK_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < m; ++k)
        if (herr[k]) BH_FAIL(BH_ERR_ASSERT, "AssertionError: hits <= 3 (two cells share more than 3 vertex positions)"
                                            " in batch item %d", idx[k]);
    // per-request offsets; entries of request k start at ent_first[k]
    std::vector<int> off(ncell + m), ent_first(m + 1, 0);
    for (int k = 0; k < m; ++k) {
        const int nft = cell_first[k + 1] - cell_first[k];
        int *o = off.data() + cell_first[k] + k;
        o[0] = 0;
        int hmax = 0;
        for (int p = 0; p < nft; ++p) {
            o[p + 1] = o[p] + hc[cell_first[k] + p];
            hmax = std::max(hmax, hc[cell_first[k] + p]);
        }
        tab[k].fill_grid = (tab[k].use_grid && hmax <= kGridFillCap) ? 1 : 0;
        ent_first[k + 1] = ent_first[k] + o[nft];
    }
    const int nent = ent_first[m];
    BH_CHECK_HIP(hipMemcpyAsync(nptr_all.p, off.data(), sizeof(int) * (ncell + m), hipMemcpyHostToDevice, s));
    entries_all.alloc(std::max(nent, 1));
    blocks_all.alloc((size_t)std::max(nent, 1) * 18);
    for (int k = 0; k < m; ++k) {
        BatchItem &it = items[idx[k]];
        tab[k].nptr = nptr_all.p + cell_first[k] + k;
        tab[k].entries = entries_all.p + ent_first[k];
        tab[k].blocks = blocks_all.p + (size_t)ent_first[k] * 18;
        it.far.nptr = tab[k].nptr;
        it.far.nent = tab[k].entries;
        it.far.nblk = tab[k].blocks;
        it.nent = ent_first[k + 1] - ent_first[k];
    }
    if (nent == 0) return;
    BH_CHECK_HIP(hipMemcpyAsync(dtab.p, tab.data(), sizeof(BatchNear) * m, hipMemcpyHostToDevice, s));
    k_classify_fill_batch<<<ncell, 256, 0, s>>>(dreq, dfirst.p, m);
    check_launch("k_classify_fill_batch");
    DevBuf<int> efirst(m + 1), counters(kBatchLists);
    DevBuf<int2> lists((size_t)kBatchLists * nent);
    BH_CHECK_HIP(hipMemcpyAsync(efirst.p, ent_first.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, s));
    BH_CHECK_HIP(hipMemsetAsync(counters.p, 0, kBatchLists * sizeof(int), s));
    k_class_lists_batch<<<(nent + 255) / 256, 256, 0, s>>>(dreq, efirst.p, m, nent, counters.p, lists.p);
    check_launch("k_class_lists_batch");
    int hcnt[kBatchLists];
    BH_CHECK_HIP(hipMemcpyAsync(hcnt, counters.p, kBatchLists * sizeof(int), hipMemcpyDeviceToHost, s));
    BH_CHECK_HIP(hipStreamSynchronize(s));
    for (int opk : {BH_OP_MW_SINGLELAYER, BH_OP_MW_DOUBLELAYER, BH_OP_MW_DOUBLELAYER_ROTATED}) {
        for (int cls : {kCF, kCE, kCV}) {
            const int l = 5 * opk + cls, cnt = hcnt[l];
            if (cnt == 0) continue;
            const int2 *lp = lists.p + (size_t)l * nent;
            const unsigned grid = (unsigned)((cnt + 3) / 4);
            ProfScope prof(ctx, "k_sauter_schwab", (double)cnt);
            if (opk == BH_OP_MW_SINGLELAYER)
                k_sauter_schwab_batch<BH_OP_MW_SINGLELAYER><<<grid, 256, 0, s>>>(dreq, lp, cnt, cls);
            else if (opk == BH_OP_MW_DOUBLELAYER)
                k_sauter_schwab_batch<BH_OP_MW_DOUBLELAYER><<<grid, 256, 0, s>>>(dreq, lp, cnt, cls);
            else
                k_sauter_schwab_batch<BH_OP_MW_DOUBLELAYER_ROTATED><<<grid, 256, 0, s>>>(dreq, lp, cnt, cls);
            check_launch("k_sauter_schwab_batch");
        }
        const int l = 5 * opk + kNear, cnt = hcnt[l];
        if (cnt == 0) continue;
        BH_REQUIRE(opk != BH_OP_MW_DOUBLELAYER_ROTATED, "near pairs of a rotated double layer");
        const int2 *lp = lists.p + (size_t)l * nent;
        ProfScope prof(ctx, "k_wilton", (double)cnt);
        if (opk == BH_OP_MW_SINGLELAYER)
            k_wilton_batch<BH_OP_MW_SINGLELAYER><<<(cnt + 3) / 4, 256, 0, s>>>(dreq, lp, cnt);
        else
            k_wilton_batch<BH_OP_MW_DOUBLELAYER><<<(cnt + 3) / 4, 256, 0, s>>>(dreq, lp, cnt);
        check_launch("k_wilton_batch");
    }
    BH_CHECK_HIP(hipStreamSynchronize(s));   // the tables go out of scope
}

// One far-tile launch per kernel configuration (operator, default rules, shifted phases).
void batch_far(Ctx &ctx, std::vector<BatchItem> &items)
{
    hipStream_t s = ctx.stream;
    struct Group {
        int opk;
        int variant;   // rule specialisation and shifted phases (see below)
        std::vector<int> members;
    };
    std::vector<Group> groups;
    for (int i = 0; i < (int)items.size(); ++i) {
        BatchItem &it = items[i];
        if (it.X->ndofs == 0) continue;
        // 0 generic rules, 1 and 2 the default far rules (3 x 4 points) without and with
        // shifted phases, 3 and 4 the rotated double layer's default (12 x 13 points)
        const bool rot = it.op.kind == BH_OP_MW_DOUBLELAYER_ROTATED;
        const bool dflt = it.far.Qo == 3 && it.far.Qi == 4 && !rot;
        const bool r67 = it.far.Qo == 12 && it.far.Qi == 13 && rot;
        const int variant = dflt ? (it.far.shifted ? 2 : 1) : (r67 ? (it.far.shifted ? 4 : 3) : 0);
        Group *g = nullptr;
        for (auto &gg : groups)
            if (gg.opk == it.op.kind && gg.variant == variant) g = &gg;
        if (!g) {
            groups.push_back({it.op.kind, variant, {}});
            g = &groups.back();
        }
        g->members.push_back(i);
    }
    std::vector<DevBuf<char>> keep;
    for (const Group &g : groups) {
        const int n = (int)g.members.size();
        std::vector<TileArgs> args(n);
        std::vector<int> first(n + 1, 0);
        double far_pp = 0.0;
        for (int k = 0; k < n; ++k) {
            BatchItem &it = items[g.members[k]];
            Space &X = *it.X;
            bh_matrix &M = *it.M;
            TileArgs &A = args[k];
            A = TileArgs{};
            A.far = it.far;
            A.row = {X.t_ptr.p, X.t_cells.p, X.t_slot.p, X.t_coef.p, (int)X.ndofs};
            A.col = A.row;
            A.Z = M.a64.p;
            A.ld = M.n;
            A.Z32 = M.lu32.p;
            A.ld32 = padded_ld(M.n);
            A.rowmajor = 1;
            A.accumulate = 0;
            A.scale_r = 1.0;
            A.scale_i = 0.0;
            if (g.opk == BH_OP_MW_SINGLELAYER && (g.variant == 1 || g.variant == 2)) {
                keep.emplace_back();
                A.rpack = pack_rows<double>(s, X, it.far, keep.back());
            }
            BH_REQUIRE((long long)first[k] + X.ntiles * X.ntiles < (1LL << 31), "batch has too many dof tiles");
            first[k + 1] = first[k] + (int)(X.ntiles * X.ntiles);
            const double nf = (double)X.mesh->nf;
            far_pp += (nf * nf - (double)it.nent) * it.far.Qo * it.far.Qi;
        }
        DevBuf<char> dargs(sizeof(TileArgs) * n), dfirst(sizeof(int) * (n + 1));
        BH_CHECK_HIP(hipMemcpyAsync(dargs.p, args.data(), sizeof(TileArgs) * n, hipMemcpyHostToDevice, s));
        BH_CHECK_HIP(hipMemcpyAsync(dfirst.p, first.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
        const TileArgs *da = (const TileArgs *)dargs.p;
        const int *df = (const int *)dfirst.p;
        const double per_pp = g.opk == BH_OP_MW_SINGLELAYER ? 170.0 : 210.0;
        ProfScope prof(ctx, "k_far_tiles", far_pp * per_pp);
        const unsigned grid = (unsigned)first[n];
#define BH_BTILES(OPK)                                                                                           \
    do {                                                                                                         \
        if (g.variant == 2) k_far_tiles_batch<double, OPK, true, 3, 4, true><<<grid, 256, 0, s>>>(da, df, n);    \
        else if (g.variant == 1) k_far_tiles_batch<double, OPK, true, 3, 4><<<grid, 256, 0, s>>>(da, df, n);     \
        else k_far_tiles_batch<double, OPK, true, 0, 0><<<grid, 256, 0, s>>>(da, df, n);                         \
    } while (0)
        if (g.opk == BH_OP_MW_SINGLELAYER) BH_BTILES(BH_OP_MW_SINGLELAYER);
        else if (g.opk == BH_OP_MW_DOUBLELAYER) BH_BTILES(BH_OP_MW_DOUBLELAYER);
        else if (g.variant == 4)
            k_far_tiles_batch<double, BH_OP_MW_DOUBLELAYER_ROTATED, true, 12, 13, true><<<grid, 256, 0, s>>>(da, df, n);
        else if (g.variant == 3)
            k_far_tiles_batch<double, BH_OP_MW_DOUBLELAYER_ROTATED, true, 12, 13><<<grid, 256, 0, s>>>(da, df, n);
        else k_far_tiles_batch<double, BH_OP_MW_DOUBLELAYER_ROTATED, true, 0, 0><<<grid, 256, 0, s>>>(da, df, n);
#undef BH_BTILES
        check_launch("k_far_tiles_batch");
        keep.push_back(std::move(dargs));
        keep.push_back(std::move(dfirst));
    }
    BH_CHECK_HIP(hipStreamSynchronize(s));
}

}  // namespace

}  // namespace bh

using namespace bh;

extern "C" int bh_assemble(bh_ctx *ctx, const bh_operator *op, const bh_quadstrat *qs, int precision, const bh_space *X,
                           const bh_space *Y, const double *scale, int accumulate, void *Z, int64_t ld,
                           int64_t row_off, int64_t col_off)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx && X && Y, "NULL handle");
    BH_REQUIRE(precision == BH_FP64 || precision == BH_FP32, "precision must be BH_FP64 or BH_FP32");
    BH_REQUIRE(X->mesh->ctx == ctx && Y->mesh->ctx == ctx, "spaces belong to another context");
    const int64_t M = X->ndofs, N = Y->ndofs;
    if (M == 0 || N == 0) return BH_OK;   // integralop.jl:72-73
    BH_REQUIRE(Z != nullptr, "Z is NULL");
    BH_REQUIRE(row_off >= 0 && col_off >= 0 && ld >= row_off + M, "ld/offsets do not fit the %lld x %lld block",
               (long long)M, (long long)N);
    const bool fp32 = precision == BH_FP32;
    const size_t es = fp32 ? 8 : 16;
    DevBuf<char> dZ((size_t)M * N * es);
    char *hz = (char *)Z + (size_t)(row_off + col_off * ld) * es;
    if (accumulate)
        BH_CHECK_HIP(hipMemcpy2DAsync(dZ.p, M * es, hz, ld * es, M * es, N, hipMemcpyHostToDevice, ctx->stream));
    const double sr = scale ? scale[0] : 1.0, si = scale ? scale[1] : 0.0;
    assemble_device(*ctx, op, qs, fp32, *const_cast<bh_space *>(X), *const_cast<bh_space *>(Y), dZ.p, M, false,
                    accumulate != 0, sr, si);
    BH_CHECK_HIP(hipMemcpy2DAsync(hz, ld * es, dZ.p, M * es, M * es, N, hipMemcpyDeviceToHost, ctx->stream));
    BH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    BH_API_END
}

extern "C" int bh_classify(bh_ctx *ctx, const bh_operator *op, const bh_quadstrat *qs, const bh_mesh *test,
                           const bh_mesh *trial, int8_t *classes)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx && test && trial && classes, "NULL argument");
    const Strategy st = parse_qs(qs);
    const OpParams p = parse_op(op);
    const size_t n = (size_t)test->nf * trial->nf;
    if (n == 0) return BH_OK;
    DevBuf<int8_t> d(n);
    const double k2 = p.gr * p.gr + p.gi * p.gi;
    k_classify_matrix<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(test->d_geo.p, (int)test->nf,
                                                                           trial->d_geo.p, (int)trial->nf, k2,
                                                                           st.mode, d.p);
    check_launch("k_classify_matrix");
    BH_CHECK_HIP(hipMemcpyAsync(classes, d.p, n, hipMemcpyDeviceToHost, ctx->stream));
    BH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    BH_API_END
}

extern "C" int bh_local_blocks(bh_ctx *ctx, const bh_operator *op, const bh_quadstrat *qs, const bh_mesh *test,
                               const bh_mesh *trial, int64_t npairs, const int32_t *test_cells,
                               const int32_t *trial_cells, double *blocks)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx && test && trial && blocks, "NULL argument");
    if (npairs <= 0) return BH_OK;
    for (int64_t k = 0; k < npairs; ++k)
        BH_REQUIRE(test_cells[k] >= 0 && test_cells[k] < test->nf && trial_cells[k] >= 0 && trial_cells[k] < trial->nf,
                   "pair %lld out of range", (long long)k);
    Plan P;
    make_plan(*ctx, op, qs, *const_cast<bh_mesh *>(test), *const_cast<bh_mesh *>(trial), false, P);
    DevBuf<int> tc(npairs), sc(npairs);
    DevBuf<double> out((size_t)npairs * 18);
    tc.upload(test_cells, npairs, ctx->stream);
    sc.upload(trial_cells, npairs, ctx->stream);
    const unsigned grid = (unsigned)((npairs + 127) / 128);
    if (P.op.kind == BH_OP_MW_SINGLELAYER)
        k_far_pairs<double, BH_OP_MW_SINGLELAYER><<<grid, 128, 0, ctx->stream>>>(P.far, tc.p, sc.p, (int)npairs, out.p);
    else if (P.op.kind == BH_OP_MW_DOUBLELAYER_ROTATED)
        k_far_pairs<double, BH_OP_MW_DOUBLELAYER_ROTATED><<<grid, 128, 0, ctx->stream>>>(P.far, tc.p, sc.p,
                                                                                         (int)npairs, out.p);
    else
        k_far_pairs<double, BH_OP_MW_DOUBLELAYER><<<grid, 128, 0, ctx->stream>>>(P.far, tc.p, sc.p, (int)npairs, out.p);
    check_launch("k_far_pairs");
    BH_CHECK_HIP(hipMemcpyAsync(blocks, out.p, sizeof(double) * 18 * npairs, hipMemcpyDeviceToHost, ctx->stream));
    BH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    BH_API_END
}

extern "C" int bh_check_strategy(const bh_operator *op, const bh_quadstrat *qs)
{
    BH_API_BEGIN
    const OpParams p = parse_op(op);
    BH_REQUIRE(qs != nullptr, "quadstrat is NULL");
    if (p.kind == BH_OP_NCROSS || p.kind == BH_OP_IDENTITY) {
        if (qs->kind != BH_QS_SINGLENUM)
            BH_FAIL(BH_ERR_UNSUPPORTED, "local operators need SingleNumQStrat (quadstrat kind %d given)", qs->kind);
        check_degree(qs->params[0], "SingleNumQStrat");
    } else {
        const Strategy st = parse_qs(qs);
        if (p.kind == BH_OP_MW_DOUBLELAYER_ROTATED && st.mode == kModeWiltonSauter)
            BH_FAIL(BH_ERR_UNSUPPORTED, "DoubleLayerRotatedMW3D has no singularity extraction");
    }
    BH_API_END
}

namespace bh {
// Resident matrices of integral operators (FP64 row-major, with the FP32 copy): the batched
// kernels serve one request as well as many, so a matrix has the same bits whatever batch
// it was assembled in.
void assemble_resident_batch(bh_ctx *ctx, int64_t count, const bh_operator *ops, const bh_quadstrat *qs,
                             const bh_space *const *spaces, bh_matrix **out)
{
    for (int64_t i = 0; i < count; ++i) out[i] = nullptr;
    if (count == 0) return;
    std::vector<std::unique_ptr<bh_matrix>> mats(count);
    std::vector<BatchItem> items(count);
    for (int64_t i = 0; i < count; ++i) {
        BH_REQUIRE(spaces[i] != nullptr, "space %lld is NULL", (long long)i);
        Space &X = *const_cast<bh_space *>(spaces[i]);
        BH_REQUIRE(X.mesh->ctx == ctx, "space %lld belongs to another context", (long long)i);
        BatchItem &it = items[i];
        it.op = parse_op(&ops[i]);
        if (it.op.kind == BH_OP_NCROSS || it.op.kind == BH_OP_IDENTITY)
            BH_FAIL(BH_ERR_UNSUPPORTED, "batch item %lld: local operators are added with bh_matrix_accumulate",
                    (long long)i);
        it.st = parse_qs(&qs[i]);
        if (it.op.kind == BH_OP_MW_DOUBLELAYER_ROTATED && it.st.mode == kModeWiltonSauter)
            BH_FAIL(BH_ERR_UNSUPPORTED, "batch item %lld: DoubleLayerRotatedMW3D has no singularity extraction",
                    (long long)i);
        it.X = &X;
        build_tiles(X, ctx->stream);
        Mesh &m = *X.mesh;
        FarArgs &F = it.far;
        F = FarArgs{};
        F.tgeo = m.d_geo.p;
        F.sgeo = m.d_geo.p;
        F.tq = quad_points(m, it.st.far_o, &F.Qo, false, ctx->stream);
        F.sq = quad_points(m, it.st.far_i, &F.Qi, false, ctx->stream);
        F.mode = it.st.mode;
        F.k2 = it.op.gr * it.op.gr + it.op.gi * it.op.gi;
        F.shifted = shifted_phases(it.op, m, m, false) ? 1 : 0;
        F.op = it.op;
        std::unique_ptr<bh_matrix> M(new bh_matrix());
        M->ctx = ctx;
        M->n = X.ndofs;
        pool_alloc(*ctx, M->a64, (size_t)M->n * M->n * 2);
        if (M->n > 0) pool_alloc(*ctx, M->lu32, (size_t)padded_ld(M->n) * M->n * 2);
        it.M = M.get();
        mats[i] = std::move(M);
    }
    std::vector<PairGrid> grids;
    DevBuf<int> nptr_all;
    DevBuf<int4> entries_all;
    DevBuf<double> blocks_all;
    batch_near(*ctx, items, grids, nptr_all, entries_all, blocks_all);
    batch_far(*ctx, items);
    for (int64_t i = 0; i < count; ++i) {
        bh_matrix *M = mats[i].release();
        M->lu32_fresh = M->n > 0;   // the tiles wrote the FP32 copy
        ctx_retain(ctx);
        out[i] = M;
    }
}
}  // namespace bh

extern "C" int bh_matrix_assemble_batch(bh_ctx *ctx, int64_t count, const bh_operator *ops, const bh_quadstrat *qs,
                                        const bh_space *const *spaces, bh_matrix **out)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx && ops && qs && spaces && out, "NULL argument");
    BH_REQUIRE(count >= 0 && count <= (1 << 20), "bad batch size %lld", (long long)count);
    assemble_resident_batch(ctx, count, ops, qs, spaces, out);
    BH_API_END
}
