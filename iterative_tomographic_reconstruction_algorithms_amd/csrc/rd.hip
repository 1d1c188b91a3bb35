// Ray-driven projector pair (semantics of radon_rd.c:15-189) for gfx950.
//
// Five sub-rays per detector bin (sub-ray i = 6b+1..6b+5 of bin b, radon_rd.c:123-185),
// each clipped to the square [-nh, nh]^2 and walked cell by cell; the length walked
// since the previous step is credited, with weight 1/5, to the cell just entered, and
// cells with row or column index 0 are skipped (radon_rd.c:44-66).  The walk repeats
// the reference build's arithmetic operation for operation (float geometry, double
// step parameters, double fused update rounded to float after every step).
//
// Forward: one thread per (bin, view) walks its five sub-rays in order.
// Back: one thread per pixel gathers, view by view and sub-ray by sub-ray in increasing
// index, the one step (if any) of each sub-ray's walk that enters the pixel, and applies
// it with the reference's rounding.  The walk is a merge of two increasing sequences of
// crossing parameters (x boundaries and y boundaries), so the step entering a cell, the
// step before it and whether the walk gets there follow from at most six of those
// parameters, computed exactly as the walk computes them.  Updates reach each pixel in
// the reference's own order: both projections are bit-identical to the reference's.
#include <algorithm>
#include <vector>

#include "epilogue.h"

namespace itomo {

static constexpr double kHalfPi = 3.141592653589793 / 2;

template <int ARITH>
__device__ __forceinline__ float rd_round(float v)
{
    if (ARITH == ITOMO_ARITH_FAST)  // gcc -ffast-math inline expansion of round()
        return truncf(__fadd_rn(v, copysignf(0.49999997f, v)));
    return (float)round((double)v);
}

// Entry/exit points of sub-ray i (radon_rd.c:123-179).
template <int ARITH>
__device__ __forceinline__ void rd_endpoints(int i, int npix, const RdView& g, float p1[2],
                                             float p2[2])
{
    const int nh = (int)(npix * 0.5);
    float t;
    if (ARITH == ITOMO_ARITH_FAST)
        t = (float)fabs(__fma_rn((double)i, 1.0 / 6.0, -(double)nh));
    else
        t = (float)fabs(__dsub_rn(__ddiv_rn((double)i, 6.0), (double)nh));
    const bool lt = g.lt_half_pi != 0;
    const bool gt = (double)g.theta > kHalfPi;
    if (i >= 3 * npix && lt) t = -t;
    if (i < 3 * npix && gt) t = -t;

    const float s = g.s, c = g.c;
    const float nhf = (float)nh, mnhf = (float)(-nh);
    float x1, y1, x2, y2;
    if (lt) {
        if (ARITH == ITOMO_ARITH_FAST) {
            y1 = __fmul_rn(__fadd_rn(t, __fmul_rn(mnhf, s)), g.rc);
            x1 = __fmul_rn(__fsub_rn(__fmul_rn(mnhf, c), t), g.rs);
            y2 = __fmul_rn(__fadd_rn(t, __fmul_rn(nhf, s)), g.rc);
            x2 = __fmul_rn(__fsub_rn(__fmul_rn(nhf, c), t), g.rs);
        } else {
            y1 = __fadd_rn(__fdiv_rn(__fmul_rn(mnhf, s), c), __fdiv_rn(t, c));
            x1 = __fsub_rn(__fdiv_rn(__fmul_rn(mnhf, c), s), __fdiv_rn(t, s));
            y2 = __fadd_rn(__fdiv_rn(__fmul_rn(nhf, s), c), __fdiv_rn(t, c));
            x2 = __fsub_rn(__fdiv_rn(__fmul_rn(nhf, c), s), __fdiv_rn(t, s));
        }
        if (y1 >= -nh && y1 <= nh) { p1[0] = mnhf; p1[1] = y1; }
        else                       { p1[0] = x1;   p1[1] = mnhf; }
        if (y2 >= -nh && y2 <= nh) { p2[0] = nhf;  p2[1] = y2; }
        else                       { p2[0] = x2;   p2[1] = nhf; }
    } else {
        if (ARITH == ITOMO_ARITH_FAST) {
            y1 = __fmul_rn(__fsub_rn(__fmul_rn(nhf, s), t), g.rc);
            x1 = __fmul_rn(__fadd_rn(t, __fmul_rn(mnhf, c)), g.rs);
            y2 = __fmul_rn(__fsub_rn(__fmul_rn(mnhf, s), t), g.rc);
            x2 = __fmul_rn(__fadd_rn(t, __fmul_rn(nhf, c)), g.rs);
        } else {
            y1 = __fsub_rn(__fdiv_rn(__fmul_rn(nhf, s), c), __fdiv_rn(t, c));
            x1 = __fadd_rn(__fdiv_rn(__fmul_rn(mnhf, c), s), __fdiv_rn(t, s));
            y2 = __fsub_rn(__fdiv_rn(__fmul_rn(mnhf, s), c), __fdiv_rn(t, c));
            x2 = __fadd_rn(__fdiv_rn(__fmul_rn(nhf, c), s), __fdiv_rn(t, s));
        }
        if (y1 >= -nh && y1 <= nh) { p1[0] = nhf;  p1[1] = y1; }
        else                       { p1[0] = x1;   p1[1] = mnhf; }
        if (y2 >= -nh && y2 <= nh) { p2[0] = mnhf; p2[1] = y2; }
        else                       { p2[0] = x2;   p2[1] = nhf; }
    }
}

// Float update of one pixel, exactly as the reference build rounds it.
template <int ARITH>
__device__ __forceinline__ float rd_back_update(float pix, float sv, float l)
{
    if (ARITH == ITOMO_ARITH_FAST)
        return (float)__fma_rn(__dmul_rn((double)sv, 0.2), (double)l, (double)pix);
    return (float)__dadd_rn((double)pix, __dmul_rn(__dmul_rn(0.2, (double)l), (double)sv));
}

template <int ARITH>
__device__ __forceinline__ float rd_fwd_update(float acc, float pix, float l)
{
    if (ARITH == ITOMO_ARITH_FAST)
        return (float)__fma_rn(__dmul_rn((double)l, 0.2), (double)pix, (double)acc);
    return (float)__dadd_rn((double)acc, __dmul_rn(__dmul_rn(0.2, (double)l), (double)pix));
}

// One sub-ray walk of the forward projection (radon_rd.c:15-71): returns the updated
// bin value.  The walk is capped at 4*npix+16 steps (a genuine walk takes at most
// 2*npix+2), so it always ends.
template <int ARITH, bool ONES>
__device__ float rd_walk(float acc, const float* __restrict__ img_in, int npix,
                         const float p1[2], const float p2[2])
{
    const int nh = (int)(npix * 0.5);
    const float d0 = __fsub_rn(p2[0], p1[0]), d1 = __fsub_rn(p2[1], p1[1]);
    // float sqrt via double: rounding a correctly rounded double square root to float
    // gives the correctly rounded float square root (53 >= 2*24 + 2).
    const float incr = (ARITH == ITOMO_ARITH_FAST)
                           ? (float)__dsqrt_rn((double)__fmaf_rn(d0, d0, __fmul_rn(d1, d1)))
                           : (float)__dsqrt_rn((double)__fadd_rn(__fmul_rn(d0, d0),
                                                                 __fmul_rn(d1, d1)));
    int ii = (int)rd_round<ARITH>(p1[0]);
    int jj = (int)rd_round<ARITH>(p1[1]);
    const float dx = (p2[0] > p1[0]) ? 1.0f : -1.0f;
    const float dy = (p2[1] > p1[1]) ? 1.0f : -1.0f;
    const double hx = 0.5 * (double)dx, hy = 0.5 * (double)dy;
    const double p1x = p1[0], p1y = p1[1], dd0 = d0, dd1 = d1;
    float alpha = 0.0f, alpha_old = 0.0f;
    float ax = (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)ii, hx), p1x), dd0);
    float ay = (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)jj, hy), p1y), dd1);
    const int cap = 4 * npix + 16;
    bool pend = false;  // an update (pv, pl) waits for the next step
    float pv = 0.0f, pl = 0.0f;
    for (int step = 0; alpha <= 1.0f && step < cap; ++step) {
        if (ax < ay) {
            alpha = ax;
            ii = (int)__fadd_rn((float)ii, dx);
            ax = (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)ii, hx), p1x), dd0);
        } else if (ax > ay) {
            alpha = ay;
            jj = (int)__fadd_rn((float)jj, dy);
            ay = (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)jj, hy), p1y), dd1);
        } else if (ax == ay) {
            alpha = ay;
            ii = (int)__fadd_rn((float)ii, dx);
            jj = (int)__fadd_rn((float)jj, dy);
            ax = (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)ii, hx), p1x), dd0);
            ay = (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)jj, hy), p1y), dd1);
        } else {
            break;  // NaN step parameter: the reference would never terminate
        }
        const float l = __fmul_rn(__fsub_rn(alpha, alpha_old), incr);
        const int i = ii + nh, j = jj + nh;
        const bool in = i > 0 && i < npix && j > 0 && j < npix;
        // this step's pixel is loaded now and added one step later, so the load is in
        // flight while the walk computes the next step; the updates keep their order
        const float v = (in && !ONES) ? img_in[(int64_t)i * npix + j] : 1.0f;
        if (pend) acc = rd_fwd_update<ARITH>(acc, pv, pl);
        pend = in;
        pv = v;
        pl = l;
        alpha_old = alpha;
    }
    if (pend) acc = rd_fwd_update<ARITH>(acc, pv, pl);
    return acc;
}

// Ray views whose neighbouring bins lie in neighbouring image rows (|sin| > |cos|: sub-ray
// i lies on y c - x s = tau(i), so one bin further moves a step of (-s, c) in (row,
// column)) record their forward entries against the transposed image, so that the 64
// lanes of a wave (64 neighbouring bins) gather from neighbouring addresses.
__device__ __forceinline__ bool rd_fwd_transposed(const RdView& g)
{
    return g.kind == RD_RAYS && fabsf(g.s) > fabsf(g.c);
}

// The cached system matrix of a plan (see rd_matrix_ensure), or all null.
struct RdMat {
    const int2* ent;
    const int64_t* off;
    const int* cnt;
};

// Replays one output's cached entries e[0], e[64], ... (n of them) into acc, in order.
// The chain of updates is serial (each rounds to float), so the latency of the entry
// loads and of the gathers they index is hidden by keeping kReplayDepth of each in
// flight: a batch's gathers and the next batch's entry loads are issued before the
// batch's updates.  Entries with the top bit set belong to a row-sum or column-sum view
// and are added plainly.
constexpr int kReplayDepth = 32;

template <int ARITH, bool ONES, bool BACK>
__device__ __forceinline__ float rd_replay(float acc, const int2* __restrict__ e, int n,
                                           const float* __restrict__ src)
{
    int2 q[kReplayDepth];
#pragma unroll
    for (int u = 0; u < kReplayDepth; ++u)
        q[u] = u < n ? e[(int64_t)u * 64] : make_int2(0, 0);
    for (int j = 0; j < n; j += kReplayDepth) {
        float v[kReplayDepth];
#pragma unroll
        for (int u = 0; u < kReplayDepth; ++u)
            v[u] = (ONES || j + u >= n) ? 1.0f : src[q[u].x & 0x7fffffff];
        int2 qn[kReplayDepth];
#pragma unroll
        for (int u = 0; u < kReplayDepth; ++u)
            qn[u] = j + kReplayDepth + u < n ? e[(int64_t)(j + kReplayDepth + u) * 64]
                                             : make_int2(0, 0);
#pragma unroll
        for (int u = 0; u < kReplayDepth; ++u) {
            if (j + u < n) {
                const float l = __int_as_float(q[u].y);
                if (q[u].x < 0)
                    acc = __fadd_rn(acc, v[u]);
                else if (BACK)
                    acc = rd_back_update<ARITH>(acc, v[u], l);
                else
                    acc = rd_fwd_update<ARITH>(acc, v[u], l);
            }
            q[u] = qn[u];
        }
    }
    return acc;
}

template <int ARITH, int EPI, bool ONES>
__global__ void __launch_bounds__(64)
rd_fwd_kernel(const RdView* __restrict__ geom, int npix, const float* __restrict__ image,
              const float* __restrict__ imageT, float* __restrict__ out, int64_t img_stride,
              int64_t sino_stride, int accumulate, EpiArgs e, RdMat mat)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = b < npix;
    const int v = blockIdx.y, z = blockIdx.z;
    const RdView g = geom[v];
    image += (int64_t)z * img_stride;
    const int64_t soff = (int64_t)z * sino_stride;
    const int64_t o = (int64_t)v * npix + b;
    double p0 = 0.0;
    if (active) {
        float acc = (EPI == EPI_STORE && accumulate) ? out[soff + o] : 0.0f;
        if (mat.ent) {
            // the bin's cached updates, in order (bitwise the walk's or the sum's), from
            // the image or its transpose as they were recorded
            const int64_t chunk = (int64_t)v * gridDim.x + blockIdx.x;
            const float* src = (!ONES && rd_fwd_transposed(g))
                                   ? imageT + (int64_t)z * img_stride : image;
            acc = rd_replay<ARITH, ONES, false>(acc, mat.ent + mat.off[chunk] + threadIdx.x,
                                                mat.cnt[(int64_t)v * npix + b], src);
        } else if (g.kind == RD_COLSUM) {
            const int i = npix - 1 - b;  // bin N-1-i <- column i, i >= 1
            if (i >= 1)
                for (int j = 0; j < npix; ++j)
                    acc = __fadd_rn(acc, ONES ? 1.0f : image[(int64_t)j * npix + i]);
        } else if (g.kind == RD_ROWSUM) {
            for (int j = 0; j < npix; ++j)
                acc = __fadd_rn(acc, ONES ? 1.0f : image[(int64_t)b * npix + j]);
        } else {
            for (int k = 1; k <= 5; ++k) {
                float p1[2], p2[2];
                rd_endpoints<ARITH>(6 * b + k, npix, g, p1, p2);
                acc = rd_walk<ARITH, ONES>(acc, image, npix, p1, p2);
            }
        }
        out[soff + o] = epi_forward<EPI>(acc, o, soff, z, e, &p0);
    }
    if (epi_has_partials(EPI) && e.partials) {
        const int blk = blockIdx.y * gridDim.x + blockIdx.x;
        block_partials3(p0, 0.0, 0.0,
                        e.partials + ((int64_t)z * e.blocks_per_slice + blk) * 3);
    }
}

// Crossing parameter of the walk's m-th boundary along one axis: the boundary between
// cells c0 + m*dc and c0 + (m+1)*dc, as the walk computes it (radon_rd.c:31-32).
__device__ __forceinline__ float rd_cross(int c0, int m, int dc, double h, float p, float d)
{
    return (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)(c0 + m * dc), h), (double)p),
                            (double)d);
}

// The step of one sub-ray's walk that enters cell (ii, jj), if the walk takes it:
// returns true and the credited length l.  The walk merges the x crossings X(0) < X(1)
// < ... with the y crossings Y(0) < Y(1) < ... (equal values make one diagonal step);
// x reaches ii after X(sx-1) and leaves it at X(sx), likewise for y, so the cell is
// entered at E = max(X(sx-1), Y(sy-1)) if E precedes both X(sx) and Y(sy).  The step
// before is the larger of the crossings preceding E; the walk takes the step while the
// one before it is <= 1 (radon_rd.c:28), and credits (E - previous) * |p2 - p1|.
template <int ARITH>
__device__ __forceinline__ bool rd_cell_step(const float p1[2], const float p2[2], int ii, int jj,
                                             float* l_out)
{
    const float d0 = __fsub_rn(p2[0], p1[0]), d1 = __fsub_rn(p2[1], p1[1]);
    const int i0 = (int)rd_round<ARITH>(p1[0]);
    const int j0 = (int)rd_round<ARITH>(p1[1]);
    const int dx = (p2[0] > p1[0]) ? 1 : -1, dy = (p2[1] > p1[1]) ? 1 : -1;
    const int sx = (ii - i0) * dx, sy = (jj - j0) * dy;
    if (sx < 0 || sy < 0 || (sx == 0 && sy == 0)) return false;  // behind, or the start cell
    const double hx = 0.5 * (double)dx, hy = 0.5 * (double)dy;
    const float ninf = __int_as_float(0xff800000);
    const float Ex = sx > 0 ? rd_cross(i0, sx - 1, dx, hx, p1[0], d0) : ninf;
    const float Ey = sy > 0 ? rd_cross(j0, sy - 1, dy, hy, p1[1], d1) : ninf;
    const float E = fmaxf(Ex, Ey);
    if (!(E < rd_cross(i0, sx, dx, hx, p1[0], d0))) return false;
    if (!(E < rd_cross(j0, sy, dy, hy, p1[1], d1))) return false;
    const float Xp = (Ex >= Ey && sx >= 2) ? rd_cross(i0, sx - 2, dx, hx, p1[0], d0) : ninf;
    const float Yp = (Ey >= Ex && sy >= 2) ? rd_cross(j0, sy - 2, dy, hy, p1[1], d1) : ninf;
    const float prev = Ex > Ey ? fmaxf(Xp, Ey) : (Ey > Ex ? fmaxf(Yp, Ex) : fmaxf(Xp, Yp));
    const float aold = (prev == ninf) ? 0.0f : prev;
    if (!(aold <= 1.0f)) return false;
    const float incr = (ARITH == ITOMO_ARITH_FAST)
                           ? (float)__dsqrt_rn((double)__fmaf_rn(d0, d0, __fmul_rn(d1, d1)))
                           : (float)__dsqrt_rn((double)__fadd_rn(__fmul_rn(d0, d0),
                                                                 __fmul_rn(d1, d1)));
    *l_out = __fmul_rn(__fsub_rn(E, aold), incr);
    return true;
}

// Back projection as a gather: one thread per pixel (16 x 16 per block), all views in
// order.  Sub-ray i lies on the line y*c - x*s = tau(i) with tau(i) = nh - i/6, except,
// for odd npix, sub-rays 6nh+1 and 6nh+2, whose t >= 0 below 3*npix keeps its sign
// (radon_rd.c:125-128) and puts them on tau = i/6 - nh.  A pixel's cell meets the lines
// with |tau - tau_c| <= (|c| + |s|)/2, tau_c = jj*c - ii*s; the candidates are those
// sub-rays plus one more on each side (a sixth of a pixel, far above the walk's
// rounding), and rd_cell_step decides each exactly.
template <int ARITH, bool ONES>
__global__ void __launch_bounds__(256)
rd_back_gather_kernel(const RdView* __restrict__ geom, int nang, int npix,
                      const float* __restrict__ sino, float* __restrict__ image,
                      int64_t sino_stride, int64_t img_stride, int accumulate)
{
    const int cc = blockIdx.x * 16 + (threadIdx.x & 15);  // image column (y cell)
    const int r = blockIdx.y * 16 + (threadIdx.x >> 4);   // image row (x cell)
    if (r >= npix || cc >= npix) return;
    image += (int64_t)blockIdx.z * img_stride;
    sino += (int64_t)blockIdx.z * sino_stride;
    const int64_t o = (int64_t)r * npix + cc;
    float pix = accumulate ? image[o] : 0.0f;
    const int nh = (int)(npix * 0.5);
    const int ii = r - nh, jj = cc - nh;
    const bool inner = r >= 1 && cc >= 1;
    for (int v = 0; v < nang; ++v) {
        const RdView g = geom[v];
        const float* row = sino + (int64_t)v * npix;
        if (g.kind == RD_COLSUM) {
            if (cc >= 1) pix = __fadd_rn(pix, ONES ? 1.0f : row[npix - 1 - cc]);
        } else if (g.kind == RD_ROWSUM) {
            pix = __fadd_rn(pix, ONES ? 1.0f : row[r]);
        } else if (inner) {
            const double sd = (double)g.s, cd = (double)g.c;
            const double tc = (double)jj * cd - (double)ii * sd;
            const double hw = 0.5 * (fabs(cd) + fabs(sd));
            int ilo = (int)floor(6.0 * ((double)nh - tc - hw)) - 1;
            int ihi = (int)ceil(6.0 * ((double)nh - tc + hw)) + 1;
            if ((npix & 1) && tc - hw - 1.0 / 3.0 <= 2.0 / 6.0 && tc + hw + 1.0 / 3.0 >= 1.0 / 6.0) {
                ilo = min(ilo, 6 * nh + 1);
                ihi = max(ihi, 6 * nh + 2);
            }
            ilo = max(ilo, 1);
            ihi = min(ihi, 6 * npix - 1);
            for (int i = ilo; i <= ihi; ++i) {
                if (i % 6 == 0) continue;
                float p1[2], p2[2], l;
                rd_endpoints<ARITH>(i, npix, g, p1, p2);
                if (!rd_cell_step<ARITH>(p1, p2, ii, jj, &l)) continue;
                pix = rd_back_update<ARITH>(pix, ONES ? 1.0f : row[i / 6], l);
            }
        }
    }
    image[o] = pix;
}

// ---- the cached system matrix ------------------------------------------------------
//
// With 288 GB of HBM per GPU the exact system matrix of a small ray-driven geometry fits
// many times over (8 bytes per walk step, once per direction), and replaying it turns
// both projections into streaming passes with full occupancy instead of latency-bound
// walks.  The entries are what the walks compute: per bin the steps of its five sub-rays
// that enter an image cell, in walk order, {pixel, credited length} (the pixel's index in
// the transposed image for views whose bins step along image rows; a row-sum or
// column-sum view's bin: its pixels {pixel | 1 << 31}, in order); per pixel, in view
// order, the row-sum and column-sum views' bins {bin | 1 << 31} and every sub-ray step
// entering it {bin, length}.  Replaying them in that order with the same update
// arithmetic gives the walks' results bit for bit.

// An entry of a row-sum or column-sum view: its input index with the top bit set.
__device__ __forceinline__ int rd_sum_flag(int index)
{
    return (int)((unsigned)index | 0x80000000u);
}

// One sub-ray walk of radon_rd.c:15-71 (the steps of rd_walk) recording the steps that
// enter an image cell: FILL == false counts them, FILL == true writes them to out[64 j].
template <int ARITH, bool FILL>
__device__ int rd_walk_record(int npix, const float p1[2], const float p2[2], int2* out, int j,
                              bool transposed)
{
    const int nh = (int)(npix * 0.5);
    const float d0 = __fsub_rn(p2[0], p1[0]), d1 = __fsub_rn(p2[1], p1[1]);
    const float incr = (ARITH == ITOMO_ARITH_FAST)
                           ? (float)__dsqrt_rn((double)__fmaf_rn(d0, d0, __fmul_rn(d1, d1)))
                           : (float)__dsqrt_rn((double)__fadd_rn(__fmul_rn(d0, d0),
                                                                 __fmul_rn(d1, d1)));
    int ii = (int)rd_round<ARITH>(p1[0]);
    int jj = (int)rd_round<ARITH>(p1[1]);
    const float dx = (p2[0] > p1[0]) ? 1.0f : -1.0f;
    const float dy = (p2[1] > p1[1]) ? 1.0f : -1.0f;
    const double hx = 0.5 * (double)dx, hy = 0.5 * (double)dy;
    const double p1x = p1[0], p1y = p1[1], dd0 = d0, dd1 = d1;
    float alpha = 0.0f, alpha_old = 0.0f;
    float ax = (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)ii, hx), p1x), dd0);
    float ay = (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)jj, hy), p1y), dd1);
    const int cap = 4 * npix + 16;
    for (int step = 0; alpha <= 1.0f && step < cap; ++step) {
        if (ax < ay) {
            alpha = ax;
            ii = (int)__fadd_rn((float)ii, dx);
            ax = (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)ii, hx), p1x), dd0);
        } else if (ax > ay) {
            alpha = ay;
            jj = (int)__fadd_rn((float)jj, dy);
            ay = (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)jj, hy), p1y), dd1);
        } else if (ax == ay) {
            alpha = ay;
            ii = (int)__fadd_rn((float)ii, dx);
            jj = (int)__fadd_rn((float)jj, dy);
            ax = (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)ii, hx), p1x), dd0);
            ay = (float)__ddiv_rn(__dsub_rn(__dadd_rn((double)jj, hy), p1y), dd1);
        } else {
            break;
        }
        const float l = __fmul_rn(__fsub_rn(alpha, alpha_old), incr);
        const int i = ii + nh, jc = jj + nh;
        if (i > 0 && i < npix && jc > 0 && jc < npix) {
            if (FILL)
                out[(int64_t)j * 64] =
                    make_int2(transposed ? jc * npix + i : i * npix + jc, __float_as_int(l));
            ++j;
        }
        alpha_old = alpha;
    }
    return j;
}

// Forward rows: one thread per (bin, view), as rd_fwd_kernel: the sum views' pixels in the
// order radon_rd.c:95-115 adds them, else the five sub-rays' steps.
template <int ARITH, bool FILL>
__global__ void __launch_bounds__(64)
rd_fwd_record_kernel(const RdView* __restrict__ geom, int npix, int* __restrict__ cnt,
                     int2* __restrict__ ent, const int64_t* __restrict__ off)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y;
    if (b >= npix) return;
    const RdView g = geom[v];
    int n = 0;
    int2* out = FILL ? ent + off[(int64_t)v * gridDim.x + blockIdx.x] + threadIdx.x : nullptr;
    if (g.kind == RD_COLSUM) {
        const int i = npix - 1 - b;   // bin N-1-i <- column i, i >= 1
        if (i >= 1)
            for (int j = 0; j < npix; ++j, ++n)
                if (FILL) out[(int64_t)n * 64] = make_int2(rd_sum_flag(j * npix + i), 0);
    } else if (g.kind == RD_ROWSUM) {
        for (int j = 0; j < npix; ++j, ++n)
            if (FILL) out[(int64_t)n * 64] = make_int2(rd_sum_flag(b * npix + j), 0);
    } else {
        for (int k = 1; k <= 5; ++k) {
            float p1[2], p2[2];
            rd_endpoints<ARITH>(6 * b + k, npix, g, p1, p2);
            n = rd_walk_record<ARITH, FILL>(npix, p1, p2, out, n, rd_fwd_transposed(g));
        }
    }
    if (!FILL) cnt[(int64_t)v * npix + b] = n;
}

// Back columns: one thread per pixel, the candidates of rd_back_gather_kernel.
template <int ARITH, bool FILL>
__global__ void __launch_bounds__(256)
rd_back_record_kernel(const RdView* __restrict__ geom, int nang, int npix,
                      int* __restrict__ cnt, int2* __restrict__ ent,
                      const int64_t* __restrict__ off)
{
    const int cc = blockIdx.x * 16 + (threadIdx.x & 15);
    const int r = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (r >= npix || cc >= npix) return;
    const int64_t o = (int64_t)r * npix + cc;
    int2* out = FILL ? ent + off[o >> 6] + (o & 63) : nullptr;
    const int nh = (int)(npix * 0.5);
    const int ii = r - nh, jj = cc - nh;
    const bool inner = r >= 1 && cc >= 1;
    int n = 0;
    for (int v = 0; v < nang; ++v) {
        const RdView g = geom[v];
        const int rowbase = v * npix;
        if (g.kind == RD_COLSUM) {
            if (cc >= 1) {
                if (FILL) out[(int64_t)n * 64] = make_int2(rd_sum_flag(rowbase + npix - 1 - cc), 0);
                ++n;
            }
        } else if (g.kind == RD_ROWSUM) {
            if (FILL) out[(int64_t)n * 64] = make_int2(rd_sum_flag(rowbase + r), 0);
            ++n;
        } else if (inner) {
            const double sd = (double)g.s, cd = (double)g.c;
            const double tc = (double)jj * cd - (double)ii * sd;
            const double hw = 0.5 * (fabs(cd) + fabs(sd));
            int ilo = (int)floor(6.0 * ((double)nh - tc - hw)) - 1;
            int ihi = (int)ceil(6.0 * ((double)nh - tc + hw)) + 1;
            if ((npix & 1) && tc - hw - 1.0 / 3.0 <= 2.0 / 6.0 && tc + hw + 1.0 / 3.0 >= 1.0 / 6.0) {
                ilo = min(ilo, 6 * nh + 1);
                ihi = max(ihi, 6 * nh + 2);
            }
            ilo = max(ilo, 1);
            ihi = min(ihi, 6 * npix - 1);
            for (int i = ilo; i <= ihi; ++i) {
                if (i % 6 == 0) continue;
                float p1[2], p2[2], l;
                rd_endpoints<ARITH>(i, npix, g, p1, p2);
                if (!rd_cell_step<ARITH>(p1, p2, ii, jj, &l)) continue;
                if (FILL) out[(int64_t)n * 64] = make_int2(rowbase + i / 6, __float_as_int(l));
                ++n;
            }
        }
    }
    if (!FILL) cnt[o] = n;
}

// Back projection from the cached columns: one thread per pixel, updates in view order
// with rd_back_gather_kernel's arithmetic.
template <int ARITH, bool ONES>
__global__ void __launch_bounds__(256)
rd_back_matrix_kernel(const int2* __restrict__ ent, const int64_t* __restrict__ off,
                      const int* __restrict__ cnt, int npix, const float* __restrict__ sino,
                      float* __restrict__ image, int64_t sino_stride, int64_t img_stride,
                      int accumulate)
{
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t n2 = (int64_t)npix * npix;
    if (o >= n2) return;
    image += (int64_t)blockIdx.y * img_stride;
    sino += (int64_t)blockIdx.y * sino_stride;
    const float pix = accumulate ? image[o] : 0.0f;
    image[o] = rd_replay<ARITH, ONES, true>(pix, ent + off[o >> 6] + (o & 63), cnt[o], sino);
}

namespace {

// Chunk offsets of a matrix from its per-output counts: outputs are grouped 64 at a time
// within rows of `row` outputs, and a chunk holds 64 x (its largest count) entries.
// Returns the number of entries.
int64_t chunk_layout(const std::vector<int>& cnt, int64_t nout, int64_t chunks_per_row,
                     int64_t row, std::vector<int64_t>* off)
{
    const int64_t nrows = nout / row;
    off->assign((size_t)(nrows * chunks_per_row), 0);
    int64_t total = 0;
    for (int64_t rr = 0; rr < nrows; ++rr)
        for (int64_t c = 0; c < chunks_per_row; ++c) {
            int m = 0;
            for (int64_t k = c * 64; k < std::min<int64_t>(row, c * 64 + 64); ++k)
                m = std::max(m, cnt[(size_t)(rr * row + k)]);
            (*off)[(size_t)(rr * chunks_per_row + c)] = total;
            total += (int64_t)m * 64;
        }
    return total;
}

}  // namespace

void rd_matrix_free(const Plan& p)
{
    void* ptrs[] = {p.rdf_ent, p.rdf_off, p.rdf_cnt, p.rdb_ent, p.rdb_off, p.rdb_cnt};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    p.rdf_ent = p.rdb_ent = nullptr;
    p.rdf_off = p.rdb_off = nullptr;
    p.rdf_cnt = p.rdb_cnt = nullptr;
    p.rd_mat_bytes = 0;
}

// hipMalloc that, when the device is out of memory, leaves no error behind: the matrix is
// then declined and the walks stay in use.
template <typename T>
static bool try_alloc(T** ptr, size_t bytes)
{
    if (hipMalloc((void**)ptr, bytes) == hipSuccess) return true;
    (void)hipGetLastError();
    *ptr = nullptr;
    return false;
}

template <int ARITH>
static int rd_matrix_build(const Plan& p, hipStream_t st)
{
    const int n = p.npix, V = p.nang;
    const int64_t n2 = (int64_t)n * n, nf = (int64_t)V * n;
    const int fcpr = (n + 63) / 64;                       // forward chunks per view
    const RdView* g = (const RdView*)p.d_geom;
    if (!try_alloc(&p.rdf_cnt, nf * sizeof(int)) || !try_alloc(&p.rdb_cnt, n2 * sizeof(int)))
        return ITOMO_OK;                                  // declined: the walks stay
    int rc;
    rd_fwd_record_kernel<ARITH, false><<<dim3(fcpr, V), 64, 0, st>>>(g, n, p.rdf_cnt, nullptr,
                                                                       nullptr);
    rd_back_record_kernel<ARITH, false><<<dim3((n + 15) / 16, (n + 15) / 16), 256, 0, st>>>(
        g, V, n, p.rdb_cnt, nullptr, nullptr);
    rc = check_hip(hipGetLastError(), "rd matrix: count launch");
    std::vector<int> cf((size_t)nf), cb((size_t)n2);
    if (!rc) rc = check_hip(hipMemcpyAsync(cf.data(), p.rdf_cnt, nf * sizeof(int),
                                           hipMemcpyDeviceToHost, st), "rd matrix: counts");
    if (!rc) rc = check_hip(hipMemcpyAsync(cb.data(), p.rdb_cnt, n2 * sizeof(int),
                                           hipMemcpyDeviceToHost, st), "rd matrix: counts");
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "rd matrix: sync");
    if (rc) return rc;
    std::vector<int64_t> of, ob;
    const int64_t tf = chunk_layout(cf, nf, fcpr, n, &of);
    // back: pixels in row-major chunks of 64 (one "row" of all n2 outputs)
    const int64_t tb = chunk_layout(cb, n2, (n2 + 63) / 64, n2, &ob);
    const size_t bytes = (size_t)(tf + tb) * sizeof(int2);
    if (bytes > kRdMatrixMaxBytes) return ITOMO_OK;      // declined: the walks stay
    if (!try_alloc(&p.rdf_ent, std::max<int64_t>(tf, 1) * sizeof(int2)) ||
        !try_alloc(&p.rdb_ent, std::max<int64_t>(tb, 1) * sizeof(int2)) ||
        !try_alloc(&p.rdf_off, of.size() * sizeof(int64_t)) ||
        !try_alloc(&p.rdb_off, ob.size() * sizeof(int64_t)))
        return ITOMO_OK;                                  // declined: the walks stay
    rc = check_hip(hipMemcpy(p.rdf_off, of.data(), of.size() * sizeof(int64_t),
                             hipMemcpyHostToDevice), "rd matrix");
    if (!rc) rc = check_hip(hipMemcpy(p.rdb_off, ob.data(), ob.size() * sizeof(int64_t),
                                      hipMemcpyHostToDevice), "rd matrix");
    if (rc) return rc;
    rd_fwd_record_kernel<ARITH, true><<<dim3(fcpr, V), 64, 0, st>>>(g, n, nullptr, p.rdf_ent,
                                                                      p.rdf_off);
    rd_back_record_kernel<ARITH, true><<<dim3((n + 15) / 16, (n + 15) / 16), 256, 0, st>>>(
        g, V, n, nullptr, p.rdb_ent, p.rdb_off);
    rc = check_hip(hipGetLastError(), "rd matrix: fill launch");
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "rd matrix: sync");
    if (rc) return rc;
    p.rd_mat_bytes = bytes;
    p.rd_mat_state = 1;
    return ITOMO_OK;
}

// A launch recorded into a stream capture keeps the matrix's addresses in the graph: from
// then on the matrix stays until the plan is destroyed (itomo_plan_rd_matrix refuses to
// drop it).
void rd_matrix_note_use(const Plan& p, hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
        p.rd_mat_pinned = 1;
}

int rd_matrix_ensure(const Plan& p, hipStream_t st)
{
    if (p.projector != ITOMO_PROJ_RD || p.rd_mat_state != 0 || scratch_region_active())
        return ITOMO_OK;
    // no synchronising build under stream capture: the walks are captured instead
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    int rc = check_hip(hipStreamIsCapturing(st, &cs), "rd matrix: capture status");
    if (rc || cs != hipStreamCaptureStatusNone) return rc;
    p.rd_mat_state = -1;
    // upper bound of the entries (a walk takes at most 2 npix + 2 steps): skip the counting
    // passes when even a fraction of it cannot fit
    const double bound = 16.0 * 5.0 * p.nang * (double)p.npix * (1.5 * p.npix);
    if (bound > 4.0 * (double)kRdMatrixMaxBytes) return ITOMO_OK;
    // entry indices are 31-bit (the top bit flags a sum view's bin)
    if ((int64_t)p.nang * p.npix >= ((int64_t)1 << 31)) return ITOMO_OK;
    rc = check_hip(hipStreamSynchronize(st), "rd matrix: sync");
    if (!rc) rc = p.arith == ITOMO_ARITH_FAST ? rd_matrix_build<ITOMO_ARITH_FAST>(p, st)
                                              : rd_matrix_build<ITOMO_ARITH_IEEE>(p, st);
    if (p.rd_mat_state != 1) rd_matrix_free(p);
    return rc;
}

namespace {

template <int ARITH, int EPI, bool ONES>
int launch_rd_fwd(const Plan& p, const float* image, float* out, int nslices,
                  int accumulate, const EpiArgs& e, hipStream_t st)
{
    dim3 block(64), grid((p.npix + 63) / 64, p.nang, nslices);
    const float* imageT = image;
    if (p.rd_mat_state == 1) {
        rd_matrix_note_use(p, st);
        if (!ONES) {   // the transposed copy the recorded entries of some views index
            Scratch s;
            int rc = get_scratch(p, nslices, &s, st);
            if (!rc) rc = launch_transpose(image, s.xT, p.npix, nslices, st);
            if (rc) return rc;
            imageT = s.xT;
        }
    }
    const RdMat mat = p.rd_mat_state == 1
                          ? RdMat{p.rdf_ent, p.rdf_off, p.rdf_cnt}
                          : RdMat{nullptr, nullptr, nullptr};
    rd_fwd_kernel<ARITH, EPI, ONES><<<grid, block, 0, st>>>(
        (const RdView*)p.d_geom, p.npix, image, imageT, out, (int64_t)p.npix * p.npix,
        (int64_t)p.nang * p.npix, accumulate, e, mat);
    return ITOMO_OK;
}

template <int ARITH>
int dispatch_rd_fwd(const Plan& p, int epi, int ones, const float* image, float* out,
                    int nslices, int accumulate, const EpiArgs& e, hipStream_t st)
{
    switch (epi) {
    case EPI_STORE:
        if (ones) return launch_rd_fwd<ARITH, EPI_STORE, true>(p, image, out, nslices, accumulate, e, st);
        return launch_rd_fwd<ARITH, EPI_STORE, false>(p, image, out, nslices, accumulate, e, st);
    case EPI_RECIP:
        return launch_rd_fwd<ARITH, EPI_RECIP, true>(p, image, out, nslices, 0, e, st);
    case EPI_SIRT_FWD:
        return launch_rd_fwd<ARITH, EPI_SIRT_FWD, false>(p, image, out, nslices, 0, e, st);
    case EPI_MLEM_FWD:
        return launch_rd_fwd<ARITH, EPI_MLEM_FWD, false>(p, image, out, nslices, 0, e, st);
    case EPI_SPS_FWD:
        return launch_rd_fwd<ARITH, EPI_SPS_FWD, false>(p, image, out, nslices, 0, e, st);
    }
    return set_error(ITOMO_EINVAL, "rd forward: unsupported epilogue %d", epi);
}

}  // namespace

int rd_forward_epi(const Plan& p, int epi, int ones, const float* image, float* out,
                   int nslices, int accumulate, const EpiArgs& e, hipStream_t st,
                   int* blocks_per_slice)
{
    const int bps = ((p.npix + 63) / 64) * p.nang;
    if (blocks_per_slice) *blocks_per_slice = bps;
    EpiArgs ee = e;
    ee.blocks_per_slice = bps;
    int rc = p.arith == ITOMO_ARITH_FAST
                 ? dispatch_rd_fwd<ITOMO_ARITH_FAST>(p, epi, ones, image, out, nslices, accumulate, ee, st)
                 : dispatch_rd_fwd<ITOMO_ARITH_IEEE>(p, epi, ones, image, out, nslices, accumulate, ee, st);
    if (rc) return rc;
    return check_hip(hipGetLastError(), "rd forward launch");
}

// Back projection into `image` (overwritten, or accumulated into).
static int rd_back_launch(const Plan& p, int ones, const float* sino, float* image, int nslices,
                          int64_t sino_stride, int64_t img_stride, int accumulate,
                          hipStream_t st)
{
    if (p.rd_mat_state == 1) {
        rd_matrix_note_use(p, st);
        const int64_t n2 = (int64_t)p.npix * p.npix;
        dim3 mg((unsigned)((n2 + 255) / 256), nslices);
#define ITOMO_RDM(A, O)                                                                  \
    rd_back_matrix_kernel<A, O><<<mg, 256, 0, st>>>(p.rdb_ent, p.rdb_off, p.rdb_cnt, p.npix, \
                                                   sino, image, sino_stride, img_stride,  \
                                                   accumulate)
        if (p.arith == ITOMO_ARITH_FAST) {
            if (ones) ITOMO_RDM(ITOMO_ARITH_FAST, true); else ITOMO_RDM(ITOMO_ARITH_FAST, false);
        } else {
            if (ones) ITOMO_RDM(ITOMO_ARITH_IEEE, true); else ITOMO_RDM(ITOMO_ARITH_IEEE, false);
        }
#undef ITOMO_RDM
        return check_hip(hipGetLastError(), "rd back (matrix) launch");
    }
    dim3 block(256), grid((p.npix + 15) / 16, (p.npix + 15) / 16, nslices);
    const RdView* g = (const RdView*)p.d_geom;
    if (p.arith == ITOMO_ARITH_FAST) {
        if (ones)
            rd_back_gather_kernel<ITOMO_ARITH_FAST, true><<<grid, block, 0, st>>>(
                g, p.nang, p.npix, sino, image, sino_stride, img_stride, accumulate);
        else
            rd_back_gather_kernel<ITOMO_ARITH_FAST, false><<<grid, block, 0, st>>>(
                g, p.nang, p.npix, sino, image, sino_stride, img_stride, accumulate);
    } else {
        if (ones)
            rd_back_gather_kernel<ITOMO_ARITH_IEEE, true><<<grid, block, 0, st>>>(
                g, p.nang, p.npix, sino, image, sino_stride, img_stride, accumulate);
        else
            rd_back_gather_kernel<ITOMO_ARITH_IEEE, false><<<grid, block, 0, st>>>(
                g, p.nang, p.npix, sino, image, sino_stride, img_stride, accumulate);
    }
    return check_hip(hipGetLastError(), "rd back launch");
}

int rd_back_overwrite(const Plan& p, int ones, const float* sino, float* image, int nslices,
                    hipStream_t st)
{
    return rd_back_launch(p, ones, sino, image, nslices, (int64_t)p.nang * p.npix,
                          (int64_t)p.npix * p.npix, 0, st);
}

int rd_forward(const Plan& p, const float* image, float* sino, int nslices, int64_t,
               int64_t, int accumulate, hipStream_t st)
{
    EpiArgs e{};
    return rd_forward_epi(p, EPI_STORE, 0, image, sino, nslices, accumulate, e, st, nullptr);
}

int rd_back(const Plan& p, const float* sino, float* image, int nslices, int64_t sino_stride,
            int64_t img_stride, int accumulate, hipStream_t st)
{
    return rd_back_launch(p, 0, sino, image, nslices, sino_stride, img_stride, accumulate, st);
}

}  // namespace itomo

extern "C" int itomo_plan_rd_matrix(itomo_plan* plan, int mode, size_t* bytes, void* stream)
{
    if (!plan || plan->projector != ITOMO_PROJ_RD || mode < ITOMO_RDM_QUERY ||
        mode > ITOMO_RDM_DROP)
        return itomo::set_error(ITOMO_EINVAL, "plan_rd_matrix: bad argument");
    int rc = ITOMO_OK;
    if (mode == ITOMO_RDM_BUILD) {
        if (itomo::scratch_region_active())
            return itomo::set_error(ITOMO_EINVAL, "plan_rd_matrix: inside a no-allocation call");
        if (plan->rd_mat_state == -1) plan->rd_mat_state = 0;
        rc = itomo::rd_matrix_ensure(*plan, (hipStream_t)stream);
    } else if (mode == ITOMO_RDM_DROP) {
        if (plan->rd_mat_pinned)
            return itomo::set_error(ITOMO_EINVAL, "plan_rd_matrix: a captured graph uses the "
                                                  "matrix; it stays until the plan is destroyed");
        rc = itomo::check_hip(hipStreamSynchronize((hipStream_t)stream), "plan_rd_matrix: sync");
        itomo::rd_matrix_free(*plan);
        plan->rd_mat_state = -1;
    }
    if (bytes) *bytes = plan->rd_mat_state == 1 ? plan->rd_mat_bytes : 0;
    return rc;
}
