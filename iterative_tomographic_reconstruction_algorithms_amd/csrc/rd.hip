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

template <int ARITH, int EPI, bool ONES>
__global__ void __launch_bounds__(64)
rd_fwd_kernel(const RdView* __restrict__ geom, int npix, const float* __restrict__ image,
              float* __restrict__ out, int64_t img_stride, int64_t sino_stride,
              int accumulate, EpiArgs e)
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
        if (g.kind == RD_COLSUM) {
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

namespace {

template <int ARITH, int EPI, bool ONES>
void launch_rd_fwd(const Plan& p, const float* image, float* out, int nslices,
                   int accumulate, const EpiArgs& e, hipStream_t st)
{
    dim3 block(64), grid((p.npix + 63) / 64, p.nang, nslices);
    rd_fwd_kernel<ARITH, EPI, ONES><<<grid, block, 0, st>>>(
        (const RdView*)p.d_geom, p.npix, image, out, (int64_t)p.npix * p.npix,
        (int64_t)p.nang * p.npix, accumulate, e);
}

template <int ARITH>
int dispatch_rd_fwd(const Plan& p, int epi, int ones, const float* image, float* out,
                    int nslices, int accumulate, const EpiArgs& e, hipStream_t st)
{
    switch (epi) {
    case EPI_STORE:
        if (ones) launch_rd_fwd<ARITH, EPI_STORE, true>(p, image, out, nslices, accumulate, e, st);
        else launch_rd_fwd<ARITH, EPI_STORE, false>(p, image, out, nslices, accumulate, e, st);
        return ITOMO_OK;
    case EPI_RECIP:
        launch_rd_fwd<ARITH, EPI_RECIP, true>(p, image, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_SIRT_FWD:
        launch_rd_fwd<ARITH, EPI_SIRT_FWD, false>(p, image, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_MLEM_FWD:
        launch_rd_fwd<ARITH, EPI_MLEM_FWD, false>(p, image, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_SPS_FWD:
        launch_rd_fwd<ARITH, EPI_SPS_FWD, false>(p, image, out, nslices, 0, e, st);
        return ITOMO_OK;
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
