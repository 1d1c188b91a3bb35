// Ray-driven projector pair (semantics of radon_rd.c:15-189) for gfx950.
//
// Five sub-rays per detector bin (sub-ray i = 6b+1..6b+5 of bin b, radon_rd.c:123-185),
// each clipped to the square [-nh, nh]^2 and walked cell by cell; the length walked
// since the previous step is credited, with weight 1/5, to the cell just entered, and
// cells with row or column index 0 are skipped (radon_rd.c:44-66).  The walk repeats
// the reference build's arithmetic operation for operation (float geometry, double
// step parameters, double fused update rounded to float after every step), so the
// forward projection is bit-identical to the reference's.
//
// Forward: one thread per (bin, view) walks its five sub-rays in order: no atomics.
// Back: one thread per (bin, view) walks the same sub-rays and applies each cell update
// with a compare-and-swap loop that performs exactly the reference's rounding
// (float(old + 0.2*sino*len), evaluated in double); only the order in which updates of
// one pixel arrive differs from the reference, so results agree to float rounding.
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

__device__ __forceinline__ void cas_update_fast(float* addr, float sv, float l, int arith)
{
    unsigned int* a = reinterpret_cast<unsigned int*>(addr);
    unsigned int old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        const float o = __uint_as_float(old);
        const float n = arith == ITOMO_ARITH_FAST ? rd_back_update<ITOMO_ARITH_FAST>(o, sv, l)
                                                  : rd_back_update<ITOMO_ARITH_IEEE>(o, sv, l);
        const unsigned int prev = atomicCAS(a, old, __float_as_uint(n));
        if (prev == old) break;
        old = prev;
    }
}

__device__ __forceinline__ void cas_add(float* addr, float v)
{
    unsigned int* a = reinterpret_cast<unsigned int*>(addr);
    unsigned int old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        const float n = __fadd_rn(__uint_as_float(old), v);
        const unsigned int prev = atomicCAS(a, old, __float_as_uint(n));
        if (prev == old) break;
        old = prev;
    }
}

// One sub-ray walk (radon_rd.c:15-71).  OPER 0 returns the updated bin value; OPER 1
// scatters into the image and returns acc unchanged.  The walk is capped at
// 4*npix+16 steps (a genuine walk takes at most 2*npix+2), so it always ends.
template <int ARITH, int OPER, bool ONES>
__device__ float rd_walk(float acc, const float* __restrict__ img_in, float* img_out,
                         int npix, const float p1[2], const float p2[2])
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
    const float sv = acc;  // OPER 1: the bin value being back-projected
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
            break;  // NaN step parameter: the reference would never terminate
        }
        const float l = __fmul_rn(__fsub_rn(alpha, alpha_old), incr);
        const int i = ii + nh, j = jj + nh;
        if (i > 0 && i < npix && j > 0 && j < npix) {
            const int64_t k = (int64_t)i * npix + j;
            if (OPER == 0)
                acc = rd_fwd_update<ARITH>(acc, ONES ? 1.0f : img_in[k], l);
            else
                cas_update_fast(img_out + k, sv, l, ARITH);
        }
        alpha_old = alpha;
    }
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
                acc = rd_walk<ARITH, 0, ONES>(acc, image, nullptr, npix, p1, p2);
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

template <int ARITH, bool ONES>
__global__ void __launch_bounds__(64)
rd_back_kernel(const RdView* __restrict__ geom, int npix, const float* __restrict__ sino,
               float* __restrict__ image, int64_t sino_stride, int64_t img_stride)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= npix) return;
    const int v = blockIdx.y;
    const RdView g = geom[v];
    image += (int64_t)blockIdx.z * img_stride;
    sino += (int64_t)blockIdx.z * sino_stride;
    const float sv = ONES ? 1.0f : sino[(int64_t)v * npix + b];
    if (g.kind == RD_COLSUM) {
        const int i = npix - 1 - b;
        if (i >= 1)
            for (int j = 0; j < npix; ++j) cas_add(image + (int64_t)j * npix + i, sv);
    } else if (g.kind == RD_ROWSUM) {
        for (int j = 0; j < npix; ++j) cas_add(image + (int64_t)b * npix + j, sv);
    } else {
        for (int k = 1; k <= 5; ++k) {
            float p1[2], p2[2];
            rd_endpoints<ARITH>(6 * b + k, npix, g, p1, p2);
            (void)rd_walk<ARITH, 1, false>(sv, nullptr, image, npix, p1, p2);
        }
    }
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

// Scatter back projection into `image` (zeroed first).
int rd_back_scatter(const Plan& p, int ones, const float* sino, float* image, int nslices,
                    hipStream_t st)
{
    int rc = fill(image, 0.0f, (int64_t)p.npix * p.npix * nslices, st);
    if (rc) return rc;
    dim3 block(64), grid((p.npix + 63) / 64, p.nang, nslices);
    const RdView* g = (const RdView*)p.d_geom;
    const int64_t ss = (int64_t)p.nang * p.npix, is = (int64_t)p.npix * p.npix;
    if (p.arith == ITOMO_ARITH_FAST) {
        if (ones) rd_back_kernel<ITOMO_ARITH_FAST, true><<<grid, block, 0, st>>>(g, p.npix, sino, image, ss, is);
        else rd_back_kernel<ITOMO_ARITH_FAST, false><<<grid, block, 0, st>>>(g, p.npix, sino, image, ss, is);
    } else {
        if (ones) rd_back_kernel<ITOMO_ARITH_IEEE, true><<<grid, block, 0, st>>>(g, p.npix, sino, image, ss, is);
        else rd_back_kernel<ITOMO_ARITH_IEEE, false><<<grid, block, 0, st>>>(g, p.npix, sino, image, ss, is);
    }
    return check_hip(hipGetLastError(), "rd back launch");
}

int rd_forward(const Plan& p, const float* image, float* sino, int nslices, int64_t,
               int64_t, int accumulate, hipStream_t st)
{
    EpiArgs e{};
    return rd_forward_epi(p, EPI_STORE, 0, image, sino, nslices, accumulate, e, st, nullptr);
}

// Back projection: scatter straight into the output when accumulating, else zero it first.
int rd_back(const Plan& p, const float* sino, float* image, int nslices, int64_t sino_stride,
            int64_t img_stride, int accumulate, hipStream_t st)
{
    if (!accumulate) return rd_back_scatter(p, 0, sino, image, nslices, st);
    dim3 block(64), grid((p.npix + 63) / 64, p.nang, nslices);
    const RdView* g = (const RdView*)p.d_geom;
    if (p.arith == ITOMO_ARITH_FAST)
        rd_back_kernel<ITOMO_ARITH_FAST, false><<<grid, block, 0, st>>>(g, p.npix, sino, image,
                                                                        sino_stride, img_stride);
    else
        rd_back_kernel<ITOMO_ARITH_IEEE, false><<<grid, block, 0, st>>>(g, p.npix, sino, image,
                                                                        sino_stride, img_stride);
    return check_hip(hipGetLastError(), "rd back launch");
}

}  // namespace itomo
