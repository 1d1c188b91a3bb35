// Pixel-driven projector pair (semantics of radon_pd.c:20-90) for gfx950.
//
// The reference walks every pixel (i, j) of every view and sends four sub-samples of
// the pixel to the detector: sample k sits at (x0 + d_k, y0 + d_k) with d_k = -1/4 for
// k = 0, 1 and +1/4 for k = 2, 3 (radon_pd.c:42-43 index delta[2k+1] for both
// coordinates, so only two points are distinct).  Its signed distance t from the
// central ray picks the bins: method 0 the nearest one, round(nh - 0.5 - t), method 1
// the two neighbours floor/ceil of nh - t with linear weights.  Every contribution is
// added to its bin (forward) or pixel (back) as float(old + w * v / 4) evaluated in
// double; w * v / 4 of two floats is exact in double, so the reference build's fused
// form and the C as written round identically.
//
// Both directions here are gathers, so each output element receives its terms in the
// reference's order and the results are bit-identical to the reference's, with no
// atomics:
//   forward: one thread per (view, bin) walks the image rows i upwards and, in each row,
//            the columns whose pixel centre lies within 1 + |(|c| - sign(c) s)|/4 +
//            margin of the bin's line (t(x0, y0) = |c| y0 - sign(c) s x0 to float
//            rounding, a bound fixed in double per row), evaluating every candidate
//            sample exactly as the reference does and keeping those that land in the bin;
//   back:    one thread per pixel walks the views in order.
// Method 0 has no range check in the reference (radon_pd.c:58-63): a sample whose bin
// l lies outside [0, npix) lands at flat index v*npix + l, i.e. in a neighbouring
// sinogram row.  That is reproduced (the forward thread of row r also gathers the
// overflow of views r-1 and r+1); flat indices outside the array, where the reference
// writes out of bounds, are skipped.
#include "epilogue.h"

namespace itomo {

constexpr int kPdFwdBlock = 256;
constexpr int kPdBackX = 64, kPdBackY = 4;

// Signed distance of sample (x, y) from the central ray (radon_pd.c:44-55).
template <int ARITH>
__device__ __forceinline__ float pd_t(const PdView& g, float x, float y)
{
    const float xs = __fmul_rn(x, g.s);
    float t, q;
    if (ARITH == ITOMO_ARITH_FAST) {  // fused difference, x*s/c as x*s*(1/c)
        t = fabsf(__fmaf_rn(-g.c, y, xs));
        q = __fmul_rn(xs, g.rc);
    } else {
        t = fabsf(__fsub_rn(xs, __fmul_rn(y, g.c)));
        q = __fdiv_rn(xs, g.c);
    }
    const bool keep = ((g.flags & (kPdLt | kPdGt)) && y > q) ||
                      ((g.flags & kPdZero) && y > 0.0f) || ((g.flags & kPdHalf) && x < 0.0f);
    return keep ? t : -t;
}

// Method 0 bin: round(nh - 0.5 - t) (radon_pd.c:58).  nh - 0.5 - t is exact in double.
template <int ARITH>
__device__ __forceinline__ int pd_nearest(int nh, float t)
{
    const double v = __dsub_rn((double)nh - 0.5, (double)t);
    if (ARITH == ITOMO_ARITH_FAST)  // gcc -ffast-math inline expansion of round()
        return (int)trunc(__dadd_rn(v, copysign(0.49999999999999994, v)));
    return (int)round(v);
}

// Method 1 split (radon_pd.c:67-71): the bins are round(lf - 0.5) and round(uf - 0.5),
// which for an integer lf >= 1 is lf itself and for lf <= 0 is lf - 1 < 0 (rejected by
// the reference's 0 < l < npix test), so a bin is valid iff 1 <= lf <= npix - 1 and is
// then lf (likewise uf).
struct PdSplit {
    float lf, uf, wl, wu;
};

template <int ARITH>
__device__ __forceinline__ PdSplit pd_split(float nhf, float t)
{
    const float t1 = __fsub_rn(nhf, t);
    PdSplit s;
    s.lf = floorf(t1);
    s.uf = ceilf(t1);
    if (ARITH == ITOMO_ARITH_FAST)  // |uf - (nh - t)| as the fast build forms it
        s.wl = fabsf(__fadd_rn(__fsub_rn(t, nhf), s.uf));
    else
        s.wl = fabsf(__fsub_rn(s.uf, t1));
    s.wu = fabsf(__fsub_rn(t1, s.lf));
    return s;
}

// float(acc + w * v / 4) with the sum rounded once in double.
__device__ __forceinline__ float pd_acc(float acc, float w, float v)
{
    return (float)__fma_rn((double)w, __dmul_rn((double)v, 0.25), (double)acc);
}

// Contribution of pixel value px at centre (x0, y0) to bin e (e >= 1 for method 1; any
// integer for method 0, e = the bin index before the flat-index overflow).
template <int ARITH, int METHOD>
__device__ __forceinline__ float pd_fwd_pixel(const PdView& g, int nh, float nhf, float x0,
                                              float y0, int e, float px, float acc)
{
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float d = h ? 0.25f : -0.25f;
        const float t = pd_t<ARITH>(g, __fadd_rn(d, x0), __fadd_rn(d, y0));
        if (METHOD == 0) {
            if (pd_nearest<ARITH>(nh, t) == e) {  // samples 2h and 2h+1 coincide
                acc = pd_acc(acc, 1.0f, px);
                acc = pd_acc(acc, 1.0f, px);
            }
        } else {
            const PdSplit s = pd_split<ARITH>(nhf, t);
            const float ef = (float)e;
            const bool hl = s.lf == ef, hu = s.uf == ef;
            // lf == uf only for an integer t1, where both weights are 0: one zero-weight
            // term then stands for the reference's two (x + 0 = x; a non-finite pixel
            // gives NaN either way)
            const float w = hl ? s.wl : s.wu;
            if (hl || hu) {
                acc = pd_acc(acc, w, px);
                acc = pd_acc(acc, w, px);
            }
        }
    }
    return acc;
}

// Sum of one view's samples that land in bin e, in the reference's (i, j, k) order.
template <int ARITH, int METHOD, bool ONES>
__device__ float pd_fwd_bin(const PdView& g, int npix, const float* __restrict__ image,
                            int e, float acc)
{
    const int nh = npix >> 1;
    const float nhf = (float)nh;
    // The samples of a pixel sit at t(x0, y0) +- dshift; float rounding of t, of the sign
    // test and of nh - t stays far below npix * 2^-18.
    const double half = 1.0 + g.dshift + 0.01 + npix * 0x1p-18;
    const double mid = (double)(nh - e);
    const double tl = mid - half, th = mid + half;
    const double C = g.cabs, S = g.sdir;
    const double xa = 0.5 - nh, xb = (double)(npix - 1 - nh) + 0.5;
    const double smin = fmin(xa * S, xb * S), smax = fmax(xa * S, xb * S);
    // rows whose centre line can meet the band: C*y0 in (tl + smin, th + smax); the
    // margin is inside the band, so the integer rounding can be tight
    int i0 = 0, i1 = npix - 1;
    if (C > 0.0) {
        const double ylo = (tl + smin) / C + (nh - 0.5), yhi = (th + smax) / C + (nh - 0.5);
        i0 = (int)fmin(fmax(0.0, ceil(ylo)), (double)npix);
        i1 = (int)fmax(-1.0, fmin((double)(npix - 1), floor(yhi)));
    }
    // column range of row i: x0 between (C*y0 - th)/S and (C*y0 - tl)/S, in column units
    // j = x0 + nh - 1/2: j_lo/hi = a*y0 + b_lo/hi
    const double rs = g.rsdir, a = C * rs, off = nh - 0.5;
    const double b_lo = (rs > 0.0 ? -th : -tl) * rs + off;
    const double b_hi = (rs > 0.0 ? -tl : -th) * rs + off;
    for (int i = i0; i <= i1; ++i) {
        const float y0 = __fadd_rn((float)(i - nh), 0.5f);
        int j0, j1;
        if (rs != 0.0) {
            j0 = (int)fmin(fmax(0.0, ceil(fma((double)y0, a, b_lo))), (double)npix);
            j1 = (int)fmax(-1.0, fmin((double)(npix - 1), floor(fma((double)y0, a, b_hi))));
        } else {
            const double yc = (double)y0 * C;
            const bool in = yc > tl && yc < th;
            j0 = in ? 0 : 1;
            j1 = in ? npix - 1 : 0;
        }
        const float* __restrict__ row = image + (int64_t)(npix - 1 - i) * npix;
        for (int j = j0; j <= j1; ++j) {
            const float px = ONES ? 1.0f : row[j];
            const float x0 = __fadd_rn((float)(j - nh), 0.5f);
            acc = pd_fwd_pixel<ARITH, METHOD>(g, nh, nhf, x0, y0, e, px, acc);
        }
    }
    return acc;
}

// One thread per (sinogram row r, bin b).  Epilogues: STORE, RECIP.
template <int ARITH, int METHOD, int EPI, bool ONES>
__global__ void __launch_bounds__(kPdFwdBlock)
pd_fwd_kernel(const PdView* __restrict__ geom, int npix, int nang,
              const float* __restrict__ image, float* __restrict__ out, int64_t img_stride,
              int64_t sino_stride, int accumulate, EpiArgs e)
{
    const int b = blockIdx.x * kPdFwdBlock + threadIdx.x;
    if (b >= npix) return;
    const int r = blockIdx.y, z = blockIdx.z;
    image += (int64_t)z * img_stride;
    const int64_t soff = (int64_t)z * sino_stride, o = (int64_t)r * npix + b;
    float acc = (EPI == EPI_STORE && accumulate) ? out[soff + o] : 0.0f;
    if (METHOD == 0) {
        // view r-1 overflows into row r at bins l = b + npix, view r+1 at l = b - npix
        for (int vs = r - 1; vs <= r + 1; ++vs)
            if (vs >= 0 && vs < nang)
                acc = pd_fwd_bin<ARITH, 0, ONES>(geom[vs], npix, image, b + (r - vs) * npix,
                                                 acc);
    } else if (b > 0) {  // bin 0 never receives anything (radon_pd.c:74)
        acc = pd_fwd_bin<ARITH, 1, ONES>(geom[r], npix, image, b, acc);
    }
    double p0 = 0.0;
    out[soff + o] = epi_forward<EPI>(acc, o, soff, z, e, &p0);
}

// One thread per pixel; the block is 64 columns x 4 rows of the image.
template <int ARITH, int METHOD, int EPI, bool ONES>
__global__ void __launch_bounds__(kPdBackX * kPdBackY)
pd_back_kernel(const PdView* __restrict__ geom, int npix, int nang,
               const float* __restrict__ sino, float* __restrict__ out, int64_t sino_stride,
               int64_t img_stride, int accumulate, EpiArgs e)
{
    const int j = blockIdx.x * kPdBackX + threadIdx.x;
    const int m = blockIdx.y * kPdBackY + threadIdx.y;  // memory row = npix-1-i
    const int z = blockIdx.z;
    sino += (int64_t)z * sino_stride;
    const int64_t ioff = (int64_t)z * img_stride, o = (int64_t)m * npix + j;
    double p1 = 0.0, p2 = 0.0;
    if (j < npix && m < npix) {
        const int nh = npix >> 1;
        const float nhf = (float)nh, npixf = (float)npix;
        const float x0 = __fadd_rn((float)(j - nh), 0.5f);
        const float y0 = __fadd_rn((float)(npix - 1 - m - nh), 0.5f);
        const int64_t ntot = (int64_t)nang * npix;
        float acc = (EPI == EPI_STORE && accumulate) ? out[ioff + o] : 0.0f;
        for (int v = 0; v < nang; ++v) {
            const PdView g = geom[v];
            const float* __restrict__ rowp = sino + (int64_t)v * npix;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float d = h ? 0.25f : -0.25f;
                const float t = pd_t<ARITH>(g, __fadd_rn(d, x0), __fadd_rn(d, y0));
                if (METHOD == 0) {
                    const int64_t f = (int64_t)v * npix + pd_nearest<ARITH>(nh, t);
                    if (f >= 0 && f < ntot) {
                        const float sv = ONES ? 1.0f : sino[f];
                        acc = pd_acc(acc, 1.0f, sv);
                        acc = pd_acc(acc, 1.0f, sv);
                    }
                } else {
                    const PdSplit s = pd_split<ARITH>(nhf, t);
                    const bool okl = s.lf >= 1.0f && s.lf < npixf;
                    const bool oku = s.uf >= 1.0f && s.uf < npixf;
                    const float sl = okl ? (ONES ? 1.0f : rowp[(int)s.lf]) : 0.0f;
                    const float su = oku ? (ONES ? 1.0f : rowp[(int)s.uf]) : 0.0f;
#pragma unroll
                    for (int rep = 0; rep < 2; ++rep) {
                        if (okl) acc = pd_acc(acc, s.wl, sl);
                        if (oku) acc = pd_acc(acc, s.wu, su);
                    }
                }
            }
        }
        const float xo = epi_updates_x(EPI) ? out[ioff + o] : 0.0f;
        out[ioff + o] = epi_back<EPI>(acc, xo, o, z, e, &p1, &p2);
    }
    if (epi_has_partials(EPI) && e.partials) {
        const int blk = blockIdx.y * gridDim.x + blockIdx.x;
        block_partials3(0.0, p1, p2, e.partials + ((int64_t)z * e.blocks_per_slice + blk) * 3);
    }
}

namespace {

template <int ARITH, int METHOD, int EPI, bool ONES>
void launch_pd_fwd(const Plan& p, const float* image, float* out, int nslices, int accumulate,
                   const EpiArgs& e, hipStream_t st)
{
    dim3 grid((p.npix + kPdFwdBlock - 1) / kPdFwdBlock, p.nang, nslices);
    pd_fwd_kernel<ARITH, METHOD, EPI, ONES><<<grid, kPdFwdBlock, 0, st>>>(
        (const PdView*)p.d_geom, p.npix, p.nang, image, out, (int64_t)p.npix * p.npix,
        (int64_t)p.nang * p.npix, accumulate, e);
}

template <int ARITH, int METHOD>
int dispatch_pd_fwd(const Plan& p, int epi, int ones, const float* image, float* out,
                    int nslices, int accumulate, const EpiArgs& e, hipStream_t st)
{
    if (epi == EPI_STORE) {
        if (ones) launch_pd_fwd<ARITH, METHOD, EPI_STORE, true>(p, image, out, nslices, accumulate, e, st);
        else launch_pd_fwd<ARITH, METHOD, EPI_STORE, false>(p, image, out, nslices, accumulate, e, st);
        return ITOMO_OK;
    }
    if (epi == EPI_RECIP && ones) {
        launch_pd_fwd<ARITH, METHOD, EPI_RECIP, true>(p, image, out, nslices, 0, e, st);
        return ITOMO_OK;
    }
    return set_error(ITOMO_EINVAL, "pd forward: unsupported epilogue %d", epi);
}

template <int ARITH, int METHOD, int EPI, bool ONES>
void launch_pd_back(const Plan& p, const float* sino, float* out, int nslices, int accumulate,
                    const EpiArgs& e, hipStream_t st)
{
    dim3 block(kPdBackX, kPdBackY);
    dim3 grid((p.npix + kPdBackX - 1) / kPdBackX, (p.npix + kPdBackY - 1) / kPdBackY, nslices);
    pd_back_kernel<ARITH, METHOD, EPI, ONES><<<grid, block, 0, st>>>(
        (const PdView*)p.d_geom, p.npix, p.nang, sino, out, (int64_t)p.nang * p.npix,
        (int64_t)p.npix * p.npix, accumulate, e);
}

template <int ARITH, int METHOD>
int dispatch_pd_back(const Plan& p, int epi, int ones, const float* sino, float* out,
                     int nslices, int accumulate, const EpiArgs& e, hipStream_t st)
{
    switch (epi) {
    case EPI_STORE:
        if (ones) launch_pd_back<ARITH, METHOD, EPI_STORE, true>(p, sino, out, nslices, accumulate, e, st);
        else launch_pd_back<ARITH, METHOD, EPI_STORE, false>(p, sino, out, nslices, accumulate, e, st);
        return ITOMO_OK;
    case EPI_RECIP:
        if (!ones) break;
        launch_pd_back<ARITH, METHOD, EPI_RECIP, true>(p, sino, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_SIRT_BACK:
        launch_pd_back<ARITH, METHOD, EPI_SIRT_BACK, false>(p, sino, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_MLEM_BACK:
        launch_pd_back<ARITH, METHOD, EPI_MLEM_BACK, false>(p, sino, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_SPS_BACK:
        launch_pd_back<ARITH, METHOD, EPI_SPS_BACK, false>(p, sino, out, nslices, 0, e, st);
        return ITOMO_OK;
    }
    return set_error(ITOMO_EINVAL, "pd back: unsupported epilogue %d", epi);
}

}  // namespace

int pd_forward_epi(const Plan& p, int epi, int ones, const float* image, float* out,
                   int nslices, int accumulate, const EpiArgs& e, hipStream_t st,
                   int* blocks_per_slice)
{
    if (epi == EPI_SIRT_FWD || epi == EPI_SPS_FWD || epi == EPI_MLEM_FWD) {
        // the projection, then the epilogue as an elementwise pass over the sinogram
        int rc = pd_forward_epi(p, EPI_STORE, ones, image, out, nslices, 0, EpiArgs{}, st,
                                nullptr);
        if (rc) return rc;
        return elementwise_fwd_epilogue(epi, out, e, (int64_t)p.nang * p.npix, nslices, st,
                                        blocks_per_slice);
    }
    if (blocks_per_slice) *blocks_per_slice = 0;
    int rc;
    if (p.arith == ITOMO_ARITH_FAST)
        rc = p.method == 0
                 ? dispatch_pd_fwd<ITOMO_ARITH_FAST, 0>(p, epi, ones, image, out, nslices, accumulate, e, st)
                 : dispatch_pd_fwd<ITOMO_ARITH_FAST, 1>(p, epi, ones, image, out, nslices, accumulate, e, st);
    else
        rc = p.method == 0
                 ? dispatch_pd_fwd<ITOMO_ARITH_IEEE, 0>(p, epi, ones, image, out, nslices, accumulate, e, st)
                 : dispatch_pd_fwd<ITOMO_ARITH_IEEE, 1>(p, epi, ones, image, out, nslices, accumulate, e, st);
    if (rc) return rc;
    return check_hip(hipGetLastError(), "pd forward launch");
}

int pd_back_epi(const Plan& p, int epi, int ones, const float* sino, float* out, int nslices,
                int accumulate, const EpiArgs& e, hipStream_t st, int* blocks_per_slice)
{
    const int bps = (int)back_blocks_per_slice(p);
    if (blocks_per_slice) *blocks_per_slice = bps;
    EpiArgs ee = e;
    ee.blocks_per_slice = bps;
    int rc;
    if (p.arith == ITOMO_ARITH_FAST)
        rc = p.method == 0
                 ? dispatch_pd_back<ITOMO_ARITH_FAST, 0>(p, epi, ones, sino, out, nslices, accumulate, ee, st)
                 : dispatch_pd_back<ITOMO_ARITH_FAST, 1>(p, epi, ones, sino, out, nslices, accumulate, ee, st);
    else
        rc = p.method == 0
                 ? dispatch_pd_back<ITOMO_ARITH_IEEE, 0>(p, epi, ones, sino, out, nslices, accumulate, ee, st)
                 : dispatch_pd_back<ITOMO_ARITH_IEEE, 1>(p, epi, ones, sino, out, nslices, accumulate, ee, st);
    if (rc) return rc;
    return check_hip(hipGetLastError(), "pd back launch");
}

int pd_forward(const Plan& p, const float* image, float* sino, int nslices, int64_t, int64_t,
               int accumulate, hipStream_t st)
{
    return pd_forward_epi(p, EPI_STORE, 0, image, sino, nslices, accumulate, EpiArgs{}, st,
                          nullptr);
}

int pd_back(const Plan& p, const float* sino, float* image, int nslices, int64_t, int64_t,
            int accumulate, hipStream_t st)
{
    return pd_back_epi(p, EPI_STORE, 0, sino, image, nslices, accumulate, EpiArgs{}, st,
                       nullptr);
}

}  // namespace itomo
