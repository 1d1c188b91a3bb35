// Slant-stacking projector pair (semantics of radon_ss.c:43-196) for gfx950.
//
// Bin k = t + nh of view v is the line x c + y s = t; it goes to sinogram row nang-1-v.
// Views with |s| <= sqrt(1/2) step the image row u (y) and sample the column
// q = (t - u s)/c; the others step the column u (x) and sample the row q = (t - u c)/s.
// Each sample adds 1/|c| (or 1/|s|) times the image at the nearest pixel (method 0) or at
// floor(q) and floor(q)+1 with weights 1-w and w (method 1).  The stepped range of every
// bin (the reference's ray_limits and rounded loop bounds, radon_ss.c:43-88,121,159) is
// tabulated per plan on the host, clipped to the image: samples outside the square are
// skipped (the reference indexes outside the image rows there).
//
// Arithmetic is the reference's, operation for operation (ARITH FAST: its -ffast-math
// build; IEEE: the C as written); every update rounds to float after a double sum.
//   forward: one thread per bin steps u in order -- the reference's own loop;
//   back:    one thread per pixel gathers, per view in order, the 2-4 bins whose line
//            can reach it, in increasing t: each (view, bin) reaches a pixel at most once,
//            so the pixel receives the reference's terms in the reference's order.
// No atomics; both directions are bit-identical to the reference wherever its samples
// stay inside the image.
#include "epilogue.h"

namespace itomo {

constexpr int kSsFwdBlock = 256;
constexpr int kSsBackX = 64, kSsBackY = 4;

// Sampled coordinate of stepped index u on line t (radon_ss.c:124,135,162,173).
template <int ARITH>
__device__ __forceinline__ float ss_q(const SsView& g, int u, float tf)
{
    if (ARITH == ITOMO_ARITH_FAST)  // (t - u a)/b as fma(-u, a, t) * (1/b)
        return __fmul_rn(__fmaf_rn(-(float)u, g.a, tf), g.rb);
    return __fsub_rn(__fdiv_rn(tf, g.b), __fdiv_rn(__fmul_rn((float)u, g.a), g.b));
}

template <int ARITH>
__device__ __forceinline__ int ss_round(float v)
{
    if (ARITH == ITOMO_ARITH_FAST)  // the fast build's float expansion of round()
        return (int)truncf(__fadd_rn(v, copysignf(0.49999997f, v)));
    return (int)round((double)v);
}

// Linear split of q: fl = floor(q), w = q - fl, wl = 1 - w as the build forms it.
template <int ARITH>
__device__ __forceinline__ void ss_split(float q, float& fl, float& w, float& wl)
{
    fl = floorf(q);
    w = __fsub_rn(q, fl);  // exact
    wl = (ARITH == ITOMO_ARITH_FAST) ? __fadd_rn(__fsub_rn(1.0f, q), fl) : __fsub_rn(1.0f, w);
}

// Forward update: bin + (w / |b|) * pix.
template <int ARITH>
__device__ __forceinline__ float ss_fwd_upd(float bin, float w, double rab, float pix)
{
    if (ARITH == ITOMO_ARITH_FAST)
        return (float)__fma_rn(__dmul_rn((double)w, rab), (double)pix, (double)bin);
    return (float)__dadd_rn((double)bin, __dmul_rn(__dmul_rn(rab, (double)w), (double)pix));
}

// Back update: pix + w * (bin / |b|).
template <int ARITH>
__device__ __forceinline__ float ss_back_upd(float pix, float w, double rab, float bin)
{
    if (ARITH == ITOMO_ARITH_FAST)
        return (float)__fma_rn((double)w, __dmul_rn((double)bin, rab), (double)pix);
    return (float)__dadd_rn((double)pix, __dmul_rn(__dmul_rn(rab, (double)w), (double)bin));
}

// Method-0 updates (weight 1): fma(v, 1/|b|, acc) in the fast build.
template <int ARITH>
__device__ __forceinline__ float ss_upd0(float acc, double rab, float v)
{
    if (ARITH == ITOMO_ARITH_FAST) return (float)__fma_rn((double)v, rab, (double)acc);
    return (float)__dadd_rn((double)acc, __dmul_rn(rab, (double)v));
}

// One thread per (sinogram row, bin).  Epilogues: STORE, RECIP.
template <int ARITH, int METHOD, int EPI, bool ONES>
__global__ void __launch_bounds__(kSsFwdBlock)
ss_fwd_kernel(const SsView* __restrict__ geom, const int2* __restrict__ range, int npix,
              int nang, const float* __restrict__ image, float* __restrict__ out,
              int64_t img_stride, int64_t sino_stride, int accumulate, EpiArgs e)
{
    const int k = blockIdx.x * kSsFwdBlock + threadIdx.x;
    if (k >= npix) return;
    const int r = blockIdx.y, z = blockIdx.z;
    const int v = nang - 1 - r;
    image += (int64_t)z * img_stride;
    const int64_t soff = (int64_t)z * sino_stride, o = (int64_t)r * npix + k;
    float acc = (EPI == EPI_STORE && accumulate) ? out[soff + o] : 0.0f;
    const int nh = npix >> 1;
    if (k < 2 * nh) {  // odd npix: the last bin has no line (radon_ss.c:109)
        const SsView g = geom[v];
        const int2 ur = range[(int64_t)v * npix + k];
        const float tf = (float)(k - nh);
        // stepped index u -> image (stepped, sampled) strides
        const int64_t ss = g.ybranch ? npix : 1, sq = g.ybranch ? 1 : npix;
        for (int u = ur.x; u < ur.y; ++u) {
            const float q = ss_q<ARITH>(g, u, tf);
            const float* __restrict__ line = image + (int64_t)(u + nh) * ss;
            if (METHOD == 0) {
                const int p = ss_round<ARITH>(q) + nh;
                if (p >= 0 && p < npix)
                    acc = ss_upd0<ARITH>(acc, g.rab, ONES ? 1.0f : line[p * sq]);
            } else {
                float fl, w, wl;
                ss_split<ARITH>(q, fl, w, wl);
                const int p = (int)fl + nh;
                if (p >= 0 && p <= npix - 1)
                    acc = ss_fwd_upd<ARITH>(acc, wl, g.rab, ONES ? 1.0f : line[p * sq]);
                if (p + 1 >= 0 && p + 1 <= npix - 1)
                    acc = ss_fwd_upd<ARITH>(acc, w, g.rab, ONES ? 1.0f : line[(p + 1) * sq]);
            }
        }
    }
    double p0 = 0.0;
    out[soff + o] = epi_forward<EPI>(acc, o, soff, z, e, &p0);
}

// One thread per pixel (memory row R, column Q); 64 columns x 4 rows per block.
template <int ARITH, int METHOD, int EPI, bool ONES>
__global__ void __launch_bounds__(kSsBackX * kSsBackY)
ss_back_kernel(const SsView* __restrict__ geom, const int2* __restrict__ range, int npix,
               int nang, const float* __restrict__ sino, float* __restrict__ out,
               int64_t sino_stride, int64_t img_stride, int accumulate, EpiArgs e)
{
    const int Q = blockIdx.x * kSsBackX + threadIdx.x;
    const int R = blockIdx.y * kSsBackY + threadIdx.y;
    const int z = blockIdx.z;
    sino += (int64_t)z * sino_stride;
    const int64_t ioff = (int64_t)z * img_stride, o = (int64_t)R * npix + Q;
    double p1 = 0.0, p2 = 0.0;
    if (Q < npix && R < npix) {
        const int nh = npix >> 1;
        // bins whose line passes within |b| * (1 + margin) of the pixel centre in the
        // sampled coordinate: t = u a + q b with q in [P - 1, P + 1]
        const double margin = 0.01 + npix * 0x1p-16;
        float acc = (EPI == EPI_STORE && accumulate) ? out[ioff + o] : 0.0f;
        for (int v = 0; v < nang; ++v) {
            const SsView g = geom[v];
            const int u = (g.ybranch ? R : Q) - nh;  // stepped
            const int P = (g.ybranch ? Q : R) - nh;  // sampled
            const double tc = (double)u * g.a + (double)P * g.b;
            const double hw = fabs((double)g.b) * (1.0 + margin) + margin;
            const int t0 = (int)fmax((double)-nh, ceil(tc - hw));
            const int t1 = (int)fmin((double)(nh - 1), floor(tc + hw));
            const float* __restrict__ row = sino + (int64_t)(nang - 1 - v) * npix;
            const int2* __restrict__ rg = range + (int64_t)v * npix;
            for (int t = t0; t <= t1; ++t) {
                const int k = t + nh;
                const int2 ur = rg[k];
                if (u < ur.x || u >= ur.y) continue;
                const float q = ss_q<ARITH>(g, u, (float)t);
                if (METHOD == 0) {
                    if (ss_round<ARITH>(q) == P)
                        acc = ss_upd0<ARITH>(acc, g.rab, ONES ? 1.0f : row[k]);
                } else {
                    float fl, w, wl;
                    ss_split<ARITH>(q, fl, w, wl);
                    const int p = (int)fl;
                    if (p == P)
                        acc = ss_back_upd<ARITH>(acc, wl, g.rab, ONES ? 1.0f : row[k]);
                    else if (p + 1 == P)
                        acc = ss_back_upd<ARITH>(acc, w, g.rab, ONES ? 1.0f : row[k]);
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
void launch_ss_fwd(const Plan& p, const float* image, float* out, int nslices, int accumulate,
                   const EpiArgs& e, hipStream_t st)
{
    dim3 grid((p.npix + kSsFwdBlock - 1) / kSsFwdBlock, p.nang, nslices);
    ss_fwd_kernel<ARITH, METHOD, EPI, ONES><<<grid, kSsFwdBlock, 0, st>>>(
        (const SsView*)p.d_geom, (const int2*)p.d_range, p.npix, p.nang, image, out,
        (int64_t)p.npix * p.npix, (int64_t)p.nang * p.npix, accumulate, e);
}

template <int ARITH, int METHOD>
int dispatch_ss_fwd(const Plan& p, int epi, int ones, const float* image, float* out,
                    int nslices, int accumulate, const EpiArgs& e, hipStream_t st)
{
    if (epi == EPI_STORE) {
        if (ones) launch_ss_fwd<ARITH, METHOD, EPI_STORE, true>(p, image, out, nslices, accumulate, e, st);
        else launch_ss_fwd<ARITH, METHOD, EPI_STORE, false>(p, image, out, nslices, accumulate, e, st);
        return ITOMO_OK;
    }
    if (epi == EPI_RECIP && ones) {
        launch_ss_fwd<ARITH, METHOD, EPI_RECIP, true>(p, image, out, nslices, 0, e, st);
        return ITOMO_OK;
    }
    return set_error(ITOMO_EINVAL, "ss forward: unsupported epilogue %d", epi);
}

template <int ARITH, int METHOD, int EPI, bool ONES>
void launch_ss_back(const Plan& p, const float* sino, float* out, int nslices, int accumulate,
                    const EpiArgs& e, hipStream_t st)
{
    dim3 block(kSsBackX, kSsBackY);
    dim3 grid((p.npix + kSsBackX - 1) / kSsBackX, (p.npix + kSsBackY - 1) / kSsBackY, nslices);
    ss_back_kernel<ARITH, METHOD, EPI, ONES><<<grid, block, 0, st>>>(
        (const SsView*)p.d_geom, (const int2*)p.d_range, p.npix, p.nang, sino, out,
        (int64_t)p.nang * p.npix, (int64_t)p.npix * p.npix, accumulate, e);
}

template <int ARITH, int METHOD>
int dispatch_ss_back(const Plan& p, int epi, int ones, const float* sino, float* out,
                     int nslices, int accumulate, const EpiArgs& e, hipStream_t st)
{
    switch (epi) {
    case EPI_STORE:
        if (ones) launch_ss_back<ARITH, METHOD, EPI_STORE, true>(p, sino, out, nslices, accumulate, e, st);
        else launch_ss_back<ARITH, METHOD, EPI_STORE, false>(p, sino, out, nslices, accumulate, e, st);
        return ITOMO_OK;
    case EPI_RECIP:
        if (!ones) break;
        launch_ss_back<ARITH, METHOD, EPI_RECIP, true>(p, sino, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_SIRT_BACK:
        launch_ss_back<ARITH, METHOD, EPI_SIRT_BACK, false>(p, sino, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_MLEM_BACK:
        launch_ss_back<ARITH, METHOD, EPI_MLEM_BACK, false>(p, sino, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_SPS_BACK:
        launch_ss_back<ARITH, METHOD, EPI_SPS_BACK, false>(p, sino, out, nslices, 0, e, st);
        return ITOMO_OK;
    }
    return set_error(ITOMO_EINVAL, "ss back: unsupported epilogue %d", epi);
}

}  // namespace

int ss_forward_epi(const Plan& p, int epi, int ones, const float* image, float* out,
                   int nslices, int accumulate, const EpiArgs& e, hipStream_t st,
                   int* blocks_per_slice)
{
    if (epi == EPI_SIRT_FWD || epi == EPI_SPS_FWD || epi == EPI_MLEM_FWD) {
        int rc = ss_forward_epi(p, EPI_STORE, ones, image, out, nslices, 0, EpiArgs{}, st,
                                nullptr);
        if (rc) return rc;
        return elementwise_fwd_epilogue(epi, out, e, (int64_t)p.nang * p.npix, nslices, st,
                                        blocks_per_slice);
    }
    if (blocks_per_slice) *blocks_per_slice = 0;
    int rc;
    if (p.arith == ITOMO_ARITH_FAST)
        rc = p.method == 0
                 ? dispatch_ss_fwd<ITOMO_ARITH_FAST, 0>(p, epi, ones, image, out, nslices, accumulate, e, st)
                 : dispatch_ss_fwd<ITOMO_ARITH_FAST, 1>(p, epi, ones, image, out, nslices, accumulate, e, st);
    else
        rc = p.method == 0
                 ? dispatch_ss_fwd<ITOMO_ARITH_IEEE, 0>(p, epi, ones, image, out, nslices, accumulate, e, st)
                 : dispatch_ss_fwd<ITOMO_ARITH_IEEE, 1>(p, epi, ones, image, out, nslices, accumulate, e, st);
    if (rc) return rc;
    return check_hip(hipGetLastError(), "ss forward launch");
}

int ss_back_epi(const Plan& p, int epi, int ones, const float* sino, float* out, int nslices,
                int accumulate, const EpiArgs& e, hipStream_t st, int* blocks_per_slice)
{
    const int bps = (int)back_blocks_per_slice(p);
    if (blocks_per_slice) *blocks_per_slice = bps;
    EpiArgs ee = e;
    ee.blocks_per_slice = bps;
    int rc;
    if (p.arith == ITOMO_ARITH_FAST)
        rc = p.method == 0
                 ? dispatch_ss_back<ITOMO_ARITH_FAST, 0>(p, epi, ones, sino, out, nslices, accumulate, ee, st)
                 : dispatch_ss_back<ITOMO_ARITH_FAST, 1>(p, epi, ones, sino, out, nslices, accumulate, ee, st);
    else
        rc = p.method == 0
                 ? dispatch_ss_back<ITOMO_ARITH_IEEE, 0>(p, epi, ones, sino, out, nslices, accumulate, ee, st)
                 : dispatch_ss_back<ITOMO_ARITH_IEEE, 1>(p, epi, ones, sino, out, nslices, accumulate, ee, st);
    if (rc) return rc;
    return check_hip(hipGetLastError(), "ss back launch");
}

int ss_forward(const Plan& p, const float* image, float* sino, int nslices, int64_t, int64_t,
               int accumulate, hipStream_t st)
{
    return ss_forward_epi(p, EPI_STORE, 0, image, sino, nslices, accumulate, EpiArgs{}, st,
                          nullptr);
}

int ss_back(const Plan& p, const float* sino, float* image, int nslices, int64_t, int64_t,
            int accumulate, hipStream_t st)
{
    return ss_back_epi(p, EPI_STORE, 0, sino, image, nslices, accumulate, EpiArgs{}, st,
                       nullptr);
}

}  // namespace itomo
