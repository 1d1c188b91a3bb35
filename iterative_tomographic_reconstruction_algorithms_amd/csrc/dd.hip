// Distance-driven projector pair (semantics of radon_dd.c:54-285) for gfx950.
//
// The reference sorts, per view and image line, the N+1 pixel boundaries and the N+1
// detector-bin boundaries and sweeps the merged list; each segment between consecutive
// keys that lies inside pixel q and bin d moves weight |segment| between them, and the
// segment that ends at the first key of index N is dropped (radon_dd.c:176-177).
// Here the same weights are evaluated in gather form: w(q, d) = min(P[q+1], D[d+1], L*)
// - max(P[q], D[d]) when positive, where L* is the start of the line's dropped segment
// (one float per view and line, tabulated when the plan is made: dd_line_end).  Keys
// are formed with exactly the float operations of the reference build, so every weight
// is bit-identical to the reference's.
//
// Accumulation order is also the reference's, so results are bitwise reproducible:
//  * forward: one thread per (view, bin) walks image lines in increasing order and,
//    inside a line, pixels in increasing order (the sweep order), using fma like the
//    reference build (-ffast-math contracts `sino += diff * img`).
//  * back: one thread per pixel walks views in list order and, inside a view, bins in
//    increasing order.
// No atomics: every output element has exactly one writer.  The SIRT/MLEM update of
// the element is applied in the same thread right after its sum is complete (see
// epilogue.h), with FP64 partial sums reduced per block in a fixed order.
#include <vector>


#include "dd_device.h"

namespace itomo {

template <int ARITH, int EPI, bool ONES, int T, int P>
__global__ void __launch_bounds__((T / P) * (T / P))
dd_back_kernel(const DdView* __restrict__ geom, const float* __restrict__ lstar, int nang,
               int npix, const float* __restrict__ sino, float* __restrict__ out,
               int64_t sino_stride, int64_t img_stride, int accumulate, EpiArgs e)
{
    dd_back_tile<ARITH, EPI, ONES, T, P>(geom, lstar, nang, npix, sino, out, sino_stride, img_stride,
                                   accumulate, e, blockIdx.x, blockIdx.y, blockIdx.z,
                                   blockIdx.y * gridDim.x + blockIdx.x);
}

// views: the view of each blockIdx.y (null: blockIdx.y itself).  FP64 partial slots are
// one per 64-bin chunk, numbered by view (dd_fwd_slots), so launches over disjoint view
// lists, also with different B, fill one table.
template <int ARITH, int EPI, bool ONES, int B, int PATH>
__global__ void __launch_bounds__(B)
dd_fwd_kernel(const DdView* __restrict__ geom, const float* __restrict__ lstar, int nang,
              int npix, const float* __restrict__ image, const float* __restrict__ imageT,
              float* __restrict__ out, int64_t img_stride, int64_t sino_stride,
              int accumulate, EpiArgs e, const int* __restrict__ views)
{
    const int view = views ? views[blockIdx.y] : (int)blockIdx.y;
    dd_fwd_block<ARITH, EPI, ONES, B, PATH>(geom, lstar, nang, npix, image, imageT, out,
                                            img_stride, sino_stride, accumulate, e, blockIdx.x,
                                            view, blockIdx.z);
}

__global__ void __launch_bounds__(256)
transpose_kernel(const float* __restrict__ in, float* __restrict__ out, int npix,
                 int64_t stride)
{
    transpose_tile(in, out, npix, stride, blockIdx.x, blockIdx.y, blockIdx.z);
}

// out = the transposes of nslices npix x npix images (npix^2 apart).
int launch_transpose(const float* in, float* out, int npix, int nslices, hipStream_t st)
{
    dim3 tb(32, 8), tg((npix + 31) / 32, (npix + 31) / 32, nslices);
    transpose_kernel<<<tg, tb, 0, st>>>(in, out, npix, (int64_t)npix * npix);
    return check_hip(hipGetLastError(), "transpose launch");
}

// lstar[v * npix + i] = L* of line i in view v (0 for views without a sweep).
template <int ARITH>
__global__ void __launch_bounds__(256)
line_end_kernel(const DdView* __restrict__ geom, int nang, int npix, float* __restrict__ lstar)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nang * npix) return;
    const int v = (int)(k / npix), i = (int)(k - (int64_t)v * npix);
    const DdView g = geom[v];
    const bool sweep = g.kind == DD_SWEEP_ROWS || g.kind == DD_SWEEP_COLS;
    lstar[k] = sweep ? dd_line_end<ARITH>(g, npix, (int)(npix * 0.5), i) : 0.0f;
}

namespace {

template <int ARITH, int EPI, bool ONES, int B, int PATH>
void launch_fwd_path(const Plan& p, const float* image, const float* imageT, float* out,
                     int nslices, int accumulate, const EpiArgs& e, const int* views,
                     int nviews, hipStream_t st)
{
    dim3 block(B), grid((p.npix + B - 1) / B, nviews, nslices);
    dd_fwd_kernel<ARITH, EPI, ONES, B, PATH><<<grid, block, 0, st>>>(
        (const DdView*)p.d_geom, p.d_lstar, p.nang, p.npix, image, imageT, out,
        (int64_t)p.npix * p.npix, (int64_t)p.nang * p.npix, accumulate, e, views);
}

// Views with more than 8 candidate pixels per bin and line go to a launch of their own
// (see kPathMain); with unit input, one launch takes every view.  The wide launch is small
// (a tenth of the views at 2048^2 x 1800) and its blocks are long, so on its own it
// leaves most of the GPU idle: it runs on the plan's second stream beside the main launch
// (both write disjoint sinogram rows), forked from and joined back into `st`.  It always
// uses 64-bin blocks: a block walks every line any of its bins meets, and a wide view's
// bins meet few lines each (about npix / w of them for bins w pixels wide, against
// npix / w + B for a block), so wider blocks only add lines that weigh zero.
template <int ARITH, int EPI, bool ONES, int B>
int launch_fwd_b(const Plan& p, const float* image, const float* imageT, float* out,
                 int nslices, int accumulate, const EpiArgs& e, hipStream_t st)
{
    if (ONES || !p.d_vlist || p.n_gather == 0) {
        launch_fwd_path<ARITH, EPI, ONES, B, kPathAll>(p, image, imageT, out, nslices,
                                                       accumulate, e, nullptr, p.nang, st);
        return ITOMO_OK;
    }
    hipStream_t wide_st = st;
    if (p.n_staged > 0 && p.aux_stream) {
        int rc = check_hip(hipEventRecord(p.ev_fork, st), "dd forward: fork");
        if (!rc) rc = check_hip(hipStreamWaitEvent(p.aux_stream, p.ev_fork, 0), "dd forward: fork");
        if (rc) return rc;
        wide_st = p.aux_stream;
    }
    // the main launch is issued first: the wide launch's long blocks then take the slots
    // the main launch's blocks leave, rather than holding slots from its start
    if (p.n_staged > 0)
        launch_fwd_path<ARITH, EPI, ONES, B, kPathMain>(p, image, imageT, out, nslices,
                                                        accumulate, e, p.d_vlist, p.n_staged, st);
    launch_fwd_path<ARITH, EPI, ONES, 64, kPathWide>(p, image, imageT, out, nslices, accumulate,
                                                     e, p.d_vlist + p.n_staged, p.n_gather,
                                                     wide_st);
    if (wide_st != st) {
        int rc = check_hip(hipEventRecord(p.ev_join, wide_st), "dd forward: join");
        if (!rc) rc = check_hip(hipStreamWaitEvent(st, p.ev_join, 0), "dd forward: join");
        if (rc) return rc;
    }
    return ITOMO_OK;
}

template <int ARITH, int EPI, bool ONES>
int launch_fwd(const Plan& p, const float* image, const float* imageT, float* out,
               int nslices, int accumulate, const EpiArgs& e, hipStream_t st)
{
    static_assert(kFwdBins[0] == 256 && kFwdBins[1] == 128 && kFwdBins[2] == 64, "variants");
    switch (p.fwd_var) {
    case 1: return launch_fwd_b<ARITH, EPI, ONES, 128>(p, image, imageT, out, nslices, accumulate, e, st);
    case 2: return launch_fwd_b<ARITH, EPI, ONES, 64>(p, image, imageT, out, nslices, accumulate, e, st);
    default: return launch_fwd_b<ARITH, EPI, ONES, 256>(p, image, imageT, out, nslices, accumulate, e, st);
    }
}

template <int ARITH, int EPI, bool ONES, int T, int P>
void launch_back_tp(const Plan& p, const float* sino, float* out, int nslices, int accumulate,
                    const EpiArgs& e, hipStream_t st)
{
    dim3 block((T / P) * (T / P)), grid((p.npix + T - 1) / T, (p.npix + T - 1) / T, nslices);
    dd_back_kernel<ARITH, EPI, ONES, T, P><<<grid, block, 0, st>>>(
        (const DdView*)p.d_geom, p.d_lstar, p.nang, p.npix, sino, out,
        (int64_t)p.nang * p.npix, (int64_t)p.npix * p.npix, accumulate, e);
}

template <int ARITH, int EPI, bool ONES>
void launch_back(const Plan& p, const float* sino, float* out, int nslices, int accumulate,
                 const EpiArgs& e, hipStream_t st)
{
    static_assert(kBackTiles[0][0] == 64 && kBackTiles[0][1] == 4 && kBackTiles[1][0] == 32 &&
                      kBackTiles[1][1] == 2 && kBackTiles[2][0] == 32 && kBackTiles[2][1] == 4,
                  "variants");
    switch (p.back_var) {
    case 1: launch_back_tp<ARITH, EPI, ONES, 32, 2>(p, sino, out, nslices, accumulate, e, st); break;
    case 2: launch_back_tp<ARITH, EPI, ONES, 32, 4>(p, sino, out, nslices, accumulate, e, st); break;
    default: launch_back_tp<ARITH, EPI, ONES, 64, 4>(p, sino, out, nslices, accumulate, e, st); break;
    }
}

template <int ARITH>
int dispatch_fwd(const Plan& p, int epi, int ones, const float* image, const float* imageT,
                 float* out, int nslices, int accumulate, const EpiArgs& e, hipStream_t st)
{
    switch (epi) {
    case EPI_STORE:
        if (ones) return launch_fwd<ARITH, EPI_STORE, true>(p, image, imageT, out, nslices, accumulate, e, st);
        return launch_fwd<ARITH, EPI_STORE, false>(p, image, imageT, out, nslices, accumulate, e, st);
    case EPI_RECIP:
        return launch_fwd<ARITH, EPI_RECIP, true>(p, image, imageT, out, nslices, 0, e, st);
    case EPI_MLEM_FWD:
        return launch_fwd<ARITH, EPI_MLEM_FWD, false>(p, image, imageT, out, nslices, 0, e, st);
    }
    return set_error(ITOMO_EINVAL, "dd forward: unsupported epilogue %d", epi);
}

template <int ARITH>
int dispatch_back(const Plan& p, int epi, int ones, const float* sino, float* out,
                  int nslices, int accumulate, const EpiArgs& e, hipStream_t st)
{
    switch (epi) {
    case EPI_STORE:
        if (ones) launch_back<ARITH, EPI_STORE, true>(p, sino, out, nslices, accumulate, e, st);
        else launch_back<ARITH, EPI_STORE, false>(p, sino, out, nslices, accumulate, e, st);
        return ITOMO_OK;
    case EPI_RECIP:
        launch_back<ARITH, EPI_RECIP, true>(p, sino, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_SIRT_BACK:
        launch_back<ARITH, EPI_SIRT_BACK, false>(p, sino, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_MLEM_BACK:
        launch_back<ARITH, EPI_MLEM_BACK, false>(p, sino, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_SPS_BACK:
        launch_back<ARITH, EPI_SPS_BACK, false>(p, sino, out, nslices, 0, e, st);
        return ITOMO_OK;
    }
    return set_error(ITOMO_EINVAL, "dd back: unsupported epilogue %d", epi);
}

}  // namespace

int dd_forward_epi(const Plan& p, int epi, int ones, const float* image, float* out,
                   int nslices, int accumulate, const EpiArgs& e, hipStream_t st,
                   int* blocks_per_slice)
{
    if (epi == EPI_SIRT_FWD || epi == EPI_SPS_FWD) {
        // The projection, then the epilogue as its own elementwise pass over the
        // sinogram.  Fused into the sweep, these two epilogues (which read a weight
        // array besides b) measured ~40% slower on the wide-bin views with identical
        // sweep code; the extra pass costs one read/write of the sinogram.  Results are
        // the same floats: the epilogue applies the same operations to the same sums.
        int rc = dd_forward_epi(p, EPI_STORE, ones, image, out, nslices, 0, EpiArgs{}, st,
                                nullptr);
        if (rc) return rc;
        return elementwise_fwd_epilogue(epi, out, e, (int64_t)p.nang * p.npix, nslices, st,
                                        blocks_per_slice);
    }
    const int bps = (int)fwd_blocks_per_slice(p);
    if (blocks_per_slice) *blocks_per_slice = bps;
    EpiArgs ee = e;
    ee.blocks_per_slice = bps;
    const float* imageT = image;
    if (!ones) {
        Scratch s;
        int rc = get_scratch(p, nslices, &s, st);
        if (rc) return rc;
        const int64_t n2 = (int64_t)p.npix * p.npix;
        dim3 tb(32, 8), tg((p.npix + 31) / 32, (p.npix + 31) / 32, nslices);
        transpose_kernel<<<tg, tb, 0, st>>>(image, s.xT, p.npix, n2);
        rc = check_hip(hipGetLastError(), "transpose launch");
        if (rc) return rc;
        imageT = s.xT;
    }
    int rc = p.arith == ITOMO_ARITH_FAST
                 ? dispatch_fwd<ITOMO_ARITH_FAST>(p, epi, ones, image, imageT, out, nslices, accumulate, ee, st)
                 : dispatch_fwd<ITOMO_ARITH_IEEE>(p, epi, ones, image, imageT, out, nslices, accumulate, ee, st);
    if (rc) return rc;
    return check_hip(hipGetLastError(), "dd forward launch");
}

int dd_back_epi(const Plan& p, int epi, int ones, const float* sino, float* out, int nslices,
                int accumulate, const EpiArgs& e, hipStream_t st, int* blocks_per_slice)
{
    const int bps = (int)back_blocks_per_slice(p);
    if (blocks_per_slice) *blocks_per_slice = bps;
    EpiArgs ee = e;
    ee.blocks_per_slice = bps;
    int rc = p.arith == ITOMO_ARITH_FAST
                 ? dispatch_back<ITOMO_ARITH_FAST>(p, epi, ones, sino, out, nslices, accumulate, ee, st)
                 : dispatch_back<ITOMO_ARITH_IEEE>(p, epi, ones, sino, out, nslices, accumulate, ee, st);
    if (rc) return rc;
    return check_hip(hipGetLastError(), "dd back launch");
}

int dd_forward(const Plan& p, const float* image, float* sino, int nslices, int64_t,
               int64_t, int accumulate, hipStream_t st)
{
    EpiArgs e{};
    return dd_forward_epi(p, EPI_STORE, 0, image, sino, nslices, accumulate, e, st, nullptr);
}

int dd_back(const Plan& p, const float* sino, float* image, int nslices, int64_t, int64_t,
            int accumulate, hipStream_t st)
{
    EpiArgs e{};
    return dd_back_epi(p, EPI_STORE, 0, sino, image, nslices, accumulate, e, st, nullptr);
}

int dd_build_view_lists(Plan& p)
{
    const DdView* g = (const DdView*)p.h_geom;
    std::vector<int> v;
    v.reserve(p.nang);
    auto wide = [&](int k) {
        return (g[k].kind == DD_SWEEP_ROWS || g[k].kind == DD_SWEEP_COLS) &&
               g[k].ncand > kFwdMainMaxCand;
    };
    for (int k = 0; k < p.nang; ++k)
        if (!wide(k)) v.push_back(k);
    p.n_staged = (int)v.size();
    for (int k = 0; k < p.nang; ++k)
        if (wide(k)) v.push_back(k);
    p.n_gather = p.nang - p.n_staged;
    int rc = check_hip(hipMalloc(&p.d_vlist, (size_t)p.nang * sizeof(int)), "plan_create: hipMalloc");
    if (rc) return rc;
    return check_hip(hipMemcpy(p.d_vlist, v.data(), (size_t)p.nang * sizeof(int),
                               hipMemcpyHostToDevice), "plan_create: hipMemcpy");
}

int dd_build_line_ends(Plan& p)
{
    const size_t n = (size_t)p.nang * p.npix;
    int rc = check_hip(hipMalloc(&p.d_lstar, n * sizeof(float)), "plan_create: hipMalloc");
    if (rc) return rc;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (p.arith == ITOMO_ARITH_FAST)
        line_end_kernel<ITOMO_ARITH_FAST><<<blocks, 256, 0, 0>>>(
            (const DdView*)p.d_geom, p.nang, p.npix, p.d_lstar);
    else
        line_end_kernel<ITOMO_ARITH_IEEE><<<blocks, 256, 0, 0>>>(
            (const DdView*)p.d_geom, p.nang, p.npix, p.d_lstar);
    rc = check_hip(hipGetLastError(), "plan_create: line-end table");
    if (rc) return rc;
    return check_hip(hipStreamSynchronize(0), "plan_create: line-end table");
}

}  // namespace itomo
