// Distance-driven projector pair (semantics of radon_dd.c:54-285) for gfx950.
//
// The reference sorts, per view and image line, the N+1 pixel boundaries and the N+1
// detector-bin boundaries and sweeps the merged list; each segment between consecutive
// keys that lies inside pixel q and bin d moves weight |segment| between them, and the
// segment that ends at the first key of index N is dropped (radon_dd.c:176-177).
// Here the same weights are evaluated in gather form: w(q, d) = min(P[q+1], D[d+1]) -
// max(P[q], D[d]) when positive, and 0 for the one pair per line whose right end is
// E = min(P[N], D[N]).  Keys are formed with exactly the float operations of the
// reference build, so every weight is bit-identical to the reference's.
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
#include "epilogue.h"

namespace itomo {

namespace {

// Bin boundary D for j = d - nh.
template <int ARITH>
__device__ __forceinline__ float key_D(const DdView& g, int j)
{
    const float x = (float)j;
    if (ARITH == ITOMO_ARITH_FAST) return __fmul_rn(x, g.rinv);
    return __fdiv_rn(x, g.b);
}

// Line offset K for line coordinate u = (i - nh) + 0.5.
template <int ARITH>
__device__ __forceinline__ float key_K(const DdView& g, float u)
{
    if (ARITH == ITOMO_ARITH_FAST) return __fmul_rn(u, g.t);
    return __fdiv_rn(__fmul_rn(u, g.a), g.b);
}

__device__ __forceinline__ float line_coord(int i, int nh) { return (float)(i - nh) + 0.5f; }

// acc += w * v as the reference build evaluates it: fused under -ffast-math (FAST),
// a rounded product then a rounded sum as written (IEEE).
template <int ARITH>
__device__ __forceinline__ float accum(float w, float v, float acc)
{
    if (ARITH == ITOMO_ARITH_FAST) return __builtin_fmaf(w, v, acc);
    return __fadd_rn(acc, __fmul_rn(w, v));
}

// Accumulate the pixel with boundaries [lo, hi] against all bins it overlaps, in
// increasing bin order.  E is the dropped right end of this line.
template <int ARITH>
__device__ __forceinline__ float pixel_gather(const DdView& g, int npix, int nh, float lo,
                                              float hi, float E,
                                              const float* __restrict__ srow, bool ones,
                                              float acc)
{
    // estimate the bin containing lo, then correct it with the exact keys
    float fe = floorf(__fmul_rn(lo, g.est));
    fe = fminf(fmaxf(fe, (float)(-nh - 1)), (float)(npix - nh));
    int d = (int)fe + nh;  // in [-1, npix]
    float Dlo = key_D<ARITH>(g, d - nh);
    float Dhi = key_D<ARITH>(g, d + 1 - nh);
    while (d > -1 && lo < Dlo) {
        --d;
        Dhi = Dlo;
        Dlo = key_D<ARITH>(g, d - nh);
    }
    while (d < npix && lo >= Dhi) {
        ++d;
        Dlo = Dhi;
        Dhi = key_D<ARITH>(g, d + 1 - nh);
    }
    while (d < npix && Dlo < hi) {
        if (d >= 0) {
            const float R = fminf(hi, Dhi), L = fmaxf(lo, Dlo);
            if (R > L && R != E)
                acc = accum<ARITH>(__fsub_rn(R, L), ones ? 1.0f : srow[d], acc);
        }
        ++d;
        Dlo = Dhi;
        Dhi = key_D<ARITH>(g, d + 1 - nh);
    }
    return acc;
}

}  // namespace

// Back projection: one thread per pixel (row r, column q) of slice blockIdx.z.
template <int ARITH, int EPI, bool ONES>
__global__ void __launch_bounds__(256)
dd_back_kernel(const DdView* __restrict__ geom, int nang, int npix,
               const float* __restrict__ sino, float* __restrict__ out,
               int64_t sino_stride, int64_t img_stride, int accumulate, EpiArgs e)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;  // column
    const int r = blockIdx.y * blockDim.y + threadIdx.y;  // row
    const bool active = q < npix && r < npix;
    const int z = blockIdx.z;
    sino += (int64_t)z * sino_stride;
    out += (int64_t)z * img_stride;
    const int nh = (int)(npix * 0.5);
    const int64_t o = (int64_t)r * npix + q;
    double p1 = 0.0, p2 = 0.0;

    if (active) {
        float acc = (EPI == EPI_STORE && accumulate) ? out[o] : 0.0f;
        const float ur = line_coord(r, nh), uq = line_coord(q, nh);
        const float xr0 = (float)(r - nh), xr1 = (float)(r + 1 - nh);
        const float xq0 = (float)(q - nh), xq1 = (float)(q + 1 - nh);
        const float xN = (float)(npix - nh);
        for (int v = 0; v < nang; ++v) {
            const DdView g = geom[v];
            const float* srow = sino + (int64_t)g.row * npix;
            if (g.kind == DD_SWEEP_ROWS) {
                const float K = key_K<ARITH>(g, ur);
                const float E = fminf(__fsub_rn(xN, K), key_D<ARITH>(g, npix - nh));
                acc = pixel_gather<ARITH>(g, npix, nh, __fsub_rn(xq0, K), __fsub_rn(xq1, K),
                                          E, srow, ONES, acc);
            } else if (g.kind == DD_SWEEP_COLS) {
                const float K = key_K<ARITH>(g, uq);
                const float E = fminf(__fsub_rn(xN, K), key_D<ARITH>(g, npix - nh));
                acc = pixel_gather<ARITH>(g, npix, nh, __fsub_rn(xr0, K), __fsub_rn(xr1, K),
                                          E, srow, ONES, acc);
            } else if (g.kind == DD_ROWSUM) {
                acc = __fadd_rn(acc, ONES ? 1.0f : srow[r]);
            } else if (g.kind == DD_COLSUM) {
                acc = __fadd_rn(acc, ONES ? 1.0f : srow[q]);
            }
        }
        if (EPI == EPI_SIRT_BACK || EPI == EPI_MLEM_BACK)
            out[o] = epi_back<EPI>(acc, out[o], o, e, &p1, &p2);
        else
            out[o] = epi_back<EPI>(acc, 0.0f, o, e, &p1, &p2);
    }
    if (epi_has_partials(EPI) && e.partials) {
        const int blk = blockIdx.y * gridDim.x + blockIdx.x;
        block_partials3(0.0, p1, p2,
                        e.partials + ((int64_t)z * e.blocks_per_slice + blk) * 3);
    }
}

// Forward projection: one thread per (bin, view blockIdx.y, slice blockIdx.z).
template <int ARITH, int EPI, bool ONES>
__global__ void __launch_bounds__(256)
dd_fwd_kernel(const DdView* __restrict__ geom, int nang, int npix,
              const float* __restrict__ image, float* __restrict__ out,
              int64_t img_stride, int64_t sino_stride, int accumulate, EpiArgs e)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = d < npix;
    const int z = blockIdx.z;
    const DdView g = geom[blockIdx.y];
    image += (int64_t)z * img_stride;
    const int64_t soff = (int64_t)z * sino_stride;
    const int64_t o = (int64_t)g.row * npix + d;
    const int nh = (int)(npix * 0.5);
    double p0 = 0.0;

    if (active) {
        float acc = (EPI == EPI_STORE && accumulate) ? out[soff + o] : 0.0f;
        if (g.kind == DD_ROWSUM) {
            const float* p = image + (int64_t)d * npix;
            for (int j = 0; j < npix; ++j) acc = __fadd_rn(acc, ONES ? 1.0f : p[j]);
        } else if (g.kind == DD_COLSUM) {
            const float* p = image + d;
            for (int j = 0; j < npix; ++j)
                acc = __fadd_rn(acc, ONES ? 1.0f : p[(int64_t)j * npix]);
        } else if (g.kind == DD_SWEEP_ROWS || g.kind == DD_SWEEP_COLS) {
            // line i: image row i (rows) or column i (columns); pixel q along it
            const int64_t line_step = (g.kind == DD_SWEEP_ROWS) ? npix : 1;
            const int64_t pix_step = (g.kind == DD_SWEEP_ROWS) ? 1 : npix;
            const float Dlo = key_D<ARITH>(g, d - nh);
            const float Dhi = key_D<ARITH>(g, d + 1 - nh);
            const float DN = key_D<ARITH>(g, npix - nh);
            const float x0 = (float)(-nh), xN = (float)(npix - nh);
            for (int i = 0; i < npix; ++i) {
                const float K = key_K<ARITH>(g, line_coord(i, nh));
                const float P0 = __fsub_rn(x0, K), PN = __fsub_rn(xN, K);
                if (!(PN > Dlo && P0 < Dhi)) continue;
                const float E = fminf(PN, DN);
                float fe = floorf(__fadd_rn(Dlo, K));
                fe = fminf(fmaxf(fe, (float)(-nh - 1)), (float)(npix - nh));
                int q = (int)fe + nh;  // in [-1, npix]
                float Plo = __fsub_rn((float)(q - nh), K);
                float Phi = __fsub_rn((float)(q + 1 - nh), K);
                while (q > -1 && Dlo < Plo) {
                    --q;
                    Phi = Plo;
                    Plo = __fsub_rn((float)(q - nh), K);
                }
                while (q < npix && Dlo >= Phi) {
                    ++q;
                    Plo = Phi;
                    Phi = __fsub_rn((float)(q + 1 - nh), K);
                }
                const float* line = image + (int64_t)i * line_step;
                while (q < npix && Plo < Dhi) {
                    if (q >= 0) {
                        const float R = fminf(Phi, Dhi), L = fmaxf(Plo, Dlo);
                        if (R > L && R != E)
                            acc = accum<ARITH>(__fsub_rn(R, L),
                                               ONES ? 1.0f : line[(int64_t)q * pix_step], acc);
                    }
                    ++q;
                    Plo = Phi;
                    Phi = __fsub_rn((float)(q + 1 - nh), K);
                }
            }
        }
        out[soff + o] = epi_forward<EPI>(acc, o, soff, e, &p0);
    }
    if (epi_has_partials(EPI) && e.partials) {
        const int blk = blockIdx.y * gridDim.x + blockIdx.x;
        block_partials3(p0, 0.0, 0.0,
                        e.partials + ((int64_t)z * e.blocks_per_slice + blk) * 3);
    }
}

namespace {

template <int ARITH, int EPI, bool ONES>
void launch_fwd(const Plan& p, const float* image, float* out, int nslices, int accumulate,
                const EpiArgs& e, hipStream_t st)
{
    dim3 block(256), grid((p.npix + 255) / 256, p.nang, nslices);
    dd_fwd_kernel<ARITH, EPI, ONES><<<grid, block, 0, st>>>(
        (const DdView*)p.d_geom, p.nang, p.npix, image, out, (int64_t)p.npix * p.npix,
        (int64_t)p.nang * p.npix, accumulate, e);
}

template <int ARITH, int EPI, bool ONES>
void launch_back(const Plan& p, const float* sino, float* out, int nslices, int accumulate,
                 const EpiArgs& e, hipStream_t st)
{
    dim3 block(64, 4), grid((p.npix + 63) / 64, (p.npix + 3) / 4, nslices);
    dd_back_kernel<ARITH, EPI, ONES><<<grid, block, 0, st>>>(
        (const DdView*)p.d_geom, p.nang, p.npix, sino, out, (int64_t)p.nang * p.npix,
        (int64_t)p.npix * p.npix, accumulate, e);
}

template <int ARITH>
int dispatch_fwd(const Plan& p, int epi, int ones, const float* image, float* out,
                 int nslices, int accumulate, const EpiArgs& e, hipStream_t st)
{
    switch (epi) {
    case EPI_STORE:
        if (ones) launch_fwd<ARITH, EPI_STORE, true>(p, image, out, nslices, accumulate, e, st);
        else launch_fwd<ARITH, EPI_STORE, false>(p, image, out, nslices, accumulate, e, st);
        return ITOMO_OK;
    case EPI_RECIP:
        launch_fwd<ARITH, EPI_RECIP, true>(p, image, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_SIRT_FWD:
        launch_fwd<ARITH, EPI_SIRT_FWD, false>(p, image, out, nslices, 0, e, st);
        return ITOMO_OK;
    case EPI_MLEM_FWD:
        launch_fwd<ARITH, EPI_MLEM_FWD, false>(p, image, out, nslices, 0, e, st);
        return ITOMO_OK;
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
    }
    return set_error(ITOMO_EINVAL, "dd back: unsupported epilogue %d", epi);
}

}  // namespace

int dd_forward_epi(const Plan& p, int epi, int ones, const float* image, float* out,
                   int nslices, int accumulate, const EpiArgs& e, hipStream_t st,
                   int* blocks_per_slice)
{
    if (blocks_per_slice) *blocks_per_slice = ((p.npix + 255) / 256) * p.nang;
    EpiArgs ee = e;
    ee.blocks_per_slice = ((p.npix + 255) / 256) * p.nang;
    int rc = p.arith == ITOMO_ARITH_FAST
                 ? dispatch_fwd<ITOMO_ARITH_FAST>(p, epi, ones, image, out, nslices, accumulate, ee, st)
                 : dispatch_fwd<ITOMO_ARITH_IEEE>(p, epi, ones, image, out, nslices, accumulate, ee, st);
    if (rc) return rc;
    return check_hip(hipGetLastError(), "dd forward launch");
}

int dd_back_epi(const Plan& p, int epi, int ones, const float* sino, float* out, int nslices,
                int accumulate, const EpiArgs& e, hipStream_t st, int* blocks_per_slice)
{
    const int bps = ((p.npix + 63) / 64) * ((p.npix + 3) / 4);
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

}  // namespace itomo
