// Projector epilogues: what a projector kernel does with the value it accumulated for
// one output element, plus deterministic FP64 partial sums for the iteration monitors.
#pragma once

#include "itomo_internal.h"

namespace itomo {

enum Epi : int {
    EPI_STORE = 0,      // out = acc (or out += acc when accumulating)
    EPI_RECIP = 1,      // out = 1 / (acc == 0 ? acc + 1e-10f : acc)   (art.py:227-237)
    EPI_SIRT_FWD = 2,   // out = R * (b - acc); part0 += (b - acc)^2          (art.py:212)
    EPI_SIRT_BACK = 3,  // x = x + C * acc; x<0 -> x*beta; part1 += dx^2; part2 += x^2
    EPI_MLEM_FWD = 4,   // ax = acc (0 -> 2^-52); out = b / ax; part0 += (ax - b)^2 (sir.py:270-271)
    EPI_MLEM_BACK = 5,  // x = x * (inv_d * acc); part1 += dx^2; part2 += x^2     (sir.py:283)
    EPI_SPS_FWD = 6,    // out = (acc - b) * w; part0 += out^2                  (sir.py:434-441)
    EPI_SPS_BACK = 7,   // x = x - acc / (den + 2^-52), quotient in FP64; part1, part2 (sir.py:535)
};

struct EpiArgs {
    const float* b;      // measured sinogram (slice stride = sino stride) -- forward epilogues
    const float* w;      // R (nang x npix) / C, inv_d (npix x npix) / SPS w, den
    int64_t w_stride;    // elements between slices of w (0: one w shared by all slices)
    float pc_beta;       // SIRT back: < 0 disables the non-negativity step
    double* partials;    // [slice][block][3] FP64 partial sums, or nullptr
    int blocks_per_slice;
};

// Block-wide deterministic sum of three doubles; thread 0 of the block writes them.
// Requires blockDim.x*blockDim.y*blockDim.z a multiple of 64 and at most 1024.
// Per-wave partial sums of a 1-D block: lane 0 of each wave writes its wave's three
// sums to dst[0..2].  No barrier, so waves that finish early leave the compute unit at
// once (the forward's lanes walk line ranges of different lengths).
__device__ __forceinline__ void wave_partials3(double a0, double a1, double a2, double* dst)
{
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_down(a0, off, 64);
        a1 += __shfl_down(a1, off, 64);
        a2 += __shfl_down(a2, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        dst[0] = a0;
        dst[1] = a1;
        dst[2] = a2;
    }
}

__device__ __forceinline__ void block_partials3(double a0, double a1, double a2,
                                                double* dst)
{
    __shared__ double red[3][16];
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_down(a0, off, 64);
        a1 += __shfl_down(a1, off, 64);
        a2 += __shfl_down(a2, off, 64);
    }
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wid = tid >> 6, nw = nthreads >> 6;
    if (lane == 0) {
        red[0][wid] = a0;
        red[1][wid] = a1;
        red[2][wid] = a2;
    }
    __syncthreads();
    if (tid == 0) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int k = 0; k < nw; ++k) {
            s0 += red[0][k];
            s1 += red[1][k];
            s2 += red[2][k];
        }
        dst[0] = s0;
        dst[1] = s1;
        dst[2] = s2;
    }
}

// Inputs of the forward-side epilogue of one output element, loaded before the
// projection sum so that their latency overlaps it.
struct EpiFwdIn {
    float b, w;
};

template <int EPI>
__device__ __forceinline__ EpiFwdIn epi_fwd_load(int64_t o, int64_t slice_off, int z,
                                                 const EpiArgs& e)
{
    EpiFwdIn in{0.0f, 0.0f};
    if (EPI == EPI_SPS_FWD || EPI == EPI_SIRT_FWD || EPI == EPI_MLEM_FWD)
        in.b = e.b[slice_off + o];
    if (EPI == EPI_SPS_FWD || EPI == EPI_SIRT_FWD) in.w = e.w[(int64_t)z * e.w_stride + o];
    return in;
}

// Forward-side epilogue for one output element with its loaded inputs.  Returns the
// value to store; writes the residual term to *p0.
template <int EPI>
__device__ __forceinline__ float epi_forward(float acc, const EpiFwdIn& in, double* p0)
{
    if (EPI == EPI_SPS_FWD) {
        const float dif = __fmul_rn(__fsub_rn(acc, in.b), in.w);
        *p0 = (double)dif * (double)dif;
        return dif;
    } else if (EPI == EPI_SIRT_FWD) {
        const float dif = __fsub_rn(in.b, acc);
        *p0 = (double)dif * (double)dif;
        return __fmul_rn(in.w, dif);
    } else if (EPI == EPI_MLEM_FWD) {
        const float bv = in.b;
        float ax = acc;
        if (ax == 0.0f) ax = __fadd_rn(ax, 2.220446049250313e-16f);
        const double dd = (double)ax - (double)bv;
        *p0 = dd * dd;
        return __fdiv_rn(bv, ax);
    } else if (EPI == EPI_RECIP) {
        float v = acc;
        if (v == 0.0f) v = __fadd_rn(v, 1e-10f);
        return __fdiv_rn(1.0f, v);
    }
    return acc;
}

// Forward-side epilogue for output element index `o` (within one slice) of slice z.
template <int EPI>
__device__ __forceinline__ float epi_forward(float acc, int64_t o, int64_t slice_off, int z,
                                             const EpiArgs& e, double* p0)
{
    return epi_forward<EPI>(acc, epi_fwd_load<EPI>(o, slice_off, z, e), p0);
}

constexpr bool epi_back_uses_w(int epi)
{
    return epi == EPI_SPS_BACK || epi == EPI_SIRT_BACK || epi == EPI_MLEM_BACK;
}

// Back-side epilogue of one pixel: `xo` is the current iterate value and `w` the
// pixel's weight (C, 1/A^T 1 or the SPS denominator), both already loaded.
template <int EPI>
__device__ __forceinline__ float epi_back_v(float acc, float xo, float w, const EpiArgs& e,
                                            double* p1, double* p2)
{
    if (EPI == EPI_SPS_BACK) {
        const double q = (double)acc / ((double)w + 2.220446049250313e-16);
        const float xn = (float)((double)xo - q);
        const double dd = (double)xn - (double)xo;
        *p1 = dd * dd;
        *p2 = (double)xo * (double)xo;
        return xn;
    } else if (EPI == EPI_SIRT_BACK) {
        float xn = __fadd_rn(xo, __fmul_rn(w, acc));
        if (e.pc_beta >= 0.0f && xn < 0.0f) xn = __fmul_rn(xn, e.pc_beta);
        const double dd = (double)xn - (double)xo;
        *p1 = dd * dd;
        *p2 = (double)xo * (double)xo;
        return xn;
    } else if (EPI == EPI_MLEM_BACK) {
        const float xn = __fmul_rn(xo, __fmul_rn(w, acc));
        const double dd = (double)xn - (double)xo;
        *p1 = dd * dd;
        *p2 = (double)xo * (double)xo;
        return xn;
    } else if (EPI == EPI_RECIP) {
        float v = acc;
        if (v == 0.0f) v = __fadd_rn(v, 1e-10f);
        return __fdiv_rn(1.0f, v);
    }
    return acc;
}

// Back-side epilogue for pixel index `o` of slice z: `xo` is the current iterate value.
template <int EPI>
__device__ __forceinline__ float epi_back(float acc, float xo, int64_t o, int z,
                                          const EpiArgs& e, double* p1, double* p2)
{
    const float w = epi_back_uses_w(EPI) ? e.w[(int64_t)z * e.w_stride + o] : 0.0f;
    return epi_back_v<EPI>(acc, xo, w, e, p1, p2);
}

constexpr bool epi_has_partials(int epi)
{
    return epi == EPI_SIRT_FWD || epi == EPI_SIRT_BACK || epi == EPI_MLEM_FWD ||
           epi == EPI_MLEM_BACK || epi == EPI_SPS_FWD || epi == EPI_SPS_BACK;
}

constexpr bool epi_updates_x(int epi)
{
    return epi == EPI_SIRT_BACK || epi == EPI_MLEM_BACK || epi == EPI_SPS_BACK;
}

// host helpers (iter.hip)
int fill(float* dst, float value, int64_t n, hipStream_t st);
int finalize_partials(const double* partials, int blocks_per_slice, int nslices,
                      int which_from, int which_to, double* stats, hipStream_t st);
int elementwise_back_epilogue(int epi, const float* acc_img, float* x, const EpiArgs& e,
                              int64_t npix2, int nslices, hipStream_t st);
int elementwise_fwd_epilogue(int epi, float* sino_io, const EpiArgs& e, int64_t nsino,
                             int nslices, hipStream_t st, int* blocks_per_slice);

// projector launchers with epilogues (dd.hip)
int dd_forward_epi(const Plan& p, int epi, int ones_input, const float* image, float* out,
                   int nslices, int accumulate, const EpiArgs& e, hipStream_t st,
                   int* blocks_per_slice);
int dd_back_epi(const Plan& p, int epi, int ones_input, const float* sino, float* out,
                int nslices, int accumulate, const EpiArgs& e, hipStream_t st,
                int* blocks_per_slice);
// any projector, with an epilogue (iter.hip)
int projector_forward_epi_any(const Plan& p, int epi, int ones, const float* image, float* out,
                              int nslices, const EpiArgs& e, hipStream_t st, int* bps);
int projector_back_epi_any(const Plan& p, int epi, int ones, const float* sino, float* x,
                           int nslices, const EpiArgs& e, hipStream_t st, int* bps);

// pd.hip
int pd_forward_epi(const Plan& p, int epi, int ones_input, const float* image, float* out,
                   int nslices, int accumulate, const EpiArgs& e, hipStream_t st,
                   int* blocks_per_slice);
int pd_back_epi(const Plan& p, int epi, int ones_input, const float* sino, float* out,
                int nslices, int accumulate, const EpiArgs& e, hipStream_t st,
                int* blocks_per_slice);
// ss.hip
int ss_forward_epi(const Plan& p, int epi, int ones_input, const float* image, float* out,
                   int nslices, int accumulate, const EpiArgs& e, hipStream_t st,
                   int* blocks_per_slice);
int ss_back_epi(const Plan& p, int epi, int ones_input, const float* sino, float* out,
                int nslices, int accumulate, const EpiArgs& e, hipStream_t st,
                int* blocks_per_slice);

}  // namespace itomo
