// SIRT constraints applied after each step (art.py:296-321): the support mask
// (utils.supp_constr, utils.py:264-266), the additional a-priori masks
// (utils.mask_constr, utils.py:279-281), non-negativity (utils.phys_constr,
// utils.py:171-176) and the resolution circle (utils.resol_circle_constr,
// utils.py:229-251), on the reconstruction square [i1, i2)^2 of each slice, in the
// reference's order.  Optionally the monitor sums of the constrained step.
#include <algorithm>

#include "epilogue.h"

namespace itomo {

constexpr int kConstrChunk = 2048;   // pixels per block of the elementwise passes

struct ConstrArgs {
    int npix, i1, m;     // image side, region origin and side (m = i2 - i1)
    int64_t n2;          // npix * npix
    int bps;             // blocks per slice
};

// x = 0 where mask == 0 (utils.supp_constr).
__global__ void __launch_bounds__(256)
constr_supp_kernel(float* __restrict__ x, const unsigned char* __restrict__ mask, ConstrArgs a)
{
    const int z = blockIdx.y;
    const int64_t m2 = (int64_t)a.m * a.m;
    float* xs = x + (int64_t)z * a.n2;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m2;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(k / a.m), c = (int)(k - (int64_t)r * a.m);
        if (mask[k] == 0) xs[(int64_t)(a.i1 + r) * a.npix + a.i1 + c] = 0.0f;
    }
}

// Per-block FP64 sum and count of x over mask == 1 (fixed grouping: reproducible).
__global__ void __launch_bounds__(256)
constr_mask_sum_kernel(const float* __restrict__ x, const unsigned char* __restrict__ mask,
                       double* __restrict__ part, ConstrArgs a)
{
    const int z = blockIdx.y;
    const int64_t m2 = (int64_t)a.m * a.m;
    const float* xs = x + (int64_t)z * a.n2;
    double s = 0.0, cnt = 0.0;
    const int64_t k0 = (int64_t)blockIdx.x * kConstrChunk;
    for (int64_t k = k0 + threadIdx.x; k < min(k0 + kConstrChunk, m2); k += blockDim.x) {
        if (mask[k] == 1) {
            const int r = (int)(k / a.m), c = (int)(k - (int64_t)r * a.m);
            s += (double)xs[(int64_t)(a.i1 + r) * a.npix + a.i1 + c];
            cnt += 1.0;
        }
    }
    block_partials3(s, cnt, 0.0, part + ((int64_t)z * a.bps + blockIdx.x) * 3);
}

// x = mean of x over mask == 1 there (utils.mask_constr), the mean rounded to float.
__global__ void __launch_bounds__(256)
constr_mask_set_kernel(float* __restrict__ x, const unsigned char* __restrict__ mask,
                       const double* __restrict__ part, ConstrArgs a)
{
    const int z = blockIdx.y;
    __shared__ float smean;
    if (threadIdx.x == 0) {
        double s = 0.0, cnt = 0.0;
        const double* p = part + (int64_t)z * a.bps * 3;
        for (int b = 0; b < a.bps; ++b) {
            s += p[b * 3];
            cnt += p[b * 3 + 1];
        }
        smean = cnt > 0.0 ? (float)(s / cnt) : 0.0f;
    }
    __syncthreads();
    const float mean = smean;
    const int64_t m2 = (int64_t)a.m * a.m;
    float* xs = x + (int64_t)z * a.n2;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m2;
         k += (int64_t)gridDim.x * blockDim.x) {
        if (mask[k] == 1) {
            const int r = (int)(k / a.m), c = (int)(k - (int64_t)r * a.m);
            xs[(int64_t)(a.i1 + r) * a.npix + a.i1 + c] = mean;
        }
    }
}

// pc constraints on the region (art.py:314-321), then the monitor sums over the whole
// image: part[0] unused, part[1] += (x - x_old)^2, part[2] += x_old^2.
__global__ void __launch_bounds__(256)
constr_pc_kernel(float* __restrict__ x, const float* __restrict__ x_old, float beta,
                 int circle, double* __restrict__ part, ConstrArgs a)
{
    const int z = blockIdx.y;
    float* xs = x + (int64_t)z * a.n2;
    const float* xo = x_old ? x_old + (int64_t)z * a.n2 : nullptr;
    // resol_circle_constr: ctr = m/2 - 1/2, radius d = ctr, zero where the distance > d
    const double ctr = a.m * 0.5 - 0.5;
    double q1 = 0.0, q2 = 0.0;
    const int64_t k0 = (int64_t)blockIdx.x * kConstrChunk;
    for (int64_t k = k0 + threadIdx.x; k < min(k0 + kConstrChunk, a.n2); k += blockDim.x) {
        const int r = (int)(k / a.npix), c = (int)(k - (int64_t)r * a.npix);
        float v = xs[k];
        const int rr = r - a.i1, cc = c - a.i1;
        if (rr >= 0 && rr < a.m && cc >= 0 && cc < a.m) {
            if (beta >= 0.0f && v < 0.0f) v = __fmul_rn(v, beta);
            if (circle) {
                const double dr = rr - ctr, dc = cc - ctr;
                if (sqrt(dr * dr + dc * dc) > ctr) v = 0.0f;
            }
            xs[k] = v;
        }
        if (xo) {
            const double o = xo[k], d = (double)v - o;
            q1 += d * d;
            q2 += o * o;
        }
    }
    if (part) block_partials3(0.0, q1, q2, part + ((int64_t)z * a.bps + blockIdx.x) * 3);
}

}  // namespace itomo

using namespace itomo;

extern "C" {

int itomo_sirt_constraints(float* x, int npix, int nslices, int i1, int i2,
                           const unsigned char* mask, const unsigned char* mask_add, int nmask,
                           float pc, const float* x_old, double* stats, double* work,
                           void* stream)
{
    if (!x || npix < 1 || nslices < 1 || nslices > 65535 || i1 < 0 || i2 > npix || i2 <= i1 ||
        nmask < 0 || (nmask > 0 && !mask_add) || (stats && (!x_old || !work)) ||
        ((nmask > 0) && !work))
        return set_error(ITOMO_EINVAL, "sirt_constraints: bad argument");
    hipStream_t st = (hipStream_t)stream;
    ConstrArgs a;
    a.npix = npix;
    a.i1 = i1;
    a.m = i2 - i1;
    a.n2 = (int64_t)npix * npix;
    const int64_t m2 = (int64_t)a.m * a.m;
    a.bps = (int)((a.n2 + kConstrChunk - 1) / kConstrChunk);
    const int eblocks = (int)std::min<int64_t>((m2 + 255) / 256, 4096);
    if (mask) {
        constr_supp_kernel<<<dim3(eblocks, nslices), 256, 0, st>>>(x, mask, a);
        int rc = check_hip(hipGetLastError(), "sirt_constraints: supp launch");
        if (rc) return rc;
    }
    for (int j = 0; j < nmask; ++j) {
        const unsigned char* mj = mask_add + (int64_t)j * m2;
        ConstrArgs am = a;
        am.bps = (int)((m2 + kConstrChunk - 1) / kConstrChunk);
        constr_mask_sum_kernel<<<dim3(am.bps, nslices), 256, 0, st>>>(x, mj, work, am);
        constr_mask_set_kernel<<<dim3(eblocks, nslices), 256, 0, st>>>(x, mj, work, am);
        int rc = check_hip(hipGetLastError(), "sirt_constraints: mask launch");
        if (rc) return rc;
    }
    // pc: 1 -> x<0 -> 0 and the circle; 2 -> the circle only; other >= 0 -> x<0 *= pc
    const bool circle = pc == 1.0f || pc == 2.0f;
    const float beta = pc == 1.0f ? 0.0f : (pc == 2.0f || pc < 0.0f ? -1.0f : pc);
    if (circle || beta >= 0.0f || stats) {
        constr_pc_kernel<<<dim3(a.bps, nslices), 256, 0, st>>>(x, stats ? x_old : nullptr, beta,
                                                              circle ? 1 : 0,
                                                              stats ? work : nullptr, a);
        int rc = check_hip(hipGetLastError(), "sirt_constraints: pc launch");
        if (rc) return rc;
        if (stats) return finalize_partials(work, a.bps, nslices, 1, 3, stats, st);
    }
    return ITOMO_OK;
}

}  // extern "C"
