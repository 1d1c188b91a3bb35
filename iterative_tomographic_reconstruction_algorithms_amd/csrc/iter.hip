// Iteration drivers of libitomo: SIRT (art.py:211-237) and MLEM (sir.py:204-283) built
// from the projector kernels with fused epilogues, plus fills, FP64 reductions and
// the elementwise epilogue used after the ray-driven scatter back-projection.
#include <mutex>

#include "epilogue.h"

namespace itomo {

__global__ void __launch_bounds__(256) fill_kernel(float* dst, float v, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = v;
}

int fill(float* dst, float value, int64_t n, hipStream_t st)
{
    if (n <= 0) return ITOMO_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    fill_kernel<<<(unsigned)blocks, 256, 0, st>>>(dst, value, n);
    return check_hip(hipGetLastError(), "fill launch");
}

// Sum partials[z][b][k] over b for k in [from, to) into stats[z*3 + k].  One block per
// slice; fixed per-thread strides and a fixed tree, so the result is reproducible.
__global__ void __launch_bounds__(256)
finalize_kernel(const double* __restrict__ partials, int bps, int from, int to,
                double* __restrict__ stats)
{
    __shared__ double red[3][256];
    const int z = blockIdx.x, t = threadIdx.x;
    const double* p = partials + (int64_t)z * bps * 3;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int b = t; b < bps; b += 256) {
        s0 += p[(int64_t)b * 3 + 0];
        s1 += p[(int64_t)b * 3 + 1];
        s2 += p[(int64_t)b * 3 + 2];
    }
    red[0][t] = s0;
    red[1][t] = s1;
    red[2][t] = s2;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) {
            red[0][t] += red[0][t + off];
            red[1][t] += red[1][t + off];
            red[2][t] += red[2][t + off];
        }
        __syncthreads();
    }
    if (t == 0)
        for (int k = from; k < to; ++k) stats[(int64_t)z * 3 + k] = red[k][0];
}

int finalize_partials(const double* partials, int bps, int nslices, int from, int to,
                      double* stats, hipStream_t st)
{
    finalize_kernel<<<nslices, 256, 0, st>>>(partials, bps, from, to, stats);
    return check_hip(hipGetLastError(), "finalize launch");
}

// Elementwise back epilogue over nslices images of n pixels: x <- epi(acc, x).

template <int EPI>
__global__ void __launch_bounds__(256)
back_epi_kernel(const float* __restrict__ acc, float* __restrict__ x, EpiArgs e, int64_t n)
{
    const int z = blockIdx.y;
    acc += (int64_t)z * n;
    x += (int64_t)z * n;
    double p1 = 0.0, p2 = 0.0;
    const int64_t base = (int64_t)blockIdx.x * kEpiChunk;
    for (int k = 0; k < kEpiChunk / 256; ++k) {
        const int64_t o = base + (int64_t)k * 256 + threadIdx.x;
        if (o < n) {
            double q1 = 0.0, q2 = 0.0;
            const float xo = epi_updates_x(EPI) ? x[o] : 0.0f;
            x[o] = epi_back<EPI>(acc[o], xo, o, z, e, &q1, &q2);
            p1 += q1;
            p2 += q2;
        }
    }
    if (epi_has_partials(EPI) && e.partials)
        block_partials3(0.0, p1, p2,
                        e.partials + ((int64_t)z * e.blocks_per_slice + blockIdx.x) * 3);
}

int elementwise_back_epilogue(int epi, const float* acc_img, float* x, const EpiArgs& e,
                              int64_t n, int nslices, hipStream_t st)
{
    EpiArgs ee = e;
    ee.blocks_per_slice = (int)((n + kEpiChunk - 1) / kEpiChunk);
    dim3 grid(ee.blocks_per_slice, nslices);
    switch (epi) {
    case EPI_STORE: back_epi_kernel<EPI_STORE><<<grid, 256, 0, st>>>(acc_img, x, ee, n); break;
    case EPI_RECIP: back_epi_kernel<EPI_RECIP><<<grid, 256, 0, st>>>(acc_img, x, ee, n); break;
    case EPI_SIRT_BACK: back_epi_kernel<EPI_SIRT_BACK><<<grid, 256, 0, st>>>(acc_img, x, ee, n); break;
    case EPI_MLEM_BACK: back_epi_kernel<EPI_MLEM_BACK><<<grid, 256, 0, st>>>(acc_img, x, ee, n); break;
    case EPI_SPS_BACK: back_epi_kernel<EPI_SPS_BACK><<<grid, 256, 0, st>>>(acc_img, x, ee, n); break;
    default: return set_error(ITOMO_EINVAL, "elementwise epilogue %d unsupported", epi);
    }
    return check_hip(hipGetLastError(), "back epilogue launch");
}

// Elementwise forward epilogue over nslices sinograms of n entries, in place.
template <int EPI>
__global__ void __launch_bounds__(256)
fwd_epi_kernel(float* __restrict__ io, EpiArgs e, int64_t n)
{
    const int z = blockIdx.y;
    const int64_t soff = (int64_t)z * n;
    double p0 = 0.0;
    const int64_t base = (int64_t)blockIdx.x * kEpiChunk;
    for (int k = 0; k < kEpiChunk / 256; ++k) {
        const int64_t o = base + (int64_t)k * 256 + threadIdx.x;
        if (o < n) {
            double q0 = 0.0;
            const EpiFwdIn in = epi_fwd_load<EPI>(o, soff, z, e);
            io[soff + o] = epi_forward<EPI>(io[soff + o], in, &q0);
            p0 += q0;
        }
    }
    if (epi_has_partials(EPI) && e.partials)
        block_partials3(p0, 0.0, 0.0,
                        e.partials + ((int64_t)z * e.blocks_per_slice + blockIdx.x) * 3);
}

int elementwise_fwd_epilogue(int epi, float* sino_io, const EpiArgs& e, int64_t n,
                             int nslices, hipStream_t st, int* blocks_per_slice)
{
    EpiArgs ee = e;
    ee.blocks_per_slice = (int)((n + kEpiChunk - 1) / kEpiChunk);
    if (blocks_per_slice) *blocks_per_slice = ee.blocks_per_slice;
    dim3 grid(ee.blocks_per_slice, nslices);
    switch (epi) {
    case EPI_SIRT_FWD: fwd_epi_kernel<EPI_SIRT_FWD><<<grid, 256, 0, st>>>(sino_io, ee, n); break;
    case EPI_SPS_FWD: fwd_epi_kernel<EPI_SPS_FWD><<<grid, 256, 0, st>>>(sino_io, ee, n); break;
    case EPI_MLEM_FWD: fwd_epi_kernel<EPI_MLEM_FWD><<<grid, 256, 0, st>>>(sino_io, ee, n); break;
    default: return set_error(ITOMO_EINVAL, "elementwise forward epilogue %d unsupported", epi);
    }
    return check_hip(hipGetLastError(), "forward epilogue launch");
}

// ---- FP64 dot product ---------------------------------------------------------------
constexpr int kDotBlocks = 1024;

__global__ void __launch_bounds__(256)
dot_partial_kernel(const float* __restrict__ a, const float* __restrict__ b, int64_t n,
                   double* __restrict__ part)
{
    double s = 0.0;
    const int64_t stride = (int64_t)kDotBlocks * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
        s += (double)a[i] * (double)b[i];
    block_partials3(s, 0.0, 0.0, part + (int64_t)blockIdx.x * 3);
}

int dot_f64(const float* a, const float* b, int64_t n, double* out, double* work,
            hipStream_t st)
{
    dot_partial_kernel<<<kDotBlocks, 256, 0, st>>>(a, b, n, work);
    int rc = check_hip(hipGetLastError(), "dot launch");
    if (rc) return rc;
    // stats layout puts the sum at out[0]; finalize writes out[z*3+k] for k=0 only
    return finalize_partials(work, kDotBlocks, 1, 0, 1, out, st);
}

// projector pieces for the ray-driven path (rd.hip)
int rd_forward_epi(const Plan& p, int epi, int ones, const float* image, float* out,
                   int nslices, int accumulate, const EpiArgs& e, hipStream_t st,
                   int* blocks_per_slice);
int rd_back_overwrite(const Plan& p, int ones, const float* sino, float* image, int nslices,
                    hipStream_t st);

int projector_forward_epi_any(const Plan& p, int epi, int ones, const float* image, float* out,
                              int nslices, const EpiArgs& e, hipStream_t st, int* bps)
{
    if (p.projector == ITOMO_PROJ_DD)
        return dd_forward_epi(p, epi, ones, image, out, nslices, 0, e, st, bps);
    if (p.projector == ITOMO_PROJ_PD)
        return pd_forward_epi(p, epi, ones, image, out, nslices, 0, e, st, bps);
    if (p.projector == ITOMO_PROJ_SS)
        return ss_forward_epi(p, epi, ones, image, out, nslices, 0, e, st, bps);
    int rc = rd_matrix_ensure(p, st);   // iterations replay the cached system matrix
    if (rc) return rc;
    return rd_forward_epi(p, epi, ones, image, out, nslices, 0, e, st, bps);
}

// Back projection followed by a back-side epilogue writing `x`.  The ray-driven path
// projects into plan scratch first and runs the epilogue as its own pass; the others fuse
// the epilogue into the gather.
int projector_back_epi_any(const Plan& p, int epi, int ones, const float* sino, float* x,
                           int nslices, const EpiArgs& e, hipStream_t st, int* bps)
{
    if (p.projector == ITOMO_PROJ_DD)
        return dd_back_epi(p, epi, ones, sino, x, nslices, 0, e, st, bps);
    if (p.projector == ITOMO_PROJ_PD)
        return pd_back_epi(p, epi, ones, sino, x, nslices, 0, e, st, bps);
    if (p.projector == ITOMO_PROJ_SS)
        return ss_back_epi(p, epi, ones, sino, x, nslices, 0, e, st, bps);
    int rc = rd_matrix_ensure(p, st);
    if (rc) return rc;
    Scratch s;
    rc = get_scratch(p, nslices, &s, st);
    if (rc) return rc;
    const int64_t n = (int64_t)p.npix * p.npix;
    rc = rd_back_overwrite(p, ones, sino, s.img, nslices, st);
    if (rc) return rc;
    if (bps) *bps = (int)((n + kEpiChunk - 1) / kEpiChunk);
    return elementwise_back_epilogue(epi, s.img, x, e, n, nslices, st);
}

}  // namespace itomo

using namespace itomo;

namespace {

int projector_forward_epi(const Plan& p, int epi, int ones, const float* image, float* out,
                          int nslices, const EpiArgs& e, hipStream_t st, int* bps)
{
    return projector_forward_epi_any(p, epi, ones, image, out, nslices, e, st, bps);
}

int projector_back_epi(const Plan& p, int epi, int ones, const float* sino, float* x,
                       int nslices, const EpiArgs& e, float*, hipStream_t st, int* bps)
{
    return projector_back_epi_any(p, epi, ones, sino, x, nslices, e, st, bps);
}

using IterScratch = Scratch;

int iter_scratch(const Plan& p, int nslices, bool, IterScratch* s, hipStream_t st)
{
    return get_scratch(p, nslices, s, st);
}

int check_iter(const itomo_plan* p, int nslices, const char* who)
{
    if (!p) return set_error(ITOMO_EINVAL, "%s: null plan", who);
    if (nslices < 1 || nslices > 65535)
        return set_error(ITOMO_EINVAL, "%s: nslices %d outside [1, 65535]", who, nslices);
    return ITOMO_OK;
}

}  // namespace

extern "C" {

int itomo_sirt_weights(const itomo_plan* p, float* R, float* C, void* stream)
{
    int rc = check_iter(p, 1, "sirt_weights");
    if (rc) return rc;
    if (!R || !C) return set_error(ITOMO_EINVAL, "sirt_weights: null output");
    hipStream_t st = (hipStream_t)stream;
    IterScratch s;
    rc = iter_scratch(*p, 1, false, &s, st);
    if (rc) return rc;
    EpiArgs e{};
    // R = 1 / A(1)   (art.py:227-230)
    rc = projector_forward_epi(*p, EPI_RECIP, 1, nullptr, R, 1, e, st, nullptr);
    if (rc) return rc;
    // C = 1 / A^T(1) (art.py:234-237)
    return projector_back_epi(*p, EPI_RECIP, 1, nullptr, C, 1, e, s.img, st, nullptr);
}

int itomo_sirt_forward(const itomo_plan* p, const float* x, const float* b, const float* R,
                       float* work, int nslices, double* stats, void* stream)
{
    int rc = check_iter(p, nslices, "sirt_forward");
    if (rc) return rc;
    if (!x || !b || !R || !work) return set_error(ITOMO_EINVAL, "sirt_forward: null array");
    hipStream_t st = (hipStream_t)stream;
    IterScratch s;
    rc = iter_scratch(*p, nslices, stats != nullptr, &s, st);
    if (rc) return rc;
    EpiArgs ef{};
    ef.b = b;
    ef.w = R;
    ef.partials = stats ? s.pf : nullptr;
    int bps = 0;
    rc = projector_forward_epi(*p, EPI_SIRT_FWD, 0, x, work, nslices, ef, st, &bps);
    if (rc) return rc;
    if (stats) rc = finalize_partials(s.pf, bps, nslices, 0, 1, stats, st);
    return rc;
}

int itomo_sirt_back(const itomo_plan* p, const float* work, const float* C, float* x,
                    int nslices, float pc_beta, double* stats, void* stream)
{
    int rc = check_iter(p, nslices, "sirt_back");
    if (rc) return rc;
    if (!x || !C || !work) return set_error(ITOMO_EINVAL, "sirt_back: null array");
    hipStream_t st = (hipStream_t)stream;
    IterScratch s;
    rc = iter_scratch(*p, nslices, stats != nullptr, &s, st);
    if (rc) return rc;
    EpiArgs eb{};
    eb.w = C;
    eb.pc_beta = pc_beta;
    eb.partials = stats ? s.pb : nullptr;
    int bps = 0;
    rc = projector_back_epi(*p, EPI_SIRT_BACK, 0, work, x, nslices, eb, s.img, st, &bps);
    if (rc) return rc;
    if (stats) rc = finalize_partials(s.pb, bps, nslices, 1, 3, stats, st);
    return rc;
}

int itomo_sirt_iterate(const itomo_plan* p, float* x, const float* b, const float* R,
                       const float* C, float* work, int nslices, float pc_beta,
                       double* stats, void* stream)
{
    int rc = itomo_sirt_forward(p, x, b, R, work, nslices, stats, stream);
    if (rc) return rc;
    return itomo_sirt_back(p, work, C, x, nslices, pc_beta, stats, stream);
}

int itomo_mlem_weights(const itomo_plan* p, float* inv_d, void* stream)
{
    int rc = check_iter(p, 1, "mlem_weights");
    if (rc) return rc;
    if (!inv_d) return set_error(ITOMO_EINVAL, "mlem_weights: null output");
    hipStream_t st = (hipStream_t)stream;
    IterScratch s;
    rc = iter_scratch(*p, 1, false, &s, st);
    if (rc) return rc;
    EpiArgs e{};
    // inv_d = 1 / A^T(1) with zeros -> 1e-10 first (sir.py:215-217, 283)
    return projector_back_epi(*p, EPI_RECIP, 1, nullptr, inv_d, 1, e, s.img, st, nullptr);
}

int itomo_mlem_forward(const itomo_plan* p, const float* x, const float* b, float* work,
                       int nslices, double* stats, void* stream)
{
    int rc = check_iter(p, nslices, "mlem_forward");
    if (rc) return rc;
    if (!x || !b || !work) return set_error(ITOMO_EINVAL, "mlem_forward: null array");
    hipStream_t st = (hipStream_t)stream;
    IterScratch s;
    rc = iter_scratch(*p, nslices, stats != nullptr, &s, st);
    if (rc) return rc;
    EpiArgs ef{};
    ef.b = b;
    ef.partials = stats ? s.pf : nullptr;
    int bps = 0;
    rc = projector_forward_epi(*p, EPI_MLEM_FWD, 0, x, work, nslices, ef, st, &bps);
    if (rc) return rc;
    if (stats) rc = finalize_partials(s.pf, bps, nslices, 0, 1, stats, st);
    return rc;
}

int itomo_mlem_back(const itomo_plan* p, const float* work, const float* inv_d, float* x,
                    int nslices, double* stats, void* stream)
{
    int rc = check_iter(p, nslices, "mlem_back");
    if (rc) return rc;
    if (!x || !inv_d || !work) return set_error(ITOMO_EINVAL, "mlem_back: null array");
    hipStream_t st = (hipStream_t)stream;
    IterScratch s;
    rc = iter_scratch(*p, nslices, stats != nullptr, &s, st);
    if (rc) return rc;
    EpiArgs eb{};
    eb.w = inv_d;
    eb.partials = stats ? s.pb : nullptr;
    int bps = 0;
    rc = projector_back_epi(*p, EPI_MLEM_BACK, 0, work, x, nslices, eb, s.img, st, &bps);
    if (rc) return rc;
    if (stats) rc = finalize_partials(s.pb, bps, nslices, 1, 3, stats, st);
    return rc;
}

int itomo_mlem_iterate(const itomo_plan* p, float* x, const float* b, const float* inv_d,
                       float* work, int nslices, double* stats, void* stream)
{
    int rc = itomo_mlem_forward(p, x, b, work, nslices, stats, stream);
    if (rc) return rc;
    return itomo_mlem_back(p, work, inv_d, x, nslices, stats, stream);
}

}  // extern "C"

// ---- SIRT step without allocation ---------------------------------------------------
// The scratch of itomo_sirt_iterate_noalloc: a fixed device array, one copy per GPU,
// held until the process ends.  Host code never passes `scratch` itself to a kernel or
// a copy (on the host the name is a stand-in, not a device address); it uses the
// address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[ITOMO_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex;
void* g_noalloc_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int noalloc_scratch(void** out)
{
    int dev = -1;
    int rc = check_hip(hipGetDevice(&dev), "sirt_iterate_noalloc: hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(ITOMO_EINVAL, "sirt_iterate_noalloc: device %d beyond %d", dev,
                         kMaxDevices);
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_noalloc_ptr[dev]) {
        void* ptr = nullptr;
        rc = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)),
                       "sirt_iterate_noalloc: hipGetSymbolAddress");
        if (rc) return rc;
        if (!ptr) return set_error(ITOMO_EHIP, "sirt_iterate_noalloc: null scratch address");
        g_noalloc_ptr[dev] = ptr;
    }
    *out = g_noalloc_ptr[dev];
    return ITOMO_OK;
}

}  // namespace

extern "C" {

int itomo_sirt_iterate_noalloc(const itomo_plan* p, float* x, const float* b, const float* R,
                               const float* C, float* work, int nslices, float pc_beta,
                               double* stats, void* stream)
{
    int rc = check_iter(p, nslices, "sirt_iterate_noalloc");
    if (rc) return rc;
    if (!x || !b || !R || !C || !work)
        return set_error(ITOMO_EINVAL, "sirt_iterate_noalloc: null array");
    const size_t need = scratch_total(*p, nslices);
    if (need > (size_t)ITOMO_NOALLOC_SCRATCH_BYTES)
        return set_error(ITOMO_EINVAL, "sirt_iterate_noalloc: %d slice(s) of %dx%d with %d "
                                       "views need %zu bytes of scratch, more than the %u "
                                       "this entry point holds",
                         nslices, p->npix, p->npix, p->nang, need,
                         (unsigned)ITOMO_NOALLOC_SCRATCH_BYTES);
    void* base = nullptr;
    rc = noalloc_scratch(&base);
    if (rc) return rc;
    // Every scratch word a step reads is written earlier in the same step (transposed
    // image, one partial-sum slot per block), so nothing needs clearing between calls.
    ScratchScope scope(ScratchRegion{base, (size_t)ITOMO_NOALLOC_SCRATCH_BYTES});
    return itomo_sirt_iterate(p, x, b, R, C, work, nslices, pc_beta, stats, stream);
}

int itomo_sirt_residual(float* ax_io, const float* b, const float* R, int64_t n, int nslices,
                        double* stats, double* work, void* stream)
{
    if (!ax_io || !b || !R || n < 1 || nslices < 1 || nslices > 65535 || (stats && !work))
        return set_error(ITOMO_EINVAL, "sirt_residual: bad argument");
    hipStream_t st = (hipStream_t)stream;
    EpiArgs e{};
    e.b = b;
    e.w = R;
    e.w_stride = 0;
    e.partials = stats ? work : nullptr;
    int bps = 0;
    int rc = elementwise_fwd_epilogue(EPI_SIRT_FWD, ax_io, e, n, nslices, st, &bps);
    if (rc || !stats) return rc;
    return finalize_partials(work, bps, nslices, 0, 1, stats, st);
}

int itomo_sirt_update(const float* atw, const float* C, float* x, int64_t n, int nslices,
                      float pc_beta, double* stats, double* work, void* stream)
{
    if (!atw || !C || !x || n < 1 || nslices < 1 || nslices > 65535 || (stats && !work))
        return set_error(ITOMO_EINVAL, "sirt_update: bad argument");
    hipStream_t st = (hipStream_t)stream;
    EpiArgs e{};
    e.w = C;
    e.w_stride = 0;
    e.pc_beta = pc_beta;
    e.partials = stats ? work : nullptr;
    int rc = elementwise_back_epilogue(EPI_SIRT_BACK, atw, x, e, n, nslices, st);
    if (rc || !stats) return rc;
    return finalize_partials(work, (int)((n + kEpiChunk - 1) / kEpiChunk), nslices, 1, 3, stats,
                             st);
}

int itomo_reciprocal(const float* v, float* out, int64_t n, void* stream)
{
    if (!v || !out || n < 1) return set_error(ITOMO_EINVAL, "reciprocal: bad argument");
    EpiArgs e{};
    return elementwise_back_epilogue(EPI_RECIP, v, out, e, n, 1, (hipStream_t)stream);
}

int itomo_dot_f64(const float* a, const float* b, int64_t n, double* out, double* work,
                  void* stream)
{
    if (!a || !b || !out || !work || n < 0)
        return set_error(ITOMO_EINVAL, "dot_f64: bad argument");
    return dot_f64(a, b, n, out, work, (hipStream_t)stream);
}

}  // extern "C"
