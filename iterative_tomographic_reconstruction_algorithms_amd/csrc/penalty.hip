// SPS penalties and the penalised, clipped SPS update (sir.py:348-560) for gfx950.
//
//   penalty_kernel        R_num / R_den of huber.c, tikhonov.c, haar.c (penalty.pyx:16-106):
//                         3x3 neighbour sums in the reference's order, arithmetic of its
//                         -ffast-math build (FAST) or of the C as written (IEEE)
//   sobel_mag_kernel      sqrt(gx^2 + gy^2) of scipy.ndimage.sobel with mode='reflect'
//                         along both axes (sir.py:339-343), the input of the Huber penalty
//   sps_update_kernel     x <- x - (num + beta R_num) / (den + beta R_den + 2^-52) with the
//                         quotient in FP64, then the clip to [0, max x] (sir.py:535-540)
//   max_final_kernel_d /
//   clip_fix_kernel       the clip's upper end: np.clip(x, 0, M) with M = max x differs from
//                         max(x, 0) only when every entry is negative (then all become M)
#include "epilogue.h"

namespace itomo {

constexpr int kPenTile = 16;   // 16 x 16 pixels per block

// Neighbour weights of penalty_weights.c:7-15, in the order k = 3 (k1 + 1) + (k2 + 1).
__device__ __forceinline__ float pen_weight(int k)
{
    return k == 4 ? 4.0f : ((k & 1) ? 1.0f : 0.70710677f);  // (float)(1/sqrt(2))
}

template <int ARITH, int KIND>
__global__ void __launch_bounds__(kPenTile * kPenTile)
penalty_kernel(const float* __restrict__ img, int nr, int nc, float delta,
               float* __restrict__ r_num, float* __restrict__ r_den)
{
    const int j = blockIdx.x * kPenTile + threadIdx.x;
    const int i = blockIdx.y * kPenTile + threadIdx.y;
    if (i >= nr || j >= nc) return;
    const int64_t off = (int64_t)blockIdx.z * nr * nc;
    img += off;
    const float c = img[(int64_t)i * nc + j];
    float rn = 0.0f, rd = 0.0f;
#pragma unroll
    for (int k1 = -1; k1 <= 1; ++k1)
#pragma unroll
        for (int k2 = -1; k2 <= 1; ++k2) {
            const int k = 3 * (k1 + 1) + (k2 + 1);
            if (i + k1 < 0 || i + k1 > nr - 1 || j + k2 < 0 || j + k2 > nc - 1) continue;
            const float nb = img[(int64_t)(i + k1) * nc + (j + k2)];
            const float d = __fsub_rn(c, nb);
            const float w = pen_weight(k);
            if (KIND == ITOMO_PENALTY_HUBER) {
                const double s = __dadd_rn((double)fabsf(__fdiv_rn(d, delta)), 1.0);
                const float pn = (float)__ddiv_rn((double)d, s);
                const float pd = (float)__ddiv_rn(1.0, s);
                if (ARITH == ITOMO_ARITH_FAST) {
                    rn = __fmaf_rn(w, pn, rn);
                    rd = __fmaf_rn(__fadd_rn(w, w), pd, rd);
                } else {
                    rn = __fadd_rn(rn, __fmul_rn(w, pn));
                    rd = __fadd_rn(rd, __fmul_rn(__fmul_rn(2.0f, w), pd));
                }
            } else if (KIND == ITOMO_PENALTY_TIKHONOV) {
                rn = (ARITH == ITOMO_ARITH_FAST) ? __fmaf_rn(__fadd_rn(w, w), d, rn)
                                                 : __fadd_rn(rn, __fmul_rn(w, __fmul_rn(2.0f, d)));
            } else {
                const float a = d >= 0.0f ? d : __fsub_rn(nb, c);
                rn = (ARITH == ITOMO_ARITH_FAST) ? __fmaf_rn(a, w, rn)
                                                 : __fadd_rn(rn, __fmul_rn(w, a));
            }
        }
    const int64_t o = off + (int64_t)i * nc + j;
    r_num[o] = rn;
    r_den[o] = KIND == ITOMO_PENALTY_HUBER ? rd : (KIND == ITOMO_PENALTY_TIKHONOV ? 2.0f : 0.0f);
}

// scipy.ndimage 'reflect' extension: index -1 -> 0, n -> n-1.
__device__ __forceinline__ int reflect1(int k, int n)
{
    return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

// First sobel pass along `axis` (weights -1, 0, 1) at (i, j), rounded to float as scipy
// stores it, computed in double.
__device__ __forceinline__ float sobel_diff(const float* __restrict__ x, int nr, int nc, int i,
                                            int j, int axis)
{
    const float lo = axis == 0 ? x[(int64_t)reflect1(i - 1, nr) * nc + j]
                               : x[(int64_t)i * nc + reflect1(j - 1, nc)];
    const float hi = axis == 0 ? x[(int64_t)reflect1(i + 1, nr) * nc + j]
                               : x[(int64_t)i * nc + reflect1(j + 1, nc)];
    return (float)__dsub_rn((double)hi, (double)lo);
}

__global__ void __launch_bounds__(kPenTile * kPenTile)
sobel_mag_kernel(const float* __restrict__ x, int nr, int nc, float* __restrict__ g)
{
    const int j = blockIdx.x * kPenTile + threadIdx.x;
    const int i = blockIdx.y * kPenTile + threadIdx.y;
    if (i >= nr || j >= nc) return;
    const int64_t off = (int64_t)blockIdx.z * nr * nc;
    x += off;
    // gx: difference along rows (axis 0), then [1, 2, 1] along columns; gy the transpose
    const float a0 = sobel_diff(x, nr, nc, i, reflect1(j - 1, nc), 0);
    const float a1 = sobel_diff(x, nr, nc, i, j, 0);
    const float a2 = sobel_diff(x, nr, nc, i, reflect1(j + 1, nc), 0);
    const float gx = (float)__dadd_rn(__dmul_rn((double)a1, 2.0), __dadd_rn((double)a0, (double)a2));
    const float b0 = sobel_diff(x, nr, nc, reflect1(i - 1, nr), j, 1);
    const float b1 = sobel_diff(x, nr, nc, i, j, 1);
    const float b2 = sobel_diff(x, nr, nc, reflect1(i + 1, nr), j, 1);
    const float gy = (float)__dadd_rn(__dmul_rn((double)b1, 2.0), __dadd_rn((double)b0, (double)b2));
    const float s = __fadd_rn(__fmul_rn(gx, gx), __fmul_rn(gy, gy));
    // correctly rounded float sqrt via double (53 >= 2*24 + 2)
    g[off + (int64_t)i * nc + j] = (float)__dsqrt_rn((double)s);
}

// x_new = x - p / q in FP64 with p = num + beta R_num, q = den + beta R_den + 2^-52;
// clip >= 0 applies max(x_new, +0).  Partials per block: [0] max of x_new before the clip,
// [1] sum (x_out - x)^2, [2] sum x^2.
template <bool PEN>
__global__ void __launch_bounds__(256)
sps_update_kernel(float* __restrict__ x, const float* __restrict__ num,
                  const float* __restrict__ den, const float* __restrict__ rn,
                  const float* __restrict__ rd, float beta, int clip, int64_t n,
                  double* __restrict__ partials, int bps)
{
    const int z = blockIdx.y;
    x += (int64_t)z * n;
    num += (int64_t)z * n;
    den += (int64_t)z * n;
    if (PEN) {
        rn += (int64_t)z * n;
        rd += (int64_t)z * n;
    }
    double mx = -INFINITY, p1 = 0.0, p2 = 0.0;
    const int64_t base = (int64_t)blockIdx.x * kEpiChunk;
    for (int k = 0; k < kEpiChunk / 256; ++k) {
        const int64_t o = base + (int64_t)k * 256 + threadIdx.x;
        if (o < n) {
            const float xo = x[o];
            double p = (double)num[o], q = (double)den[o];
            if (PEN) {
                p = __dadd_rn(p, __dmul_rn((double)beta, (double)rn[o]));
                q = __dadd_rn(q, __dmul_rn((double)beta, (double)rd[o]));
            }
            q = __dadd_rn(q, 2.220446049250313e-16);
            float xn = (float)__dsub_rn((double)xo, __ddiv_rn(p, q));
            mx = fmax(mx, (double)xn);
            if (clip && xn < 0.0f) xn = 0.0f;
            x[o] = xn;
            const double dd = (double)xn - (double)xo;
            p1 += dd * dd;
            p2 += (double)xo * (double)xo;
        }
    }
    // block reduction: max in slot 0, sums in 1 and 2 (fixed order)
    __shared__ double red[3][4];
    for (int s = 32; s > 0; s >>= 1) {
        mx = fmax(mx, __shfl_down(mx, s, 64));
        p1 += __shfl_down(p1, s, 64);
        p2 += __shfl_down(p2, s, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wid] = mx;
        red[1][wid] = p1;
        red[2][wid] = p2;
    }
    __syncthreads();
    if (threadIdx.x == 0 && partials) {
        double a = red[0][0], b = 0.0, c = 0.0;
        for (int k = 0; k < 4; ++k) {
            a = fmax(a, red[0][k]);
            b += red[1][k];
            c += red[2][k];
        }
        double* dst = partials + ((int64_t)z * bps + blockIdx.x) * 3;
        dst[0] = a;
        dst[1] = b;
        dst[2] = c;
    }
}

// Max of partials[z][b][0] over every slice and block into *out (one block).
__global__ void __launch_bounds__(256)
max_final_kernel_d(const double* __restrict__ partials, int64_t count, double* __restrict__ out)
{
    __shared__ double red[256];
    double m = -INFINITY;
    for (int64_t k = threadIdx.x; k < count; k += 256) m = fmax(m, partials[k * 3]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// np.clip(x, 0, M) after max(x, 0): only when M < 0 does anything change (every entry
// becomes M).  Every block reads M and returns at once otherwise.
__global__ void __launch_bounds__(256)
clip_fix_kernel(float* __restrict__ x, int64_t n, const double* __restrict__ m)
{
    const double mv = m[0];
    if (!(mv < 0.0)) return;
    const float f = (float)mv;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256)
        x[k] = f;
}

template <int ARITH>
static void launch_penalty(int kind, const float* img, int nr, int nc, int ns, float delta,
                           float* rn, float* rd, hipStream_t st)
{
    dim3 block(kPenTile, kPenTile), grid((nc + kPenTile - 1) / kPenTile,
                                         (nr + kPenTile - 1) / kPenTile, ns);
    if (kind == ITOMO_PENALTY_HUBER)
        penalty_kernel<ARITH, ITOMO_PENALTY_HUBER><<<grid, block, 0, st>>>(img, nr, nc, delta, rn, rd);
    else if (kind == ITOMO_PENALTY_TIKHONOV)
        penalty_kernel<ARITH, ITOMO_PENALTY_TIKHONOV><<<grid, block, 0, st>>>(img, nr, nc, delta, rn, rd);
    else
        penalty_kernel<ARITH, ITOMO_PENALTY_HAAR><<<grid, block, 0, st>>>(img, nr, nc, delta, rn, rd);
}

static int check_dims(int nr, int nc, int ns, const char* who)
{
    if (nr < 1 || nc < 1 || ns < 1 || ns > 65535 || (int64_t)nr * nc > ((int64_t)1 << 31))
        return set_error(ITOMO_EINVAL, "%s: bad image size %d x %d x %d", who, ns, nr, nc);
    return ITOMO_OK;
}

}  // namespace itomo

using namespace itomo;

extern "C" {

int itomo_penalty(const float* image, int nr, int nc, int nslices, int kind, float delta,
                  int arith, float* r_num, float* r_den, void* stream)
{
    int rc = check_dims(nr, nc, nslices, "penalty");
    if (rc) return rc;
    if (!image || !r_num || !r_den) return set_error(ITOMO_EINVAL, "penalty: null array");
    if (kind != ITOMO_PENALTY_HUBER && kind != ITOMO_PENALTY_TIKHONOV && kind != ITOMO_PENALTY_HAAR)
        return set_error(ITOMO_EINVAL, "penalty: unknown kind %d", kind);
    if (arith != ITOMO_ARITH_FAST && arith != ITOMO_ARITH_IEEE)
        return set_error(ITOMO_EINVAL, "penalty: unknown arithmetic %d", arith);
    hipStream_t st = (hipStream_t)stream;
    if (arith == ITOMO_ARITH_FAST)
        launch_penalty<ITOMO_ARITH_FAST>(kind, image, nr, nc, nslices, delta, r_num, r_den, st);
    else
        launch_penalty<ITOMO_ARITH_IEEE>(kind, image, nr, nc, nslices, delta, r_num, r_den, st);
    return check_hip(hipGetLastError(), "penalty launch");
}

int itomo_sobel_magnitude(const float* x, int nr, int nc, int nslices, float* g, void* stream)
{
    int rc = check_dims(nr, nc, nslices, "sobel_magnitude");
    if (rc) return rc;
    if (!x || !g || x == g) return set_error(ITOMO_EINVAL, "sobel_magnitude: bad arrays");
    dim3 block(kPenTile, kPenTile), grid((nc + kPenTile - 1) / kPenTile,
                                         (nr + kPenTile - 1) / kPenTile, nslices);
    sobel_mag_kernel<<<grid, block, 0, (hipStream_t)stream>>>(x, nr, nc, g);
    return check_hip(hipGetLastError(), "sobel launch");
}

int itomo_sps_step(const itomo_plan* p, float* x, const float* b, const float* w,
                   const float* den, float* work, float* img_work, int nslices, int penalty,
                   float beta, float delta, int clip, double* stats, void* stream)
{
    if (!p || !x || !b || !w || !den || !work || !img_work || nslices < 1 || nslices > 65535)
        return set_error(ITOMO_EINVAL, "sps_step: bad argument");
    if (penalty != ITOMO_PENALTY_NONE && penalty != ITOMO_PENALTY_HUBER &&
        penalty != ITOMO_PENALTY_TIKHONOV && penalty != ITOMO_PENALTY_HAAR)
        return set_error(ITOMO_EINVAL, "sps_step: unknown penalty %d", penalty);
    hipStream_t st = (hipStream_t)stream;
    const int n = p->npix;
    const int64_t n2 = (int64_t)n * n, ni = n2 * nslices;
    float* num = img_work;
    float* rn = img_work + ni;
    float* rd = img_work + 2 * ni;
    float* g = img_work + 3 * ni;
    Scratch s;
    int rc = get_scratch(*p, nslices, &s, st);
    if (rc) return rc;
    // penalty terms of the iterate entering the step (sir.py:455-530)
    if (penalty == ITOMO_PENALTY_HUBER) {
        if ((rc = itomo_sobel_magnitude(x, n, n, nslices, g, stream))) return rc;
        if ((rc = itomo_penalty(g, n, n, nslices, penalty, delta, ITOMO_ARITH_FAST, rn, rd, stream)))
            return rc;
    } else if (penalty != ITOMO_PENALTY_NONE) {
        if ((rc = itomo_penalty(x, n, n, nslices, penalty, delta, ITOMO_ARITH_FAST, rn, rd, stream)))
            return rc;
    }
    // work = (A x - b) w, stats[0] = sum work^2 (sir.py:428-441)
    EpiArgs ef{};
    ef.b = b;
    ef.w = w;
    ef.w_stride = (int64_t)p->nang * n;
    ef.partials = stats ? s.pf : nullptr;
    int bps = 0;
    rc = projector_forward_epi_any(*p, EPI_SPS_FWD, 0, x, work, nslices, ef, st, &bps);
    if (rc) return rc;
    if (stats && (rc = finalize_partials(s.pf, bps, nslices, 0, 1, stats, st))) return rc;
    // num = A^T work (sir.py:444-452)
    EpiArgs eb{};
    rc = projector_back_epi_any(*p, EPI_STORE, 0, work, num, nslices, eb, st, nullptr);
    if (rc) return rc;
    // update + clip (sir.py:535-540)
    const int ubps = (int)((n2 + kEpiChunk - 1) / kEpiChunk);
    dim3 grid(ubps, nslices);
    if (penalty == ITOMO_PENALTY_NONE)
        sps_update_kernel<false><<<grid, 256, 0, st>>>(x, num, den, nullptr, nullptr, beta, clip,
                                                       n2, s.pb, ubps);
    else
        sps_update_kernel<true><<<grid, 256, 0, st>>>(x, num, den, rn, rd, beta, clip, n2, s.pb,
                                                      ubps);
    if ((rc = check_hip(hipGetLastError(), "sps update launch"))) return rc;
    if (stats && (rc = finalize_partials(s.pb, ubps, nslices, 1, 3, stats, st))) return rc;
    if (clip) {
        max_final_kernel_d<<<1, 256, 0, st>>>(s.pb, (int64_t)ubps * nslices, s.pf);
        if ((rc = check_hip(hipGetLastError(), "sps max launch"))) return rc;
        int64_t blocks = (ni + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        clip_fix_kernel<<<(unsigned)blocks, 256, 0, st>>>(x, ni, s.pf);
        if ((rc = check_hip(hipGetLastError(), "sps clip launch"))) return rc;
    }
    return ITOMO_OK;
}

}  // extern "C"
