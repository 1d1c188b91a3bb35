// Gridding (gridrec) projector pair for gfx950: the reference's Fourier-regridding
// back projection / filtered back projection (gridrec_v4_backproj.c:64-446) and forward
// projection (gridrec_v4_forwproj.c:59-442), with this library's own FFT.
//
// Back projection, per slice:
//   1. each sinogram row, edge-padded to pdim samples (the row's end values, or zeros for
//      the first-derivative Radon transform), is FFT-ed (rows kernel below);
//   2. Cdata1/Cdata2 of every polar sample (angle n, frequency w < pdim/2): the filtered
//      coefficient at -w and its Hermitian partner (backproj.c:271-300);
//   3. polar -> Cartesian gridding as a GATHER: every grid point sums, over the angles
//      whose ray passes within the kernel window, the kernel-weighted Cdata1 of the polar
//      samples whose window holds it, and the Cdata2 of those whose mirrored window holds
//      it (the reference scatters both, backproj.c:333-350).  One writer per grid point,
//      no atomics: deterministic;
//   4. a 2-D inverse FFT of the pdim x pdim grid (rows, transpose, rows);
//   5. deapodisation, the centring shift and the crop to npix x npix (backproj.c:371-418).
// Forward projection reverses it: deapodised image on the grid, 2-D inverse FFT, a
// gather of each polar sample from its kernel window (forwproj.c:326-360, already a
// gather in the reference), the centre phase, a 1-D FFT per angle, the crop.
//
// The FFT is a mixed-radix (4, 2, 3, 5) Stockham transform of one row per block in LDS
// (two ping-pong buffers of L complex floats, L <= 4096), twiddles from a table computed
// in double on the host.  Float arithmetic throughout, as the reference's fftwf.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "itomo_internal.h"

namespace itomo {

constexpr int kFftMaxLen = 4096;
constexpr int kFftMaxStages = 16;

struct FftPlan {
    int L;
    int nst;
    int radix[kFftMaxStages];
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(__fsub_rn(__fmul_rn(a.x, b.x), __fmul_rn(a.y, b.y)),
                       __fadd_rn(__fmul_rn(a.x, b.y), __fmul_rn(a.y, b.x)));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b)
{
    return make_float2(__fadd_rn(a.x, b.x), __fadd_rn(a.y, b.y));
}
__device__ __forceinline__ float2 csub(float2 a, float2 b)
{
    return make_float2(__fsub_rn(a.x, b.x), __fsub_rn(a.y, b.y));
}
// i*sign*a
__device__ __forceinline__ float2 cisg(float2 a, float sg) { return make_float2(-sg * a.y, sg * a.x); }

// In-place DFT of r <= 5 values with W = exp(sg * 2 pi i / r).
__device__ __forceinline__ void small_dft(float2* v, int r, float sg)
{
    if (r == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (r == 4) {
        const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const float2 s13 = cadd(v[1], v[3]), d13 = cisg(csub(v[1], v[3]), sg);
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    } else if (r == 3) {
        const float h = 0.8660254037844386f;  // sin(2 pi / 3)
        const float2 t1 = cadd(v[1], v[2]);
        const float2 t2 = cisg(csub(v[1], v[2]), sg * h);
        const float2 m = make_float2(__fsub_rn(v[0].x, 0.5f * t1.x), __fsub_rn(v[0].y, 0.5f * t1.y));
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, t2);
        v[2] = csub(m, t2);
    } else {  // r == 5: direct sums with the fifth roots of unity
        const float c1 = 0.30901699437494745f, c2 = -0.8090169943749475f;
        const float s1 = 0.9510565162951535f, s2 = 0.5877852522924731f;
        const float2 w[5] = {make_float2(1.f, 0.f), make_float2(c1, sg * s1),
                             make_float2(c2, sg * s2), make_float2(c2, -sg * s2),
                             make_float2(c1, -sg * s1)};
        float2 out[5];
        for (int s = 0; s < 5; ++s) {
            float2 acc = v[0];
            for (int t = 1; t < 5; ++t) acc = cadd(acc, cmul(v[t], w[(s * t) % 5]));
            out[s] = acc;
        }
        for (int s = 0; s < 5; ++s) v[s] = out[s];
    }
}

// FFT of each of `nrows` rows of L complex values (in place), sign -1 (FFTW_FORWARD) or
// +1 (FFTW_BACKWARD), unnormalised.  tw[q] = exp(-2 pi i q / L).
__global__ void __launch_bounds__(256)
fft_rows_kernel(float2* __restrict__ data, FftPlan f, const float2* __restrict__ tw, int sign)
{
    extern __shared__ float2 fft_sm[];
    const int L = f.L;
    float2* a = fft_sm;
    float2* b = fft_sm + L;
    float2* row = data + (int64_t)blockIdx.x * L;
    const int tid = threadIdx.x, nt = blockDim.x;
    const float sg = (float)sign;
    for (int i = tid; i < L; i += nt) a[i] = row[i];
    __syncthreads();
    int m = 1;
    for (int st = 0; st < f.nst; ++st) {
        const int r = f.radix[st];
        const int Lr = L / r;
        const int tstep = L / (m * r);
        for (int j = tid; j < Lr; j += nt) {
            const int k = j % m;
            float2 v[5];
            for (int t = 0; t < r; ++t) {
                float2 x = a[j + t * Lr];
                if (t > 0 && k > 0) {
                    float2 w = tw[t * k * tstep];
                    if (sign > 0) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[t] = x;
            }
            small_dft(v, r, sg);
            const int base = (j - k) * r + k;
            for (int s = 0; s < r; ++s) b[base + s * m] = v[s];
        }
        __syncthreads();
        float2* tmp = a;
        a = b;
        b = tmp;
        m *= r;
    }
    for (int i = tid; i < L; i += nt) row[i] = a[i];
}

// out[c][r] = in[r][c] for n x n complex, one 32 x 32 tile per block (32 x 8 threads).
__global__ void __launch_bounds__(256)
ctranspose_kernel(const float2* __restrict__ in, float2* __restrict__ out, int n)
{
    __shared__ float2 tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + threadIdx.x;
        if (r < n && c < n) tile[k][threadIdx.x] = in[(int64_t)r * n + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = c0 + k, c = r0 + threadIdx.x;
        if (r < n && c < n) out[(int64_t)r * n + c] = tile[threadIdx.x][k];
    }
}

// ---- gridrec -----------------------------------------------------------------------

struct GrParams {
    int npix, nang, pdim, ph, padleft, interp, radon_degree, nlut;
    float hb, tblb;   // back projection: half window lconv/2, LUT scale 2 ltbl / lconv
    float hf, tblf;   // forward projection (forwproj.c:164 fixes lconv = 2 * 7 / pi)
};

// convolv_nn / convolv_lin (filters.c:285-296) at x = |offset| * tblspcg.
__device__ __forceinline__ float gr_kernel(const float* __restrict__ lut, int nlut, int interp,
                                           float x)
{
    if (interp == 0) {
        const int i = min((int)floorf(__fadd_rn(x, 0.5f)), nlut - 1);
        return lut[i];
    }
    const float xi = floorf(x);
    const float d = __fsub_rn(x, xi);
    const int i = min((int)xi, nlut - 2);
    return __fadd_rn(__fmul_rn(__fsub_rn(1.0f, d), lut[i]), __fmul_rn(d, lut[i + 1]));
}

// Edge-padded projection rows (backproj.c:231-266) as complex input of the row FFT.
__global__ void __launch_bounds__(256)
gr_pad_kernel(const float* __restrict__ sino, float2* __restrict__ rows, GrParams p)
{
    const int n = blockIdx.y;
    const float* s = sino + (int64_t)n * p.npix;
    float2* out = rows + (int64_t)n * p.pdim;
    const float left = p.radon_degree ? 0.0f : s[0];
    const float right = p.radon_degree ? 0.0f : s[p.npix - 1];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < p.pdim; i += gridDim.x * blockDim.x) {
        float v;
        if (i < p.padleft) v = left;
        else if (i < p.padleft + p.npix) v = s[i - p.padleft];
        else v = right;
        out[i] = make_float2(v, 0.0f);
    }
}

// Cdata1 / Cdata2 of each polar sample (backproj.c:271-300): with P the unnormalised FFT
// of the padded row, f the filter, Ctmp = P / pdim:
//   w > 0: C1 = f * Ctmp[pdim - w], C2 = conj(f) * Ctmp[w];   w = 0: C1 = Re f * Ctmp[0], C2 = 0.
__global__ void __launch_bounds__(256)
gr_cdata_kernel(const float2* __restrict__ rows, const float2* __restrict__ filt,
                float2* __restrict__ c1, float2* __restrict__ c2, GrParams p)
{
    const int n = blockIdx.y;
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= p.ph) return;
    const float2* P = rows + (int64_t)n * p.pdim;
    const float inv = 1.0f / (float)p.pdim;
    const float2 f = filt[w];
    const float2 t2 = make_float2(__fmul_rn(inv, P[w].x), __fmul_rn(inv, P[w].y));
    float2 d1, d2;
    if (w != 0) {
        const int m = p.pdim - w;
        const float2 t3 = make_float2(__fmul_rn(inv, P[m].x), __fmul_rn(inv, P[m].y));
        d1 = cmul(f, t3);
        d2 = cmul(make_float2(f.x, -f.y), t2);
    } else {
        d1 = make_float2(__fmul_rn(f.x, t2.x), __fmul_rn(f.x, t2.y));
        d2 = make_float2(0.0f, 0.0f);
    }
    c1[(int64_t)n * p.ph + w] = d1;
    c2[(int64_t)n * p.ph + w] = d2;
}

// Polar -> Cartesian gridding, gathered per grid point (a = iu row, b = iv column).
__global__ void __launch_bounds__(256)
gr_grid_back_kernel(const float2* __restrict__ c1, const float2* __restrict__ c2,
                    const float2* __restrict__ trig, const float* __restrict__ lut,
                    float2* __restrict__ H, GrParams p)
{
    const int b = blockIdx.x * 16 + (threadIdx.x & 15);
    const int a = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (a >= p.pdim || b >= p.pdim) return;
    const float cen = (float)p.ph;
    const float af = (float)a, bf = (float)b;
    const float da = __fsub_rn(af, cen), db = __fsub_rn(bf, cen);
    const bool mirror_ok = a >= 1 && b >= 1;
    const float ma = (float)(p.pdim - a), mb = (float)(p.pdim - b);
    const float h = p.hb, tbl = p.tblb;
    float2 acc = make_float2(0.0f, 0.0f);
    for (int n = 0; n < p.nang; ++n) {
        const float S = trig[n].x, C = trig[n].y;
        const float span = __fadd_rn(__fmul_rn(h, __fadd_rn(fabsf(S), fabsf(C))), 0.5f);
        const float perp = __fadd_rn(-__fmul_rn(da, S), __fmul_rn(db, C));
        if (fabsf(perp) > span) continue;
        const float par = __fadd_rn(__fmul_rn(da, C), __fmul_rn(db, S));
        // direct: the window of (n, w) holds (a, b); every such w lies within `span` of par
        {
            const int w0 = max(0, (int)ceilf(__fsub_rn(par, span)));
            const int w1 = min(p.ph - 1, (int)floorf(__fadd_rn(par, span)));
            for (int w = w0; w <= w1; ++w) {
                const float U = __fadd_rn(__fmul_rn((float)w, C), cen);
                const float V = __fadd_rn(__fmul_rn((float)w, S), cen);
                if (af < __fsub_rn(U, h) || af > __fadd_rn(U, h) || bf < __fsub_rn(V, h) ||
                    bf > __fadd_rn(V, h))
                    continue;
                const float cv = __fmul_rn(
                    gr_kernel(lut, p.nlut, p.interp, __fmul_rn(fabsf(__fsub_rn(U, af)), tbl)),
                    gr_kernel(lut, p.nlut, p.interp, __fmul_rn(fabsf(__fsub_rn(V, bf)), tbl)));
                const float2 d = c1[(int64_t)n * p.ph + w];
                acc.x = __fadd_rn(acc.x, __fmul_rn(cv, d.x));
                acc.y = __fadd_rn(acc.y, __fmul_rn(cv, d.y));
            }
        }
        // mirrored: the window of (n, w), w >= 1, holds (pdim - a, pdim - b)
        if (mirror_ok) {
            const int w0 = max(1, (int)ceilf(__fsub_rn(-par, span)));
            const int w1 = min(p.ph - 1, (int)floorf(__fadd_rn(-par, span)));
            for (int w = w0; w <= w1; ++w) {
                const float U = __fadd_rn(__fmul_rn((float)w, C), cen);
                const float V = __fadd_rn(__fmul_rn((float)w, S), cen);
                if (ma < __fsub_rn(U, h) || ma > __fadd_rn(U, h) || mb < __fsub_rn(V, h) ||
                    mb > __fadd_rn(V, h))
                    continue;
                const float cv = __fmul_rn(
                    gr_kernel(lut, p.nlut, p.interp, __fmul_rn(fabsf(__fsub_rn(U, ma)), tbl)),
                    gr_kernel(lut, p.nlut, p.interp, __fmul_rn(fabsf(__fsub_rn(V, mb)), tbl)));
                const float2 d = c2[(int64_t)n * p.ph + w];
                acc.x = __fadd_rn(acc.x, __fmul_rn(cv, d.x));
                acc.y = __fadd_rn(acc.y, __fmul_rn(cv, d.y));
            }
        }
    }
    H[(int64_t)a * p.pdim + b] = acc;
}

// Deapodisation, centring shift and crop (backproj.c:371-418).  T holds the 2-D
// transform transposed: T[k][j] = H'[j][k].
__global__ void __launch_bounds__(256)
gr_crop_kernel(const float2* __restrict__ T, const float* __restrict__ deapod,
               float* __restrict__ image, GrParams p)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= p.npix) return;
    const int j = c + p.padleft;
    const int k = (p.npix - 1 - r) + p.padleft;
    const int iu = (j + p.ph) % p.pdim, iv = (k + p.ph) % p.pdim;
    const float d = __fmul_rn(deapod[j], deapod[k]);
    image[(int64_t)r * p.npix + c] = __fmul_rn(d, T[(int64_t)iv * p.pdim + iu].x);
}

// Forward: the deapodised image on the shifted grid, zeros elsewhere (forwproj.c:212-246).
__global__ void __launch_bounds__(256)
gr_place_kernel(const float* __restrict__ image, const float* __restrict__ deapod,
                float2* __restrict__ H, GrParams p)
{
    const int iv = blockIdx.x * blockDim.x + threadIdx.x;
    const int iu = blockIdx.y;
    if (iv >= p.pdim) return;
    const int j = (iu + p.ph) % p.pdim, k = (iv + p.ph) % p.pdim;
    float v = 0.0f;
    if (j >= p.padleft && j < p.padleft + p.npix && k >= p.padleft && k < p.padleft + p.npix) {
        const int64_t idx = (int64_t)(p.npix - 1 - (k - p.padleft)) * p.npix + (j - p.padleft);
        v = __fmul_rn(__fmul_rn(deapod[j], deapod[k]), image[idx]);
    }
    H[(int64_t)iu * p.pdim + iv] = make_float2(v, 0.0f);
}

// Forward: each polar sample gathers its kernel window from the transformed grid (T[iv][iu]
// = H'[iu][iv]), then the centre phase (forwproj.c:326-385) into the row to be FFT-ed.
__global__ void __launch_bounds__(256)
gr_grid_fwd_kernel(const float2* __restrict__ T, const float2* __restrict__ trig,
                   const float* __restrict__ lut, const float2* __restrict__ phase,
                   float2* __restrict__ rows, GrParams p)
{
    const int n = blockIdx.y;
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= p.ph) return;
    const int pd = p.pdim;
    const float S = trig[n].x, C = trig[n].y;
    const float cen = (float)p.ph;
    const float U = __fadd_rn(__fmul_rn((float)w, C), cen);
    const float V = __fadd_rn(__fmul_rn((float)w, S), cen);
    const float h = p.hf, tbl = p.tblf;
    const int iul = max(0, (int)ceilf(__fsub_rn(U, h)));
    const int iuh = min(pd - 1, (int)floorf(__fadd_rn(U, h)));
    const int ivl = max(0, (int)ceilf(__fsub_rn(V, h)));
    const int ivh = min(pd - 1, (int)floorf(__fadd_rn(V, h)));
    float kv[8];
    for (int iv = ivl, k = 0; iv <= ivh && k < 8; ++iv, ++k)
        kv[k] = gr_kernel(lut, p.nlut, p.interp, __fmul_rn(fabsf(__fsub_rn(V, (float)iv)), tbl));
    float2 d1 = make_float2(0.f, 0.f), d2 = make_float2(0.f, 0.f);
    for (int iu = iul; iu <= iuh; ++iu) {
        const float ku = gr_kernel(lut, p.nlut, p.interp, __fmul_rn(fabsf(__fsub_rn(U, (float)iu)), tbl));
        for (int iv = ivl, k = 0; iv <= ivh && k < 8; ++iv, ++k) {
            const float cv = __fmul_rn(ku, kv[k]);
            const float2 x = T[(int64_t)iv * pd + iu];
            d1.x = __fadd_rn(d1.x, __fmul_rn(x.x, cv));
            d1.y = __fadd_rn(d1.y, __fmul_rn(x.y, cv));
            if (iu != 0 && iv != 0 && w != 0) {
                const float2 y = T[(int64_t)(pd - iv) * pd + (pd - iu)];
                d2.x = __fadd_rn(d2.x, __fmul_rn(y.x, cv));
                d2.y = __fadd_rn(d2.y, __fmul_rn(y.y, cv));
            }
        }
    }
    const float2 ct = phase[w];        // (CtmpR, CtmpI) of forwproj.c:363-364
    float2* row = rows + (int64_t)n * pd;
    if (w != 0) {
        row[pd - w] = cmul(ct, d1);
        row[w] = cmul(make_float2(ct.x, -ct.y), d2);
    } else {
        row[0] = cmul(ct, d1);
        row[p.ph] = make_float2(0.f, 0.f);
    }
}

// Forward: S[n][k] = Re(row[padleft + k]) / pdim (forwproj.c:396-397).
__global__ void __launch_bounds__(256)
gr_extract_kernel(const float2* __restrict__ rows, float* __restrict__ sino, GrParams p)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (k >= p.npix) return;
    const float inv = 1.0f / (float)p.pdim;
    sino[(int64_t)n * p.npix + k] = __fmul_rn(inv, rows[(int64_t)n * p.pdim + p.padleft + k].x);
}

}  // namespace itomo

using namespace itomo;

struct itomo_gridrec {
    GrParams p;
    FftPlan fft;
    float* d_lut = nullptr;
    float* d_deapod = nullptr;
    float2* d_trig = nullptr;
    float2* d_filt_at = nullptr;
    float2* d_filt_fbp = nullptr;
    float2* d_phase = nullptr;
    float2* d_tw = nullptr;
    float2* d_H = nullptr;      // pdim x pdim
    float2* d_T = nullptr;      // pdim x pdim (transpose target)
    float2* d_rows = nullptr;   // nang x pdim
    float2* d_c1 = nullptr;     // nang x pdim/2
    float2* d_c2 = nullptr;
};

namespace {

bool fft_factor(int L, FftPlan* f)
{
    f->L = L;
    f->nst = 0;
    int n = L;
    const int radices[4] = {4, 2, 3, 5};
    for (int r : radices)
        while (n % r == 0) {
            if (f->nst >= kFftMaxStages) return false;
            f->radix[f->nst++] = r;
            n /= r;
        }
    return n == 1;
}

int fft_rows(float2* data, const FftPlan& f, const float2* tw, int nrows, int sign, hipStream_t st)
{
    if (nrows <= 0) return ITOMO_OK;
    const size_t lds = 2 * (size_t)f.L * sizeof(float2);
    fft_rows_kernel<<<nrows, 256, lds, st>>>(data, f, tw, sign);
    return check_hip(hipGetLastError(), "fft rows launch");
}

// 2-D FFT of the n x n grid in H, the result transposed into T (T[k][j] = H'[j][k]).
int fft_2d_t(itomo_gridrec* g, int sign, hipStream_t st)
{
    const int n = g->p.pdim;
    int rc = fft_rows(g->d_H, g->fft, g->d_tw, n, sign, st);
    if (rc) return rc;
    dim3 tb(32, 8), tg((n + 31) / 32, (n + 31) / 32);
    ctranspose_kernel<<<tg, tb, 0, st>>>(g->d_H, g->d_T, n);
    rc = check_hip(hipGetLastError(), "transpose launch");
    if (rc) return rc;
    return fft_rows(g->d_T, g->fft, g->d_tw, n, sign, st);
}

template <typename T>
int upload(T** dst, const void* src, size_t count, const char* what)
{
    int rc = check_hip(hipMalloc((void**)dst, count * sizeof(T)), what);
    if (rc) return rc;
    if (src) rc = check_hip(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice), what);
    return rc;
}

}  // namespace

extern "C" {

int itomo_fft_rows(void* data, int L, int nrows, int sign, void* stream)
{
    if (!data || nrows < 0 || (sign != 1 && sign != -1))
        return set_error(ITOMO_EINVAL, "fft_rows: bad argument");
    FftPlan f;
    if (L < 1 || L > kFftMaxLen || !fft_factor(L, &f))
        return set_error(ITOMO_EINVAL, "fft_rows: length %d is not a product of 2, 3 and 5 "
                                       "up to %d", L, kFftMaxLen);
    std::vector<float2> tw(L);
    for (int q = 0; q < L; ++q) {
        const double a = -2.0 * M_PI * (double)q / (double)L;
        tw[q] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    float2* d_tw = nullptr;
    hipStream_t st = (hipStream_t)stream;
    int rc = check_hip(hipMalloc(&d_tw, L * sizeof(float2)), "fft_rows: hipMalloc");
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(d_tw, tw.data(), L * sizeof(float2), hipMemcpyHostToDevice, st),
                   "fft_rows: upload");
    if (!rc) rc = fft_rows((float2*)data, f, d_tw, nrows, sign, st);
    const int rc2 = check_hip(hipStreamSynchronize(st), "fft_rows: sync");
    (void)hipFree(d_tw);
    return rc ? rc : rc2;
}

int itomo_gridrec_create(itomo_gridrec** out, int npix, int nang, int pdim,
                         const float* trig_host, const float* lut_host, int nlut, int ltbl,
                         const float* deapod_host, float lconv_back, float lconv_fwd,
                         int interp, int radon_degree, const float* filt_at_host,
                         const float* filt_fbp_host, const float* phase_host)
{
    if (!out) return set_error(ITOMO_EINVAL, "gridrec_create: null output");
    *out = nullptr;
    if (npix < 2 || nang < 1 || pdim < npix || pdim % 4 != 0 || !trig_host || !lut_host ||
        nlut < 2 || ltbl < 1 || !deapod_host || !filt_at_host || !filt_fbp_host || !phase_host ||
        !(lconv_back > 0.f) || !(lconv_fwd > 0.f) || (interp != 0 && interp != 1))
        return set_error(ITOMO_EINVAL, "gridrec_create: bad argument");
    itomo_gridrec* g = new (std::nothrow) itomo_gridrec();
    if (!g) return set_error(ITOMO_ENOMEM, "gridrec_create: out of host memory");
    if (pdim > kFftMaxLen || !fft_factor(pdim, &g->fft)) {
        delete g;
        return set_error(ITOMO_EINVAL, "gridrec_create: grid size %d is not a product of 2, 3 "
                                       "and 5 up to %d", pdim, kFftMaxLen);
    }
    GrParams& p = g->p;
    p.npix = npix;
    p.nang = nang;
    p.pdim = pdim;
    p.ph = pdim / 2;
    p.padleft = (int)((pdim - npix) * 0.5);
    p.interp = interp;
    p.radon_degree = radon_degree != 0;
    p.nlut = nlut;
    p.hb = lconv_back * 0.5f;
    p.tblb = (float)(2 * ltbl) / lconv_back;
    p.hf = lconv_fwd * 0.5f;
    p.tblf = (float)(2 * ltbl) / lconv_fwd;
    std::vector<float2> tw(pdim);
    for (int q = 0; q < pdim; ++q) {
        const double a = -2.0 * M_PI * (double)q / (double)pdim;
        tw[q] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    const size_t n2 = (size_t)pdim * pdim;
    int rc = upload(&g->d_lut, lut_host, nlut, "gridrec_create: lut");
    if (!rc) rc = upload(&g->d_deapod, deapod_host, pdim, "gridrec_create: deapod");
    if (!rc) rc = upload(&g->d_trig, trig_host, nang, "gridrec_create: trig");
    if (!rc) rc = upload(&g->d_filt_at, filt_at_host, p.ph, "gridrec_create: filter");
    if (!rc) rc = upload(&g->d_filt_fbp, filt_fbp_host, p.ph, "gridrec_create: filter");
    if (!rc) rc = upload(&g->d_phase, phase_host, p.ph, "gridrec_create: phase");
    if (!rc) rc = upload(&g->d_tw, tw.data(), pdim, "gridrec_create: twiddles");
    if (!rc) rc = upload<float2>(&g->d_H, nullptr, n2, "gridrec_create: grid");
    if (!rc) rc = upload<float2>(&g->d_T, nullptr, n2, "gridrec_create: grid");
    if (!rc) rc = upload<float2>(&g->d_rows, nullptr, (size_t)nang * pdim, "gridrec_create: rows");
    if (!rc) rc = upload<float2>(&g->d_c1, nullptr, (size_t)nang * p.ph, "gridrec_create: polar");
    if (!rc) rc = upload<float2>(&g->d_c2, nullptr, (size_t)nang * p.ph, "gridrec_create: polar");
    if (rc) {
        itomo_gridrec_destroy(g);
        return rc;
    }
    *out = g;
    return ITOMO_OK;
}

int itomo_gridrec_destroy(itomo_gridrec* g)
{
    if (!g) return ITOMO_OK;
    void* ptrs[] = {g->d_lut, g->d_deapod, g->d_trig, g->d_filt_at, g->d_filt_fbp, g->d_phase,
                    g->d_tw, g->d_H, g->d_T, g->d_rows, g->d_c1, g->d_c2};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    delete g;
    return ITOMO_OK;
}

int itomo_gridrec_back(itomo_gridrec* g, const float* sino, float* image, int nslices,
                       int filtered, void* stream)
{
    if (!g || !sino || !image || nslices < 1)
        return set_error(ITOMO_EINVAL, "gridrec_back: bad argument");
    const GrParams& p = g->p;
    hipStream_t st = (hipStream_t)stream;
    const float2* filt = filtered ? g->d_filt_fbp : g->d_filt_at;
    for (int z = 0; z < nslices; ++z) {
        const float* s = sino + (int64_t)z * p.nang * p.npix;
        float* img = image + (int64_t)z * p.npix * p.npix;
        gr_pad_kernel<<<dim3((p.pdim + 255) / 256, p.nang), 256, 0, st>>>(s, g->d_rows, p);
        int rc = check_hip(hipGetLastError(), "gridrec pad launch");
        if (!rc) rc = fft_rows(g->d_rows, g->fft, g->d_tw, p.nang, -1, st);
        if (rc) return rc;
        gr_cdata_kernel<<<dim3((p.ph + 255) / 256, p.nang), 256, 0, st>>>(g->d_rows, filt,
                                                                          g->d_c1, g->d_c2, p);
        gr_grid_back_kernel<<<dim3((p.pdim + 15) / 16, (p.pdim + 15) / 16), 256, 0, st>>>(
            g->d_c1, g->d_c2, g->d_trig, g->d_lut, g->d_H, p);
        rc = check_hip(hipGetLastError(), "gridrec gridding launch");
        if (!rc) rc = fft_2d_t(g, +1, st);
        if (rc) return rc;
        gr_crop_kernel<<<dim3((p.npix + 255) / 256, p.npix), 256, 0, st>>>(g->d_T, g->d_deapod,
                                                                          img, p);
        rc = check_hip(hipGetLastError(), "gridrec crop launch");
        if (rc) return rc;
    }
    return ITOMO_OK;
}

int itomo_gridrec_forward(itomo_gridrec* g, const float* image, float* sino, int nslices,
                          void* stream)
{
    if (!g || !sino || !image || nslices < 1)
        return set_error(ITOMO_EINVAL, "gridrec_forward: bad argument");
    const GrParams& p = g->p;
    hipStream_t st = (hipStream_t)stream;
    for (int z = 0; z < nslices; ++z) {
        const float* img = image + (int64_t)z * p.npix * p.npix;
        float* s = sino + (int64_t)z * p.nang * p.npix;
        gr_place_kernel<<<dim3((p.pdim + 255) / 256, p.pdim), 256, 0, st>>>(img, g->d_deapod,
                                                                           g->d_H, p);
        int rc = check_hip(hipGetLastError(), "gridrec place launch");
        if (!rc) rc = fft_2d_t(g, +1, st);
        if (rc) return rc;
        gr_grid_fwd_kernel<<<dim3((p.ph + 255) / 256, p.nang), 256, 0, st>>>(
            g->d_T, g->d_trig, g->d_lut, g->d_phase, g->d_rows, p);
        rc = check_hip(hipGetLastError(), "gridrec gather launch");
        if (!rc) rc = fft_rows(g->d_rows, g->fft, g->d_tw, p.nang, -1, st);
        if (rc) return rc;
        gr_extract_kernel<<<dim3((p.npix + 255) / 256, p.nang), 256, 0, st>>>(g->d_rows, s, p);
        rc = check_hip(hipGetLastError(), "gridrec extract launch");
        if (rc) return rc;
    }
    return ITOMO_OK;
}

}  // extern "C"
