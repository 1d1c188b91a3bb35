// Vector kernels of the CG (cg.py:83-176), ADMM-TV (admm.py:460-604, cg.py:256-310) and
// SPS (sir.py:348-540) iterations: finite differences and their exact adjoint, the
// fused CG update steps with FP64 inner products, the ADMM multiplier step, and the
// SPS weights.  Scalars never leave the device: each step reads the FP64 sums written
// by the previous one, so a whole iteration can be queued (or captured) at once.
//
// Per-slice scalar block (double scal[4*z .. 4*z+3]):
//   [0] rr     = <r, r>           (current residual)
//   [1] pHp    = <p, H p>
//   [2] rr_new = <r_new, r_new>
//   [3] tv     = sum (G x - u)^2  (ADMM multiplier residual)
// The CG loops of the reference stop once ||r|| <= 1e-12 (cg.py:101,128-142 / 203,215-231 /
// 271,274-292; only an exactly solved system, such as an all-zero slice, gets there).  The
// device keeps that decision per slice: after the step that reaches it, rr is set to -1,
// and a negative rr makes the remaining steps of the loop leave x, r and p alone.  The next
// residual kernel (cg_dot / tv_residual / l1_residual) writes a fresh rr >= 0.
#include "epilogue.h"

namespace itomo {

constexpr int kVecBlocks = 256;   // blocks per slice of the vector kernels
constexpr double kSpacing1 = 2.220446049250313e-16;  // np.spacing(1)
constexpr double kCgStop = 1e-12;                      // cg.py:101, 203, 271 (eps)

__device__ __forceinline__ bool cg_stops(double rn) { return sqrt(rn) <= kCgStop; }

// ---- np.gradient (edge_order=1, unit spacing) and its adjoint ------------------------
// Along one axis of length n: interior (f[i+1]-f[i-1])/2, edges f[1]-f[0], f[n-1]-f[n-2].
// As a matrix D: D[0][0]=-1, D[0][1]=1; D[n-1][n-2]=-1, D[n-1][n-1]=1;
// D[a][a-1]=-1/2, D[a][a+1]=1/2 for 0<a<n-1.

__device__ __forceinline__ float grad1(const float* f, int n, int i, int64_t stride)
{
    if (i == 0) return __fsub_rn(f[stride], f[0]);
    if (i == n - 1) return __fsub_rn(f[(int64_t)(n - 1) * stride], f[(int64_t)(n - 2) * stride]);
    return __fdiv_rn(__fsub_rn(f[(int64_t)(i + 1) * stride], f[(int64_t)(i - 1) * stride]), 2.0f);
}

__device__ __forceinline__ float dcoef(int a, int b, int n)
{
    if (a == 0) return (b == 0) ? -1.0f : ((b == 1) ? 1.0f : 0.0f);
    if (a == n - 1) return (b == n - 1) ? 1.0f : ((b == n - 2) ? -1.0f : 0.0f);
    return (b == a + 1) ? 0.5f : ((b == a - 1) ? -0.5f : 0.0f);
}

// (D^T g)[b] along one axis: sum over a in {b-1, b, b+1}.
__device__ __forceinline__ float gradT1(const float* g, int n, int b, int64_t stride)
{
    float s = 0.0f;
    for (int a = max(b - 1, 0); a <= min(b + 1, n - 1); ++a) {
        const float c = dcoef(a, b, n);
        if (c != 0.0f) s = __fadd_rn(s, __fmul_rn(c, g[(int64_t)a * stride]));
    }
    return s;
}

// G^T (g0, g1) at (i, j) = (D^T g0)[i] along rows + (D^T g1)[j] along columns
// (cg.py:66-71: dot(g1, Gx) + dot(Gy, g0)).
__device__ __forceinline__ float gradT_at(const float* g0, const float* g1, int n, int i, int j)
{
    const float a = gradT1(g0 + j, n, i, n);
    const float b = gradT1(g1 + (int64_t)i * n, n, j, 1);
    return __fadd_rn(b, a);
}

__global__ void __launch_bounds__(256)
grad_kernel(const float* __restrict__ x, float* __restrict__ g, int n)
{
    const int z = blockIdx.y;
    const int64_t n2 = (int64_t)n * n;
    x += (int64_t)z * n2;
    float* g0 = g + (int64_t)z * 2 * n2;
    float* g1 = g0 + n2;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n2;
         k += (int64_t)gridDim.x * 256) {
        const int i = (int)(k / n), j = (int)(k - (int64_t)i * n);
        g0[k] = grad1(x + j, n, i, n);
        g1[k] = grad1(x + (int64_t)i * n, n, j, 1);
    }
}

__global__ void __launch_bounds__(256)
gradT_kernel(const float* __restrict__ g, float* __restrict__ out, int n)
{
    const int z = blockIdx.y;
    const int64_t n2 = (int64_t)n * n;
    const float* g0 = g + (int64_t)z * 2 * n2;
    const float* g1 = g0 + n2;
    out += (int64_t)z * n2;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n2;
         k += (int64_t)gridDim.x * 256) {
        const int i = (int)(k / n), j = (int)(k - (int64_t)i * n);
        out[k] = gradT_at(g0, g1, n, i, j);
    }
}

// ---- CG pieces -----------------------------------------------------------------------

// Hp += lambd1 * p + mu * G^T g   (TV terms when g != nullptr), partial <p, Hp>.
__global__ void __launch_bounds__(256)
cg_hp_kernel(float* __restrict__ Hp, const float* __restrict__ p, const float* __restrict__ g,
             int n, float lambd1, float mu, double* __restrict__ partials)
{
    const int z = blockIdx.y;
    const int64_t n2 = (int64_t)n * n;
    Hp += (int64_t)z * n2;
    p += (int64_t)z * n2;
    const float* g0 = g ? g + (int64_t)z * 2 * n2 : nullptr;
    double s = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n2;
         k += (int64_t)gridDim.x * 256) {
        float h = Hp[k];
        if (g0) {
            // cg.py:278: At(A p) + lambd1 * p + mu * Gt(G(p))
            const int i = (int)(k / n), j = (int)(k - (int64_t)i * n);
            h = __fadd_rn(__fadd_rn(h, __fmul_rn(lambd1, p[k])),
                          __fmul_rn(mu, gradT_at(g0, g0 + n2, n, i, j)));
            Hp[k] = h;
        } else if (lambd1 != 0.0f) {
            // cg.py:211: At(A p) + (lambd1 + mu) * p, the sum passed as lambd1
            h = __fadd_rn(h, __fmul_rn(lambd1, p[k]));
            Hp[k] = h;
        }
        s += (double)p[k] * (double)h;
    }
    block_partials3(s, 0.0, 0.0, partials + ((int64_t)z * kVecBlocks + blockIdx.x) * 3);
}

// x += a p; r -= a Hp with a = float(rr / (pHp + eps)) (cg.py:126-140); partial <r, r>.
__global__ void __launch_bounds__(256)
cg_update_kernel(float* __restrict__ x, float* __restrict__ r, const float* __restrict__ p,
                 const float* __restrict__ Hp, int64_t n2, const double* __restrict__ scal,
                 double* __restrict__ partials)
{
    const int z = blockIdx.y;
    const double rr = scal[4 * z + 0], pHp = scal[4 * z + 1];
    const float a = (float)(rr / (pHp + kSpacing1));
    x += (int64_t)z * n2;
    r += (int64_t)z * n2;
    p += (int64_t)z * n2;
    Hp += (int64_t)z * n2;
    double s = 0.0;
    // rr < 0: this slice's loop has stopped (x and r stay, <r, r> reads 0)
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; rr >= 0.0 && k < n2;
         k += (int64_t)gridDim.x * 256) {
        x[k] = __fadd_rn(x[k], __fmul_rn(a, p[k]));
        const float rn = __fsub_rn(r[k], __fmul_rn(a, Hp[k]));
        r[k] = rn;
        s += (double)rn * (double)rn;
    }
    block_partials3(s, 0.0, 0.0, partials + ((int64_t)z * kVecBlocks + blockIdx.x) * 3);
}

// p = r + float(rr_new / (rr [+ eps])) p  (cg.py:149 with eps, cg.py:302 without).
__global__ void __launch_bounds__(256)
cg_direction_kernel(float* __restrict__ p, const float* __restrict__ r, int64_t n2,
                    const double* __restrict__ scal, int with_eps)
{
    const int z = blockIdx.y;
    const double rr = scal[4 * z + 0], rn = scal[4 * z + 2];
    if (rr < 0.0 || cg_stops(rn)) return;   // stopped before this direction update
    const float bta = (float)(rn / (rr + (with_eps ? kSpacing1 : 0.0)));
    p += (int64_t)z * n2;
    r += (int64_t)z * n2;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n2;
         k += (int64_t)gridDim.x * 256)
        p[k] = __fadd_rn(r[k], __fmul_rn(bta, p[k]));
}

// per-slice <a, b> partials
__global__ void __launch_bounds__(256)
vdot_kernel(const float* __restrict__ a, const float* __restrict__ b, int64_t n,
            double* __restrict__ partials)
{
    const int z = blockIdx.y;
    a += (int64_t)z * n;
    b += (int64_t)z * n;
    double s = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256)
        s += (double)a[k] * (double)b[k];
    block_partials3(s, 0.0, 0.0, partials + ((int64_t)z * kVecBlocks + blockIdx.x) * 3);
}

// rr <- rr_new, or the stop marker -1 once ||r_new|| <= 1e-12 (kept until the next
// residual kernel).
__global__ void scal_shift_kernel(double* scal, int nslices)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nslices || scal[4 * z + 0] < 0.0) return;
    const double rn = scal[4 * z + 2];
    scal[4 * z + 0] = cg_stops(rn) ? -1.0 : rn;
}

// Copy the finalized sums (stats layout [z*3 + 0]) into scal[4*z + slot].
__global__ void scal_store_kernel(const double* tmp, double* scal, int slot, int nslices)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z < nslices) scal[4 * z + slot] = tmp[3 * z];
}

// r = atb + G^T(mu u - alpha) - (atax + lambd1 x + G^T gx)   (cg.py:258), partial <r, r>.
__global__ void __launch_bounds__(256)
tv_residual_kernel(float* __restrict__ r, const float* __restrict__ atb,
                   const float* __restrict__ atax, const float* __restrict__ x,
                   const float* __restrict__ gx, const float* __restrict__ v, int n,
                   float lambd1, double* __restrict__ partials)
{
    const int z = blockIdx.y;
    const int64_t n2 = (int64_t)n * n;
    r += (int64_t)z * n2;
    atb += (int64_t)z * n2;
    atax += (int64_t)z * n2;
    x += (int64_t)z * n2;
    const float* gx0 = gx + (int64_t)z * 2 * n2;
    const float* v0 = v + (int64_t)z * 2 * n2;
    double s = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n2;
         k += (int64_t)gridDim.x * 256) {
        const int i = (int)(k / n), j = (int)(k - (int64_t)i * n);
        const float lhs = __fadd_rn(atb[k], gradT_at(v0, v0 + n2, n, i, j));
        const float rhs = __fadd_rn(__fadd_rn(atax[k], __fmul_rn(lambd1, x[k])),
                                    gradT_at(gx0, gx0 + n2, n, i, j));
        const float rv = __fsub_rn(lhs, rhs);
        r[k] = rv;
        s += (double)rv * (double)rv;
    }
    block_partials3(s, 0.0, 0.0, partials + ((int64_t)z * kVecBlocks + blockIdx.x) * 3);
}

// v = mu u - alpha (the right-hand side term of the inner solve, admm.py:532-534).
__global__ void __launch_bounds__(256)
tv_rhs_kernel(float* __restrict__ v, const float* __restrict__ u, const float* __restrict__ al,
              int64_t n, float mu)
{
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256)
        v[k] = __fsub_rn(__fmul_rn(mu, u[k]), al[k]);
}

// ADMM multiplier step (admm.py:460-461, 565-593): g = G x;
// u = soft(g + alpha/mu, lambd2/mu) = max(0, y - t) - max(0, -y - t) (utils.py:214-216);
// alpha += g - u; partial sum (g - u)^2.
__global__ void __launch_bounds__(256)
admm_tv_update_kernel(const float* __restrict__ x, float* __restrict__ u,
                      float* __restrict__ al, int n, float mu, float lambd2,
                      double* __restrict__ partials)
{
    const int z = blockIdx.y;
    const int64_t n2 = (int64_t)n * n;
    x += (int64_t)z * n2;
    u += (int64_t)z * 2 * n2;
    al += (int64_t)z * 2 * n2;
    const float t = __fdiv_rn(lambd2, mu);
    double s = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n2;
         k += (int64_t)gridDim.x * 256) {
        const int i = (int)(k / n), j = (int)(k - (int64_t)i * n);
        const float g[2] = {grad1(x + j, n, i, n), grad1(x + (int64_t)i * n, n, j, 1)};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int64_t kk = (int64_t)c * n2 + k;
            const float y = __fadd_rn(g[c], __fdiv_rn(al[kk], mu));
            const float un = __fsub_rn(fmaxf(0.0f, __fsub_rn(y, t)),
                                       fmaxf(0.0f, __fsub_rn(-y, t)));
            u[kk] = un;
            const float rr = __fsub_rn(g[c], un);
            al[kk] = __fadd_rn(al[kk], rr);
            s += (double)rr * (double)rr;
        }
    }
    block_partials3(s, 0.0, 0.0, partials + ((int64_t)z * kVecBlocks + blockIdx.x) * 3);
}

// Lasso-L1 inner-solve residual (cg.py:191): r = (A^T b + v) - (A^T A x + c x) with
// c = lambd1 + mu; partial <r, r>.
__global__ void __launch_bounds__(256)
l1_residual_kernel(float* __restrict__ r, const float* __restrict__ atb,
                   const float* __restrict__ atax, const float* __restrict__ x,
                   const float* __restrict__ v, int64_t n2, float c,
                   double* __restrict__ partials)
{
    const int z = blockIdx.y;
    const int64_t off = (int64_t)z * n2;
    double s = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n2;
         k += (int64_t)gridDim.x * 256) {
        const float rv = __fsub_rn(__fadd_rn(atb[off + k], v[off + k]),
                                   __fadd_rn(atax[off + k], __fmul_rn(c, x[off + k])));
        r[off + k] = rv;
        s += (double)rv * (double)rv;
    }
    block_partials3(s, 0.0, 0.0, partials + ((int64_t)z * kVecBlocks + blockIdx.x) * 3);
}

// Lasso-L1 multiplier step (admm.py:380-386): u = soft(x + alpha/mu, lambd2/mu)
// (utils.hard_thres, utils.py:214-216), alpha += x - u; partial sum (x - u)^2.
__global__ void __launch_bounds__(256)
admm_l1_update_kernel(const float* __restrict__ x, float* __restrict__ u,
                      float* __restrict__ al, int64_t n2, float mu, float lambd2,
                      double* __restrict__ partials)
{
    const int z = blockIdx.y;
    const int64_t off = (int64_t)z * n2;
    const float t = __fdiv_rn(lambd2, mu);
    double s = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n2;
         k += (int64_t)gridDim.x * 256) {
        const float xv = x[off + k];
        const float y = __fadd_rn(xv, __fdiv_rn(al[off + k], mu));
        const float un = __fsub_rn(fmaxf(0.0f, __fsub_rn(y, t)), fmaxf(0.0f, __fsub_rn(-y, t)));
        u[off + k] = un;
        const float rr = __fsub_rn(xv, un);
        al[off + k] = __fadd_rn(al[off + k], rr);
        s += (double)rr * (double)rr;
    }
    block_partials3(s, 0.0, 0.0, partials + ((int64_t)z * kVecBlocks + blockIdx.x) * 3);
}

// ---- SPS setup (sir.py:366-384) --------------------------------------------------------
__global__ void __launch_bounds__(256)
max_kernel(const float* __restrict__ b, int64_t n, float* __restrict__ part)
{
    __shared__ float red[256];
    float m = -INFINITY;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256)
        m = fmaxf(m, b[k]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) max_final_kernel(float* part, int nb)
{
    __shared__ float red[256];
    float m = -INFINITY;
    for (int k = threadIdx.x; k < nb; k += 256) m = fmaxf(m, part[k]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[0] = red[0];
}

// w = exp(-b / b_max) (sir.py:369); aux = w * A(1) (sir.py:375, A(1) shared by slices).
__global__ void __launch_bounds__(256)
sps_w_kernel(const float* __restrict__ b, float* __restrict__ w, float* __restrict__ aux,
             const float* __restrict__ a1, const float* __restrict__ bmax, int64_t per,
             int64_t n)
{
    const float bm = bmax[0];
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256) {
        const float a = a1[k % per];  // read before w[k] is written: a1 may alias w
        // exp in double rounded to float: numpy's float32 exp (the reference) is not
        // correctly rounded either; this is, to within the double exp's error
        const float wv = (float)exp(-(double)__fdiv_rn(b[k], bm));
        w[k] = wv;
        aux[k] = __fmul_rn(wv, a);
    }
}

static unsigned grid1(int64_t n)
{
    int64_t b = (n + 255) / 256;
    return (unsigned)(b < kVecBlocks ? (b < 1 ? 1 : b) : kVecBlocks);
}

}  // namespace itomo

using namespace itomo;

static int vec_check(int npix, int nslices, const char* who)
{
    if (npix < 2 || nslices < 1 || nslices > 65535)
        return set_error(ITOMO_EINVAL, "%s: need npix >= 2 and 1 <= nslices <= 65535", who);
    return ITOMO_OK;
}

extern "C" {

int itomo_grad(const float* x, int npix, int nslices, float* g, void* stream)
{
    int rc = vec_check(npix, nslices, "grad");
    if (rc) return rc;
    if (!x || !g) return set_error(ITOMO_EINVAL, "grad: null array");
    dim3 grid(grid1((int64_t)npix * npix), nslices);
    grad_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(x, g, npix);
    return check_hip(hipGetLastError(), "grad launch");
}

int itomo_grad_adjoint(const float* g, int npix, int nslices, float* out, void* stream)
{
    int rc = vec_check(npix, nslices, "grad_adjoint");
    if (rc) return rc;
    if (!g || !out) return set_error(ITOMO_EINVAL, "grad_adjoint: null array");
    dim3 grid(grid1((int64_t)npix * npix), nslices);
    gradT_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(g, out, npix);
    return check_hip(hipGetLastError(), "grad_adjoint launch");
}

static int finalize_into(double* partials, int nslices, double* scal, int slot, hipStream_t st)
{
    // partial sums live at partials[0 .. nslices*kVecBlocks*3); the finalized per-slice
    // sums go to a small tail area, then into the scalar slot
    double* tmp = partials + (size_t)nslices * kVecBlocks * 3;
    int rc = finalize_partials(partials, kVecBlocks, nslices, 0, 1, tmp, st);
    if (rc) return rc;
    scal_store_kernel<<<(nslices + 255) / 256, 256, 0, st>>>(tmp, scal, slot, nslices);
    return check_hip(hipGetLastError(), "scal store launch");
}

int itomo_cg_hp(float* Hp, const float* p, const float* g, int npix, int nslices,
                float lambd1, float mu, double* scal, double* work, void* stream)
{
    int rc = vec_check(npix, nslices, "cg_hp");
    if (rc) return rc;
    if (!Hp || !p || !scal || !work) return set_error(ITOMO_EINVAL, "cg_hp: null array");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(kVecBlocks, nslices);
    cg_hp_kernel<<<grid, 256, 0, st>>>(Hp, p, g, npix, lambd1, mu, work);
    if ((rc = check_hip(hipGetLastError(), "cg_hp launch"))) return rc;
    return finalize_into(work, nslices, scal, 1, st);
}

int itomo_cg_update(float* x, float* r, const float* p, const float* Hp, int64_t n,
                    int nslices, double* scal, double* work, void* stream)
{
    if (!x || !r || !p || !Hp || !scal || !work || n < 1 || nslices < 1 || nslices > 65535)
        return set_error(ITOMO_EINVAL, "cg_update: bad argument");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(kVecBlocks, nslices);
    cg_update_kernel<<<grid, 256, 0, st>>>(x, r, p, Hp, n, scal, work);
    int rc = check_hip(hipGetLastError(), "cg_update launch");
    if (rc) return rc;
    return finalize_into(work, nslices, scal, 2, st);
}

int itomo_cg_direction(float* p, const float* r, int64_t n, int nslices, int with_eps,
                       double* scal, void* stream)
{
    if (!p || !r || !scal || n < 1 || nslices < 1 || nslices > 65535)
        return set_error(ITOMO_EINVAL, "cg_direction: bad argument");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(grid1(n), nslices);
    cg_direction_kernel<<<grid, 256, 0, st>>>(p, r, n, scal, with_eps);
    int rc = check_hip(hipGetLastError(), "cg_direction launch");
    if (rc) return rc;
    scal_shift_kernel<<<(nslices + 255) / 256, 256, 0, st>>>(scal, nslices);
    return check_hip(hipGetLastError(), "scal shift launch");
}

int itomo_cg_dot(const float* a, const float* b, int64_t n, int nslices, int slot,
                 double* scal, double* work, void* stream)
{
    if (!a || !b || !scal || !work || n < 1 || nslices < 1 || nslices > 65535 || slot < 0 ||
        slot > 3)
        return set_error(ITOMO_EINVAL, "cg_dot: bad argument");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(kVecBlocks, nslices);
    vdot_kernel<<<grid, 256, 0, st>>>(a, b, n, work);
    int rc = check_hip(hipGetLastError(), "cg_dot launch");
    if (rc) return rc;
    return finalize_into(work, nslices, scal, slot, st);
}

int itomo_tv_residual(float* r, const float* atb, const float* atax, const float* x,
                      const float* gx, const float* v, int npix, int nslices, float lambd1,
                      double* scal, double* work, void* stream)
{
    int rc = vec_check(npix, nslices, "tv_residual");
    if (rc) return rc;
    if (!r || !atb || !atax || !x || !gx || !v || !scal || !work)
        return set_error(ITOMO_EINVAL, "tv_residual: null array");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(kVecBlocks, nslices);
    tv_residual_kernel<<<grid, 256, 0, st>>>(r, atb, atax, x, gx, v, npix, lambd1, work);
    if ((rc = check_hip(hipGetLastError(), "tv_residual launch"))) return rc;
    return finalize_into(work, nslices, scal, 0, st);
}

int itomo_tv_rhs(float* v, const float* u, const float* alpha, int64_t n, float mu,
                 void* stream)
{
    if (!v || !u || !alpha || n < 1) return set_error(ITOMO_EINVAL, "tv_rhs: bad argument");
    tv_rhs_kernel<<<grid1(n), 256, 0, (hipStream_t)stream>>>(v, u, alpha, n, mu);
    return check_hip(hipGetLastError(), "tv_rhs launch");
}

int itomo_l1_residual(float* r, const float* atb, const float* atax, const float* x,
                      const float* v, int npix, int nslices, float c, double* scal,
                      double* work, void* stream)
{
    int rc = vec_check(npix, nslices, "l1_residual");
    if (rc) return rc;
    if (!r || !atb || !atax || !x || !v || !scal || !work)
        return set_error(ITOMO_EINVAL, "l1_residual: null array");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(kVecBlocks, nslices);
    l1_residual_kernel<<<grid, 256, 0, st>>>(r, atb, atax, x, v, (int64_t)npix * npix, c, work);
    if ((rc = check_hip(hipGetLastError(), "l1_residual launch"))) return rc;
    return finalize_into(work, nslices, scal, 0, st);
}

int itomo_admm_l1_update(const float* x, float* u, float* alpha, int npix, int nslices,
                         float mu, float lambd2, double* scal, double* work, void* stream)
{
    int rc = vec_check(npix, nslices, "admm_l1_update");
    if (rc) return rc;
    if (!x || !u || !alpha || !scal || !work)
        return set_error(ITOMO_EINVAL, "admm_l1_update: null array");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(kVecBlocks, nslices);
    admm_l1_update_kernel<<<grid, 256, 0, st>>>(x, u, alpha, (int64_t)npix * npix, mu, lambd2,
                                                work);
    if ((rc = check_hip(hipGetLastError(), "admm_l1_update launch"))) return rc;
    return finalize_into(work, nslices, scal, 3, st);
}

int itomo_admm_tv_update(const float* x, float* u, float* alpha, int npix, int nslices,
                         float mu, float lambd2, double* scal, double* work, void* stream)
{
    int rc = vec_check(npix, nslices, "admm_tv_update");
    if (rc) return rc;
    if (!x || !u || !alpha || !scal || !work)
        return set_error(ITOMO_EINVAL, "admm_tv_update: null array");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(kVecBlocks, nslices);
    admm_tv_update_kernel<<<grid, 256, 0, st>>>(x, u, alpha, npix, mu, lambd2, work);
    if ((rc = check_hip(hipGetLastError(), "admm_tv_update launch"))) return rc;
    return finalize_into(work, nslices, scal, 3, st);
}

}  // extern "C"

// ---- SPS ---------------------------------------------------------------------------------

namespace itomo {
int projector_forward_epi_any(const Plan& p, int epi, int ones, const float* image, float* out,
                              int nslices, const EpiArgs& e, hipStream_t st, int* bps);
int projector_back_epi_any(const Plan& p, int epi, int ones, const float* sino, float* x,
                           int nslices, const EpiArgs& e, hipStream_t st, int* bps);
}

extern "C" {

// w = exp(-b / max b), den = A^T(w * A(1)) per slice (sir.py:366-384).
// work: nslices x nang x npix scratch (receives w * A(1)).  bmax_work: DEVICE float[1024].
int itomo_sps_weights(const itomo_plan* p, const float* b, int nslices, float* w, float* den,
                      float* work, float* bmax_work, void* stream)
{
    if (!p || !b || !w || !den || !work || !bmax_work || nslices < 1 || nslices > 65535)
        return set_error(ITOMO_EINVAL, "sps_weights: bad argument");
    hipStream_t st = (hipStream_t)stream;
    const int64_t ns = (int64_t)p->nang * p->npix, n = ns * nslices;
    max_kernel<<<1024, 256, 0, st>>>(b, n, bmax_work);
    int rc = check_hip(hipGetLastError(), "max launch");
    if (rc) return rc;
    max_final_kernel<<<1, 256, 0, st>>>(bmax_work, 1024);
    if ((rc = check_hip(hipGetLastError(), "max final launch"))) return rc;
    // A(1) goes to w's first slice; slices are then finished from the last to the first,
    // so A(1) is read before slice 0 overwrites it (element by element, same thread)
    EpiArgs e{};
    rc = projector_forward_epi_any(*p, EPI_STORE, 1, nullptr, w, 1, e, st, nullptr);
    if (rc) return rc;
    for (int z = nslices - 1; z >= 0; --z) {
        const int64_t off = (int64_t)z * ns;
        sps_w_kernel<<<grid1(ns), 256, 0, st>>>(b + off, w + off, work + off, w, bmax_work, ns,
                                                ns);
        if ((rc = check_hip(hipGetLastError(), "sps w launch"))) return rc;
    }
    // den = A^T(w * A(1)), plain back projection per slice
    return projector_back_epi_any(*p, EPI_STORE, 0, work, den, nslices, e, st, nullptr);
}

// One SPS step without penalty (beta = 0): work = (A x - b) * w; x -= A^T(work) / (den + eps)
// with the quotient in FP64 (sir.py:428-452, 535).  stats as for SIRT, [0] = sum work^2.
int itomo_sps_iterate(const itomo_plan* p, float* x, const float* b, const float* w,
                      const float* den, float* work, int nslices, double* stats, void* stream)
{
    if (!p || !x || !b || !w || !den || !work || nslices < 1 || nslices > 65535)
        return set_error(ITOMO_EINVAL, "sps_iterate: bad argument");
    hipStream_t st = (hipStream_t)stream;
    Scratch s;
    int rc = get_scratch(*p, nslices, &s, st);
    if (rc) return rc;
    EpiArgs ef{};
    ef.b = b;
    ef.w = w;
    ef.w_stride = (int64_t)p->nang * p->npix;
    ef.partials = stats ? s.pf : nullptr;
    int bps = 0;
    rc = projector_forward_epi_any(*p, EPI_SPS_FWD, 0, x, work, nslices, ef, st, &bps);
    if (rc) return rc;
    if (stats && (rc = finalize_partials(s.pf, bps, nslices, 0, 1, stats, st))) return rc;
    EpiArgs eb{};
    eb.w = den;
    eb.w_stride = (int64_t)p->npix * p->npix;
    eb.partials = stats ? s.pb : nullptr;
    rc = projector_back_epi_any(*p, EPI_SPS_BACK, 0, work, x, nslices, eb, st, &bps);
    if (rc) return rc;
    if (stats) rc = finalize_partials(s.pb, bps, nslices, 1, 3, stats, st);
    return rc;
}

}  // extern "C"
