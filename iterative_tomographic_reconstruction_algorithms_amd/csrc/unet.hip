// Learned warm start: the kernels of a small U-Net (3x3 convolutions, ReLU, 2x2 max
// pooling, nearest-neighbour upsampling with skip concatenation, a 1x1 head added to
// the input image), forward and backward, plus the MSE loss and Adam.
//
// Activations and weights are BF16 (raw 16-bit storage, round to nearest even), every
// product is accumulated in FP32 on the matrix cores: a 3x3 convolution is an implicit
// GEMM, M = pixels, N = output channels, K = 9 * input channels, evaluated with
// v_mfma_f32_16x16x32_bf16.  Layout is NHWC (channels fastest) so that the 8 K-values a
// lane feeds to one MFMA are 8 consecutive channels of one pixel and one tap: a single
// 16-byte load.  Master weights, gradients and Adam moments are FP32.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (64 lanes, l = lane):
//   A (16 x 32, rows i, K k):  lane l holds A[l % 16][8 * (l / 16) + e], e = 0..7
//   B (32 x 16, K k, cols j):  lane l holds B[8 * (l / 16) + e][l % 16]
//   D (16 x 16):               lane l holds D[4 * (l / 16) + r][l % 16], r = 0..3
#include <algorithm>

#include "itomo_internal.h"

namespace itomo {
namespace {

typedef unsigned short bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

union Frag {
    uint4 u;
    bf16x8 v;
};

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ bf16_t f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 zero8()
{
    Frag f;
    f.u = make_uint4(0, 0, 0, 0);
    return f.v;
}

__device__ __forceinline__ bf16x8 load8(const bf16_t* p)
{
    Frag f;
    f.u = *(const uint4*)p;
    return f.v;
}

// 8 BF16 values of one 16-byte piece as floats, and back (round to nearest even)
__device__ __forceinline__ void unpack8(uint4 u, float* f)
{
    const unsigned w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        f[2 * i] = __uint_as_float(w[i] << 16);
        f[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
    }
}

__device__ __forceinline__ uint4 pack8(const float* f)
{
    unsigned w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = f2bf(f[2 * i]) | ((unsigned)f2bf(f[2 * i + 1]) << 16);
    return make_uint4(w[0], w[1], w[2], w[3]);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---- 3x3 convolution, stride 1, zero padding 1 -------------------------------------
//
// Y[p][co] = act(bias[co] + sum_{tap, ci} X[p + off(tap)][ci] * Wt[co][tap][ci]),
// optionally multiplied by (mask[p][co] > 0).  Block = 4 waves = 128 pixels x 32 output
// channels; each wave a 32 x 32 tile (2 x 2 MFMA tiles).  VEC: Cin % 8 == 0, so a lane's
// 8 K-values share one tap and are one 16-byte load; otherwise element loads.
template <bool VEC>
__global__ void __launch_bounds__(256)
conv3x3_kernel(const bf16_t* __restrict__ X, const bf16_t* __restrict__ Wt,
               const float* __restrict__ bias, const bf16_t* __restrict__ mask,
               bf16_t* __restrict__ Y, int N, int H, int W, int Cin, int Cout, int relu)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int64_t P = (int64_t)N * H * W;
    const int64_t p0 = (int64_t)blockIdx.x * 128 + wave * 32;
    const int co0 = blockIdx.y * 32;

    int ya[2], xa[2];
    int64_t rowa[2];  // (n*H) * W base of the pixel's image
    bool va[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int64_t p = p0 + t * 16 + li;
        va[t] = p < P;
        const int64_t pp = va[t] ? p : 0;
        const int64_t hw = (int64_t)H * W;
        const int64_t n = pp / hw;
        const int64_t r = pp - n * hw;
        ya[t] = (int)(r / W);
        xa[t] = (int)(r - (int64_t)ya[t] * W);
        rowa[t] = n * hw;
    }
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int K = 9 * Cin;
    for (int k0 = 0; k0 < K; k0 += 32) {
        const int kk = k0 + 8 * lg;
        bf16x8 a[2], b[2];
        if (VEC) {
            const bool kin = kk < K;
            const int tap = kin ? kk / Cin : 0, ci = kk - tap * Cin;
            const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int y = ya[t] + dy, x = xa[t] + dx;
                const bool ok = kin && va[t] && y >= 0 && y < H && x >= 0 && x < W;
                a[t] = ok ? load8(X + (rowa[t] + (int64_t)y * W + x) * Cin + ci) : zero8();
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int co = co0 + t * 16 + li;
                b[t] = (kin && co < Cout) ? load8(Wt + ((int64_t)co * 9 + tap) * Cin + ci)
                                          : zero8();
            }
        } else {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                bf16_t ea[8], eb[8];
                const int co = co0 + t * 16 + li;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = kk + e;
                    ea[e] = 0;
                    eb[e] = 0;
                    if (k < K) {
                        const int tap = k / Cin, ci = k - tap * Cin;
                        const int y = ya[t] + tap / 3 - 1, x = xa[t] + tap % 3 - 1;
                        if (va[t] && y >= 0 && y < H && x >= 0 && x < W)
                            ea[e] = X[(rowa[t] + (int64_t)y * W + x) * Cin + ci];
                        if (co < Cout) eb[e] = Wt[(int64_t)co * K + k];
                    }
                }
                Frag fa, fb;
                fa.u = make_uint4(ea[0] | ((unsigned)ea[1] << 16), ea[2] | ((unsigned)ea[3] << 16),
                                  ea[4] | ((unsigned)ea[5] << 16), ea[6] | ((unsigned)ea[7] << 16));
                fb.u = make_uint4(eb[0] | ((unsigned)eb[1] << 16), eb[2] | ((unsigned)eb[3] << 16),
                                  eb[4] | ((unsigned)eb[5] << 16), eb[6] | ((unsigned)eb[7] << 16));
                a[t] = fa.v;
                b[t] = fb.v;
            }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
    }

#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int co = co0 + j * 16 + li;
            if (co >= Cout) continue;
            const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t p = p0 + i * 16 + 4 * lg + r;
                if (p >= P) continue;
                float v = acc[i][j][r] + bv;
                if (relu) v = fmaxf(v, 0.0f);
                const int64_t o = p * Cout + co;
                if (mask && !(bf2f(mask[o]) > 0.0f)) v = 0.0f;
                Y[o] = f2bf(v);
            }
        }
}

// The same convolution for one input channel (the network's first layer, K = 9): one
// thread per pixel, FP32 sums, eight output channels per 16-byte store; the weights sit in
// LDS as floats.  Cout % 8 == 0, at most 64.
__global__ void __launch_bounds__(256)
conv3x3_c1_kernel(const bf16_t* __restrict__ X, const bf16_t* __restrict__ Wt,
                  const float* __restrict__ bias, const bf16_t* __restrict__ mask,
                  bf16_t* __restrict__ Y, int N, int H, int W, int Cout, int relu)
{
    __shared__ float sw[64 * 9];
    __shared__ float sb[64];
    for (int k = threadIdx.x; k < Cout * 9; k += blockDim.x) sw[k] = bf2f(Wt[k]);
    for (int k = threadIdx.x; k < Cout; k += blockDim.x) sb[k] = bias ? bias[k] : 0.0f;
    __syncthreads();
    const int64_t P = (int64_t)N * H * W;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int64_t hw = (int64_t)H * W;
    const int64_t n = p / hw;
    const int r = (int)(p - n * hw);
    const int y = r / W, x = r - (r / W) * W;
    float v[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
        v[t] = (yy >= 0 && yy < H && xx >= 0 && xx < W) ? bf2f(X[n * hw + (int64_t)yy * W + xx])
                                                        : 0.0f;
    }
    for (int c0 = 0; c0 < Cout; c0 += 8) {
        Frag m;
        if (mask) m.u = *(const uint4*)(mask + p * Cout + c0);
        const bf16_t* mh = (const bf16_t*)&m.u;
        unsigned short o[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float a = sb[c0 + e];
#pragma unroll
            for (int t = 0; t < 9; ++t) a = fmaf(v[t], sw[(c0 + e) * 9 + t], a);
            if (relu) a = fmaxf(a, 0.0f);
            if (mask && !(bf2f(mh[e]) > 0.0f)) a = 0.0f;
            o[e] = f2bf(a);
        }
        *(uint4*)(Y + p * Cout + c0) =
            make_uint4(o[0] | ((unsigned)o[1] << 16), o[2] | ((unsigned)o[3] << 16),
                       o[4] | ((unsigned)o[5] << 16), o[6] | ((unsigned)o[7] << 16));
    }
}

// The same convolution for Cin % 32 == 0, tiled through LDS.  Block = 4 waves = a tile of
// (4 WM) x 16 pixels of one image x 16 WN output channels; wave w computes tile rows
// w WM .. w WM + WM - 1 (one 16-pixel MFMA M-tile each) for all 16 WN channels.  Input
// channels are taken 32 at a time: the chunk's (4 WM + 2) x 18 halo of input pixels and
// its 16 WN x 9 x 32 weights are staged in LDS (every global load of a chunk is issued
// before any is stored), then the 9 taps run from LDS in order: a wave reads its WM + 2
// halo rows at the three column offsets once per chunk (A-fragments, 16 bytes per lane;
// each serves up to three taps), and per tap WN B-fragments, for WM x WN MFMAs per tap.  An LDS row (one staged pixel, or one
// output channel and tap of the weights) holds 32 channels padded to 40 (80 bytes), so
// the 16-byte reads of 16 consecutive lanes spread over the banks.  Pixels outside the
// image read zeros (the padding of the convolution); outputs outside it are not stored.
constexpr int kConvChunk = 32;   // input channels per staged chunk
constexpr int kConvRow = 40;     // bf16 per LDS row (32 + 8 padding)

template <int WM, int WN>
__global__ void __launch_bounds__(256)
conv3x3_tiled_kernel(const bf16_t* __restrict__ X, const bf16_t* __restrict__ Wt,
                     const float* __restrict__ bias, const bf16_t* __restrict__ mask,
                     bf16_t* __restrict__ Y, int H, int W, int Cin, int Cout, int relu,
                     int co_blocks, int vec_out)
{
    constexpr int TH = 4 * WM, HY = TH + 2, HX = 18, NB = 16 * WN;
    constexpr int NXP = HY * HX * 4;       // 16-byte pieces of a staged halo chunk
    constexpr int NWP = NB * 9 * 4;        // 16-byte pieces of a staged weight chunk
    constexpr int PX = (NXP + 255) / 256, PW = (NWP + 255) / 256;
    __shared__ __attribute__((aligned(16))) bf16_t sX[HY * HX * kConvRow];
    __shared__ __attribute__((aligned(16))) bf16_t sW[NB * 9 * kConvRow];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int n = blockIdx.z / co_blocks, cob = blockIdx.z - n * co_blocks;
    const int y0 = blockIdx.y * TH, x0 = blockIdx.x * 16, co0 = cob * NB;
    const int64_t img = (int64_t)n * H * W;

    f32x4 acc[WM][WN];
#pragma unroll
    for (int t = 0; t < WM; ++t)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[t][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int c0 = 0; c0 < Cin; c0 += kConvChunk) {
        uint4 vx[PX], vw[PW];
#pragma unroll
        for (int r = 0; r < PX; ++r) {
            const int k = tid + r * 256;
            vx[r] = make_uint4(0, 0, 0, 0);
            if (k < NXP) {
                const int pix = k >> 2, q = k & 3;
                const int hy = pix / HX, hx = pix - hy * HX;
                const int y = y0 - 1 + hy, x = x0 - 1 + hx;
                if (y >= 0 && y < H && x >= 0 && x < W)
                    vx[r] = *(const uint4*)(X + (img + (int64_t)y * W + x) * Cin + c0 + 8 * q);
            }
        }
#pragma unroll
        for (int r = 0; r < PW; ++r) {
            const int k = tid + r * 256;
            vw[r] = make_uint4(0, 0, 0, 0);
            if (k < NWP) {
                const int row = k >> 2, q = k & 3;      // row = local channel * 9 + tap
                const int co = co0 + row / 9;
                if (co < Cout)
                    vw[r] = *(const uint4*)(Wt + ((int64_t)co0 * 9 + row) * Cin + c0 + 8 * q);
            }
        }
        __syncthreads();   // the previous chunk is consumed
#pragma unroll
        for (int r = 0; r < PX; ++r) {
            const int k = tid + r * 256;
            if (k < NXP) *(uint4*)&sX[(k >> 2) * kConvRow + 8 * (k & 3)] = vx[r];
        }
#pragma unroll
        for (int r = 0; r < PW; ++r) {
            const int k = tid + r * 256;
            if (k < NWP) *(uint4*)&sW[(k >> 2) * kConvRow + 8 * (k & 3)] = vw[r];
        }
        __syncthreads();
        // the wave's (WM + 2) halo rows at the three column offsets serve all nine taps:
        // read once, then the taps in order 0..8
        bf16x8 a[3][WM + 2];
#pragma unroll
        for (int tx = 0; tx < 3; ++tx)
#pragma unroll
            for (int t = 0; t < WM + 2; ++t)
                a[tx][t] = load8(&sX[((wave * WM + t) * HX + li + tx) * kConvRow + 8 * lg]);
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int ty = tap / 3, tx = tap - 3 * (tap / 3);
            bf16x8 b[WN];
#pragma unroll
            for (int j = 0; j < WN; ++j)
                b[j] = load8(&sW[((j * 16 + li) * 9 + tap) * kConvRow + 8 * lg]);
#pragma unroll
            for (int t = 0; t < WM; ++t)
#pragma unroll
                for (int j = 0; j < WN; ++j) acc[t][j] = mfma(a[tx][t + ty], b[j], acc[t][j]);
        }
    }

    if (vec_out) {
        // the tile's outputs (bias and ReLU applied, rounded to BF16) go through LDS, in
        // the weights' buffer, as [pixel][channel] rows of NB + 8; then 16-byte pieces
        // (8 channels of one pixel) are masked and written whole
        constexpr int OS = NB + 8;
        static_assert(TH * 16 * OS <= NB * 9 * kConvRow, "output staging fits");
        __syncthreads();   // every wave is done with the chunk's weights
#pragma unroll
        for (int j = 0; j < WN; ++j) {
            const int col = j * 16 + li;
            const float bv = (bias && co0 + col < Cout) ? bias[co0 + col] : 0.0f;
#pragma unroll
            for (int t = 0; t < WM; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    float v = acc[t][j][r] + bv;
                    if (relu) v = fmaxf(v, 0.0f);
                    sW[((wave * WM + t) * 16 + 4 * lg + r) * OS + col] = f2bf(v);
                }
        }
        __syncthreads();
        constexpr int NP = TH * 16 * (NB / 8);
        for (int k = tid; k < NP; k += 256) {
            const int px = k / (NB / 8), cg = k - px * (NB / 8);
            const int y = y0 + (px >> 4), x = x0 + (px & 15), co = co0 + 8 * cg;
            if (y >= H || x >= W || co >= Cout) continue;
            Frag v;
            v.u = *(const uint4*)&sW[px * OS + 8 * cg];
            const int64_t o = (img + (int64_t)y * W + x) * Cout + co;
            if (mask) {
                Frag m;
                m.u = *(const uint4*)(mask + o);
                bf16_t* vh = (bf16_t*)&v.u;
                const bf16_t* mh = (const bf16_t*)&m.u;
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    if (!(bf2f(mh[e]) > 0.0f)) vh[e] = 0;
            }
            *(uint4*)(Y + o) = v.u;
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < WN; ++j) {
        const int co = co0 + j * 16 + li;
        if (co >= Cout) continue;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int t = 0; t < WM; ++t) {
            const int y = y0 + wave * WM + t;
            if (y >= H) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int x = x0 + 4 * lg + r;
                if (x >= W) continue;
                float v = acc[t][j][r] + bv;
                if (relu) v = fmaxf(v, 0.0f);
                const int64_t o = (img + (int64_t)y * W + x) * Cout + co;
                if (mask && !(bf2f(mask[o]) > 0.0f)) v = 0.0f;
                Y[o] = f2bf(v);
            }
        }
    }
}

// Wf[ci][tap][co] = Wt[co][8 - tap][ci]: the weights of the transposed convolution
// (gradient with respect to the input of a 3x3, pad-1 convolution).
__global__ void flip_kernel(const bf16_t* __restrict__ Wt, bf16_t* __restrict__ Wf, int Cin,
                            int Cout)
{
    const int64_t n = (int64_t)Cout * 9 * Cin;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int co = (int)(k % Cout);
        const int64_t r = k / Cout;
        const int tap = (int)(r % 9);
        const int ci = (int)(r / 9);
        Wf[k] = Wt[((int64_t)co * 9 + (8 - tap)) * Cin + ci];
    }
}

// dW[co][tap][ci] += sum_p dY[p][co] * X[p + off(tap)][ci] over the pixels
// [blockIdx.z * chunk, +chunk).  GEMM M = Cout, N = 9*Cin, K = pixels; block = 64 x 64
// outputs (4 waves, 2 x 2 of 32 x 32), 32 pixels per step staged transposed in LDS so
// that each lane's 8 K-values (8 consecutive pixels) are one 16-byte LDS read.
__global__ void __launch_bounds__(256)
wgrad_kernel(const bf16_t* __restrict__ X, const bf16_t* __restrict__ dY,
             float* __restrict__ dW, int N, int H, int W, int Cin, int Cout, int64_t chunk)
{
    __shared__ __attribute__((aligned(16))) bf16_t sA[64][40];  // [co][pixel]
    __shared__ __attribute__((aligned(16))) bf16_t sB[64][40];  // [col][pixel]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int wr = wave >> 1, wc = wave & 1;
    const int ncol = 9 * Cin;
    const int col0 = blockIdx.x * 64, co0 = blockIdx.y * 64;
    const int64_t P = (int64_t)N * H * W, hw = (int64_t)H * W;
    const int64_t pb = (int64_t)blockIdx.z * chunk;
    const int64_t pe = min(P, pb + chunk);

    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // staging role: pixel sp of the 32, channel/column group sg*8 .. +7
    const int sp = tid >> 3, sg = (tid & 7) * 8;
    for (int64_t ps = pb; ps < pe; ps += 32) {
        const int64_t p = ps + sp;
        const bool pin = p < pe;
        // dY[p][co0 + sg .. +7] -> sA[sg + e][sp]
        {
            bf16_t v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int co = co0 + sg + e;
                v[e] = (pin && co < Cout) ? dY[p * Cout + co] : (bf16_t)0;
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) sA[sg + e][sp] = v[e];
        }
        // X[p + off(tap)][ci] for columns col0 + sg .. +7 -> sB[sg + e][sp]
        {
            int64_t n = 0;
            int y = 0, x = 0;
            if (pin) {
                n = p / hw;
                const int64_t r = p - n * hw;
                y = (int)(r / W);
                x = (int)(r - (int64_t)y * W);
            }
            bf16_t v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int col = col0 + sg + e;
                v[e] = 0;
                if (pin && col < ncol) {
                    const int tap = col / Cin, ci = col - tap * Cin;
                    const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                        v[e] = X[((n * H + yy) * W + xx) * Cin + ci];
                }
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) sB[sg + e][sp] = v[e];
        }
        __syncthreads();
        bf16x8 a[2], b[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            a[t] = load8(&sA[wr * 32 + t * 16 + li][8 * lg]);
            b[t] = load8(&sB[wc * 32 + t * 16 + li][8 * lg]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = col0 + wc * 32 + j * 16 + li;
            if (col >= ncol) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = co0 + wr * 32 + i * 16 + 4 * lg + r;
                if (co < Cout) atomicAdd(&dW[(int64_t)co * ncol + col], acc[i][j][r]);
            }
        }
}

// The weight gradient for Cin % 32 == 0 and Cout % 32 == 0, tiled through LDS.  GEMM
// M = output channels (a block's 16 WM), N = 9 taps x 32 input channels (one chunk),
// K = pixels in tiles of 8 x 16 (128).  The MFMA wants 8 consecutive K values per lane,
// i.e. 8 pixels of one channel, while activations are stored channels-fastest, so both
// operands are transposed (8 x 8 blocks in registers) while they are staged: dY of the
// tile as sD[channel][pixel],
// and the tile's 10 x 18 input halo three times, shifted by the tap's column offset, as
// sX[shift][channel][halo row][16 columns], so that every fragment read is one aligned
// 16-byte LDS read.  Wave (wr, wc) computes the WM 16-channel tiles from channel 16 WM wr,
// for taps 0..8 and input channels 16 wc .. 16 wc + 15.  Blocks take pixel tiles
// blockIdx.z, blockIdx.z + gridDim.z, ... and add their sums to dW atomically; the bias
// gradient comes out of the same pass (see bias_here).
constexpr int kWgTileRows = 8;
constexpr int kWgDRow = 136;              // bf16 per sD row: 128 pixels + 8 padding
constexpr int kWgXRow = 24;               // bf16 per sX row: 16 columns + 8 padding
constexpr int kWgXChan = 10 * kWgXRow + 8;  // bf16 per channel of one shifted copy

template <int WM>
__global__ void __launch_bounds__(256)
wgrad_tiled_kernel(const bf16_t* __restrict__ X, const bf16_t* __restrict__ dY,
                   float* __restrict__ dW, float* __restrict__ db, int N, int H, int W,
                   int Cin, int Cout, int ci_chunks)
{
    constexpr int MB = 32 * WM;            // output channels per block (2 waves x WM tiles)
    __shared__ __attribute__((aligned(16))) bf16_t sD[MB * kWgDRow];
    __shared__ __attribute__((aligned(16))) bf16_t sX[3 * 32 * kWgXChan];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int wr = wave >> 1, wc = wave & 1;
    const int cob = blockIdx.x / ci_chunks, cic = blockIdx.x - cob * ci_chunks;
    const int co0 = cob * MB, c0 = cic * 32;
    const int tiles_x = (W + 15) / 16, tiles_y = (H + kWgTileRows - 1) / kWgTileRows;
    const int64_t ntiles = (int64_t)N * tiles_y * tiles_x;

    f32x4 acc[WM][9];
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int t = 0; t < 9; ++t) acc[i][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // the bias gradient: the same A fragments against a B of ones (every column of D is
    // the sum over the K pixels), in the blocks of the first input-channel chunk, by the
    // waves of the first channel half
    const bool bias_here = db && cic == 0 && wc == 0;
    Frag ones;
    ones.u = make_uint4(0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u);
    f32x4 accb[WM];
#pragma unroll
    for (int i = 0; i < WM; ++i) accb[i] = f32x4{0.f, 0.f, 0.f, 0.f};

    // Staging in units of 8 channels: a dY unit is 8 pixels of one tile row (tile row,
    // half row, channel group), a halo unit 10 pixels of one halo row (channel group, halo
    // row, half row: halo columns 8 h .. 8 h + 9), which give the half rows of all three
    // shifted copies (the copy shifted by sh holds halo columns sh .. sh + 15).  A thread
    // has at most one unit (16 MB / 8 + 80 <= 256); it loads the 16-byte channel pieces of
    // its pixels into registers, transposes them and writes 16-byte LDS rows of 8 pixels of
    // one channel.  The next tile's loads are issued before the current tile's MFMAs, so
    // their latency hides behind them.  Zeros outside the image.
    constexpr int UD = 16 * (MB / 8);     // dY units
    constexpr int UX = 4 * 10 * 2;        // halo units
    static_assert(UD + UX <= 256, "one staging unit per thread");
    const int u = tid;
    const bool is_d = u < UD, is_x = !is_d && u < UD + UX;
    // dY unit: tile row ur, half uh, channel group ug; halo unit: halo row ur, half uh,
    // channel group ug
    const int ur = is_d ? ((u & 15) >> 1) : (((u - UD) >> 1) % 10);
    const int uh = is_d ? (u & 1) : ((u - UD) & 1);
    const int ug = is_d ? (u >> 4) : (((u - UD) >> 1) / 10);
    // Loads are unconditional, from addresses clamped into the image, and the pieces of
    // pixels outside it are zeroed when the registers are consumed (bit i of `valid`):
    // a zero written into a register while a load to it is in flight would make the wave
    // wait for every load it has issued.
    const bf16_t* src = is_d ? dY : X;
    const int cst = is_d ? Cout : Cin;
    const int cof = (is_d ? co0 : c0) + 8 * ug;
    const int oy = is_d ? ur : ur - 1, ox = is_d ? 8 * uh : 8 * uh - 1;
    uint4 in[10];
    unsigned valid = 0;
    auto load_unit = [&](int64_t tl) {
        const int n = (int)(tl / ((int64_t)tiles_y * tiles_x));
        const int rem = (int)(tl - (int64_t)n * tiles_y * tiles_x);
        const int y = (rem / tiles_x) * kWgTileRows + oy, x0 = (rem % tiles_x) * 16 + ox;
        const bool rowok = y >= 0 && y < H;
        const int64_t row = (int64_t)n * H + min(max(y, 0), H - 1);
        valid = is_d ? 0x300u : 0u;   // a dY unit has 8 pixels
        if (is_d || is_x) {
#pragma unroll
            for (int i = 0; i < 10; ++i) {
                if (i < 8 || is_x) {
                    const int x = x0 + i;
                    in[i] = *(const uint4*)(src + (row * W + min(max(x, 0), W - 1)) * cst + cof);
                    valid |= (rowok && x >= 0 && x < W) ? (1u << i) : 0u;
                }
            }
        }
    };
    // word m of channel c's row from pixels a (low half) and b (high half)
    auto pair = [](const uint4& a, const uint4& b, int c) -> unsigned {
        const unsigned lo = ((const unsigned*)&a)[c >> 1], hi = ((const unsigned*)&b)[c >> 1];
        return (c & 1) ? ((lo >> 16) | (hi & 0xffff0000u)) : ((lo & 0xffffu) | (hi << 16));
    };
    if ((int64_t)blockIdx.z < ntiles) load_unit(blockIdx.z);

    for (int64_t tile = blockIdx.z; tile < ntiles; tile += gridDim.z) {
        __syncthreads();   // the previous tile is consumed
        if (valid != 0x3ffu) {
#pragma unroll
            for (int i = 0; i < 10; ++i)
                if (!((valid >> i) & 1u)) in[i] = make_uint4(0, 0, 0, 0);
        }
        if (is_d) {
            bf16_t* dst = &sD[(8 * ug) * kWgDRow + ur * 16 + 8 * uh];
#pragma unroll
            for (int c = 0; c < 8; ++c)
                *(uint4*)(dst + c * kWgDRow) =
                    make_uint4(pair(in[0], in[1], c), pair(in[2], in[3], c),
                               pair(in[4], in[5], c), pair(in[6], in[7], c));
        } else if (is_x) {
            bf16_t* dst = &sX[(8 * ug) * kWgXChan + ur * kWgXRow + 8 * uh];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                // shifts 0 and 2 share the even pairs, shift 1 takes the odd ones
                unsigned e[5], o[4];
#pragma unroll
                for (int m = 0; m < 5; ++m) e[m] = pair(in[2 * m], in[2 * m + 1], c);
#pragma unroll
                for (int m = 0; m < 4; ++m) o[m] = pair(in[2 * m + 1], in[2 * m + 2], c);
                *(uint4*)(dst + c * kWgXChan) = make_uint4(e[0], e[1], e[2], e[3]);
                *(uint4*)(dst + (32 + c) * kWgXChan) = make_uint4(o[0], o[1], o[2], o[3]);
                *(uint4*)(dst + (64 + c) * kWgXChan) = make_uint4(e[1], e[2], e[3], e[4]);
            }
        }
        __syncthreads();
        if (tile + gridDim.z < ntiles) load_unit(tile + gridDim.z);
        // K steps of 32 pixels: tile rows 2 s, 2 s + 1; lane group lg holds row
        // 2 s + (lg >> 1), columns 8 (lg & 1) .. + 7
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int row = 2 * s + (lg >> 1), col = 8 * (lg & 1);
            bf16x8 a[WM];
#pragma unroll
            for (int i = 0; i < WM; ++i)
                a[i] = load8(&sD[((wr * WM + i) * 16 + li) * kWgDRow + row * 16 + col]);
            if (bias_here) {
#pragma unroll
                for (int i = 0; i < WM; ++i) accb[i] = mfma(a[i], ones.v, accb[i]);
            }
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int dy = t / 3, sh = t - 3 * (t / 3);
                const bf16x8 b =
                    load8(&sX[(sh * 32 + 16 * wc + li) * kWgXChan + (row + dy) * kWgXRow + col]);
#pragma unroll
                for (int i = 0; i < WM; ++i) acc[i][t] = mfma(a[i], b, acc[i][t]);
            }
        }
    }
    // acc[i][t] lane: channel co0 + (wr WM + i) 16 + 4 lg + r, tap t, input channel
    // c0 + 16 wc + li
    const int ncol = 9 * Cin;
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = co0 + (wr * WM + i) * 16 + 4 * lg + r;
                if (co < Cout)
                    atomicAdd(&dW[(int64_t)co * ncol + t * Cin + c0 + 16 * wc + li], acc[i][t][r]);
            }
    if (bias_here && li == 0) {
#pragma unroll
        for (int i = 0; i < WM; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = co0 + (wr * WM + i) * 16 + 4 * lg + r;
                if (co < Cout) atomicAdd(&db[co], accb[i][r]);
            }
    }
}

// The weight and bias gradients of the one-input-channel layer (Cout == 32): dW[co][tap]
// and db[co].  Work items are 64-pixel segments of one image row; a block takes segments
// blockIdx.x, + gridDim.x, ...  Per segment the three input rows it needs (66 columns,
// zeros outside the image) are staged in LDS as floats; thread (pixel px, channel group
// cq) loads the pixel's 8 dY channels 8 cq .. 8 cq + 7 as one 16-byte piece (four
// threads read a pixel's 64 bytes, a wave 1 KB in a row) and adds g * x for the 9 taps
// and g for the bias in FP32.  At the end the lanes of equal channel group are added with
// shuffles, the four waves in LDS, and one atomic per block and output.
__global__ void __launch_bounds__(256)
wgrad_c1_kernel(const bf16_t* __restrict__ X, const bf16_t* __restrict__ dY,
                float* __restrict__ dW, float* __restrict__ db, int N, int H, int W)
{
    __shared__ float sx[3][68];
    __shared__ float red[4][32 * 10];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int px = tid >> 2, cq = tid & 3;
    float a[8][10];
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int t = 0; t < 10; ++t) a[c][t] = 0.0f;
    const int spr = (W + 63) / 64;
    const int64_t nseg = (int64_t)N * H * spr;
    for (int64_t sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
        const int64_t row = sg / spr;               // n * H + y
        const int x0 = (int)(sg - row * spr) * 64;
        const int y = (int)(row % H);
        const int x = x0 + px;
        // the pixel's gradient first, so that its latency overlaps the staging
        uint4 gv = make_uint4(0, 0, 0, 0);
        if (x < W) gv = *(const uint4*)(dY + (row * W + x) * 32 + 8 * cq);
        __syncthreads();   // the previous segment's rows are consumed
        if (tid < 3 * 66) {
            const int r = tid / 66, c = tid - r * 66;
            const int yy = y + r - 1, xx = x0 + c - 1;
            sx[r][c] = (yy >= 0 && yy < H && xx >= 0 && xx < W)
                           ? bf2f(X[(row + r - 1) * W + xx]) : 0.0f;
        }
        __syncthreads();
        float g[8];
        unpack8(gv, g);
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const float v = sx[t / 3][px + t % 3];
#pragma unroll
            for (int c = 0; c < 8; ++c) a[c][t] = fmaf(g[c], v, a[c][t]);
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) a[c][9] += g[c];
    }
#pragma unroll
    for (int off = 4; off < 64; off <<= 1)
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
            for (int t = 0; t < 10; ++t) a[c][t] += __shfl_xor(a[c][t], off);
    if (lane < 4) {
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
            for (int t = 0; t < 10; ++t) red[wave][(8 * lane + c) * 10 + t] = a[c][t];
    }
    __syncthreads();
    for (int k = tid; k < 32 * 10; k += blockDim.x) {
        const float sum = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
        const int c = k / 10, t = k - c * 10;
        if (t < 9)
            atomicAdd(&dW[c * 9 + t], sum);
        else if (db)
            atomicAdd(&db[c], sum);
    }
}

// db[c] += sum over the block's pixels of dY[p][c]; thread c of a block of C threads.
__global__ void bias_grad_kernel(const bf16_t* __restrict__ dY, float* __restrict__ db,
                                 int64_t P, int C, int64_t chunk)
{
    const int c = threadIdx.x;
    if (c >= C) return;
    const int64_t pb = (int64_t)blockIdx.x * chunk, pe = min(P, pb + chunk);
    float s = 0.0f;
    for (int64_t p = pb; p < pe; ++p) s += bf2f(dY[p * C + c]);
    atomicAdd(&db[c], s);
}

// ---- pooling, upsampling, concatenation ---------------------------------------------

// Y = 2x2 max pool of X (H, W even); idx = position of the maximum (first on ties).
__global__ void pool_fwd_kernel(const bf16_t* __restrict__ X, bf16_t* __restrict__ Y,
                                unsigned char* __restrict__ idx, int N, int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n_out = (int64_t)N * Ho * Wo * C;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_out;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(k % C);
        int64_t r = k / C;
        const int x = (int)(r % Wo);
        r /= Wo;
        const int y = (int)(r % Ho);
        const int64_t n = r / Ho;
        float best = 0.0f;
        int bi = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int yy = 2 * y + (q >> 1), xx = 2 * x + (q & 1);
            const float v = bf2f(X[((n * H + yy) * W + xx) * C + c]);
            if (q == 0 || v > best) {
                best = v;
                bi = q;
            }
        }
        Y[k] = f2bf(best);
        idx[k] = (unsigned char)bi;
    }
}

// dX = (skip + scatter of dY to the argmax) * (Xm > 0); skip may be null.
__global__ void pool_bwd_kernel(const bf16_t* __restrict__ dY, const unsigned char* __restrict__ idx,
                                const bf16_t* __restrict__ skip, const bf16_t* __restrict__ Xm,
                                bf16_t* __restrict__ dX, int N, int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n_in = (int64_t)N * H * W * C;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_in;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(k % C);
        int64_t r = k / C;
        const int x = (int)(r % W);
        r /= W;
        const int y = (int)(r % H);
        const int64_t n = r / H;
        const int64_t o = ((n * Ho + y / 2) * Wo + x / 2) * C + c;
        float g = skip ? bf2f(skip[k]) : 0.0f;
        if (idx[o] == (unsigned char)((y & 1) * 2 + (x & 1))) g += bf2f(dY[o]);
        if (!(bf2f(Xm[k]) > 0.0f)) g = 0.0f;
        dX[k] = f2bf(g);
    }
}

// Y[n][y][x] = concat(A[n][y/2][x/2][0:Ca], B[n][y][x][0:Cb]); Y has H x W, A H/2 x W/2.
__global__ void upcat_fwd_kernel(const bf16_t* __restrict__ A, const bf16_t* __restrict__ B,
                                 bf16_t* __restrict__ Y, int N, int H, int W, int Ca, int Cb)
{
    const int C = Ca + Cb;
    const int64_t n_out = (int64_t)N * H * W * C;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_out;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(k % C);
        const int64_t pix = k / C;
        const int x = (int)(pix % W);
        const int64_t r = pix / W;
        const int y = (int)(r % H);
        const int64_t n = r / H;
        Y[k] = c < Ca ? A[((n * (H / 2) + y / 2) * (W / 2) + x / 2) * Ca + c]
                      : B[pix * Cb + (c - Ca)];
    }
}

// dA[n][y][x][c] = (sum of dY over the 2x2 block, channel c) * (A > 0);
// dB[p][c] = dY[p][Ca + c].
__global__ void upcat_bwd_kernel(const bf16_t* __restrict__ dY, const bf16_t* __restrict__ A,
                                 bf16_t* __restrict__ dA, bf16_t* __restrict__ dB, int N, int H,
                                 int W, int Ca, int Cb)
{
    const int C = Ca + Cb, Ho = H / 2, Wo = W / 2;
    const int64_t na = (int64_t)N * Ho * Wo * Ca, nb = (int64_t)N * H * W * Cb;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < na + nb;
         k += (int64_t)gridDim.x * blockDim.x) {
        if (k < na) {
            const int c = (int)(k % Ca);
            int64_t r = k / Ca;
            const int x = (int)(r % Wo);
            r /= Wo;
            const int y = (int)(r % Ho);
            const int64_t n = r / Ho;
            float g = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int yy = 2 * y + (q >> 1), xx = 2 * x + (q & 1);
                g += bf2f(dY[((n * H + yy) * W + xx) * C + c]);
            }
            if (!(bf2f(A[k]) > 0.0f)) g = 0.0f;
            dA[k] = f2bf(g);
        } else {
            const int64_t kb = k - na;
            const int c = (int)(kb % Cb);
            const int64_t pix = kb / Cb;
            dB[kb] = dY[pix * C + Ca + c];
        }
    }
}

// The same four layout kernels on 16-byte pieces of 8 channels, for channel counts that
// are multiples of 8 and 16-byte aligned tensors: one block row per image row of the
// batch (blockIdx.y = n * rows + y, 32-bit index arithmetic), threads over (column,
// channel group).  Each element is computed exactly as in the kernels above.
__global__ void __launch_bounds__(256)
pool_fwd_vec_kernel(const bf16_t* __restrict__ X, bf16_t* __restrict__ Y,
                    unsigned char* __restrict__ idx, int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2, G = C / 8;
    const int row = blockIdx.y, n = row / Ho, y = row - n * Ho;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < Wo * G; t += gridDim.x * blockDim.x) {
        const int x = t / G, g = t - x * G;
        float best[8], v[8];
        unsigned char bi[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int yy = 2 * y + (q >> 1), xx = 2 * x + (q & 1);
            unpack8(*(const uint4*)(X + (((int64_t)n * H + yy) * W + xx) * C + 8 * g), v);
#pragma unroll
            for (int e = 0; e < 8; ++e)
                if (q == 0 || v[e] > best[e]) {
                    best[e] = v[e];
                    bi[e] = (unsigned char)q;
                }
        }
        const int64_t o = ((int64_t)row * Wo + x) * C + 8 * g;
        *(uint4*)(Y + o) = pack8(best);
        *(uint2*)(idx + o) = make_uint2(bi[0] | (bi[1] << 8) | (bi[2] << 16) | ((unsigned)bi[3] << 24),
                                        bi[4] | (bi[5] << 8) | (bi[6] << 16) | ((unsigned)bi[7] << 24));
    }
}

__global__ void __launch_bounds__(256)
pool_bwd_vec_kernel(const bf16_t* __restrict__ dY, const unsigned char* __restrict__ idx,
                    const bf16_t* __restrict__ skip, const bf16_t* __restrict__ Xm,
                    bf16_t* __restrict__ dX, int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2, G = C / 8;
    const int row = blockIdx.y, n = row / H, y = row - n * H;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < W * G; t += gridDim.x * blockDim.x) {
        const int x = t / G, g = t - x * G;
        const int64_t k = ((int64_t)row * W + x) * C + 8 * g;
        const int64_t o = (((int64_t)n * Ho + y / 2) * Wo + x / 2) * C + 8 * g;
        float gs[8], dy[8], xm[8];
        if (skip) unpack8(*(const uint4*)(skip + k), gs);
        else
#pragma unroll
            for (int e = 0; e < 8; ++e) gs[e] = 0.0f;
        unpack8(*(const uint4*)(dY + o), dy);
        unpack8(*(const uint4*)(Xm + k), xm);
        const uint2 iv = *(const uint2*)(idx + o);
        const unsigned char mine = (unsigned char)((y & 1) * 2 + (x & 1));
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const unsigned char ie = (unsigned char)(((e < 4 ? iv.x : iv.y) >> (8 * (e & 3))) & 0xff);
            float gv = gs[e];
            if (ie == mine) gv += dy[e];
            if (!(xm[e] > 0.0f)) gv = 0.0f;
            gs[e] = gv;
        }
        *(uint4*)(dX + k) = pack8(gs);
    }
}

__global__ void __launch_bounds__(256)
upcat_fwd_vec_kernel(const bf16_t* __restrict__ A, const bf16_t* __restrict__ B,
                     bf16_t* __restrict__ Y, int H, int W, int Ca, int Cb)
{
    const int C = Ca + Cb, G = C / 8;
    const int row = blockIdx.y, n = row / H, y = row - n * H;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < W * G; t += gridDim.x * blockDim.x) {
        const int x = t / G, g = t - x * G;
        const int64_t pix = (int64_t)row * W + x;
        const int c = 8 * g;
        const uint4 v = c < Ca ? *(const uint4*)(A + (((int64_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * Ca + c)
                               : *(const uint4*)(B + pix * Cb + (c - Ca));
        *(uint4*)(Y + pix * C + c) = v;
    }
}

// blockIdx.y < N * Ho: the low-resolution rows (dA); above: the full-resolution rows (dB).
__global__ void __launch_bounds__(256)
upcat_bwd_vec_kernel(const bf16_t* __restrict__ dY, const bf16_t* __restrict__ A,
                     bf16_t* __restrict__ dA, bf16_t* __restrict__ dB, int N, int H, int W,
                     int Ca, int Cb)
{
    const int C = Ca + Cb, Ho = H / 2, Wo = W / 2;
    if ((int)blockIdx.y < N * Ho) {
        const int row = blockIdx.y, n = row / Ho, y = row - n * Ho, G = Ca / 8;
        for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < Wo * G;
             t += gridDim.x * blockDim.x) {
            const int x = t / G, g = t - x * G;
            float acc[8], v[8], a[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) acc[e] = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int yy = 2 * y + (q >> 1), xx = 2 * x + (q & 1);
                unpack8(*(const uint4*)(dY + (((int64_t)n * H + yy) * W + xx) * C + 8 * g), v);
#pragma unroll
                for (int e = 0; e < 8; ++e) acc[e] += v[e];
            }
            const int64_t k = ((int64_t)row * Wo + x) * Ca + 8 * g;
            unpack8(*(const uint4*)(A + k), a);
#pragma unroll
            for (int e = 0; e < 8; ++e)
                if (!(a[e] > 0.0f)) acc[e] = 0.0f;
            *(uint4*)(dA + k) = pack8(acc);
        }
    } else {
        const int row = blockIdx.y - N * Ho, G = Cb / 8;
        for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < W * G; t += gridDim.x * blockDim.x) {
            const int x = t / G, g = t - x * G;
            const int64_t pix = (int64_t)row * W + x;
            *(uint4*)(dB + pix * Cb + 8 * g) = *(const uint4*)(dY + pix * C + Ca + 8 * g);
        }
    }
}

// head_fwd_kernel with the 32 features of a pixel read as four 16-byte pieces.
__global__ void __launch_bounds__(256)
head_fwd_vec_kernel(const bf16_t* __restrict__ F, const float* __restrict__ hw,
                    const float* __restrict__ hb, const float* __restrict__ img,
                    float* __restrict__ out, int64_t P)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P;
         p += (int64_t)gridDim.x * blockDim.x) {
        float s = hb[0];
#pragma unroll
        for (int c8 = 0; c8 < 4; ++c8) {
            float f[8];
            unpack8(*(const uint4*)(F + p * 32 + 8 * c8), f);
#pragma unroll
            for (int e = 0; e < 8; ++e) s = fmaf(hw[8 * c8 + e], f[e], s);
        }
        out[p] = img[p] + s;
    }
}

// ---- head, loss, optimizer ----------------------------------------------------------

// out[p] = img[p] + hb + sum_c hw[c] * F[p][c]   (float result)
__global__ void head_fwd_kernel(const bf16_t* __restrict__ F, const float* __restrict__ hw,
                                const float* __restrict__ hb, const float* __restrict__ img,
                                float* __restrict__ out, int64_t P, int C)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P;
         p += (int64_t)gridDim.x * blockDim.x) {
        float s = hb[0];
        for (int c = 0; c < C; ++c) s = fmaf(hw[c], bf2f(F[p * C + c]), s);
        out[p] = img[p] + s;
    }
}

// dF[p][c] = dout[p] * hw[c] * (F > 0); dhw[c] += sum_p dout[p] F[p][c]; dhb += sum dout.
// Block of 256 threads; per-block sums in LDS, one atomic per block and channel.
__global__ void __launch_bounds__(256)
head_bwd_kernel(const bf16_t* __restrict__ F, const float* __restrict__ hw,
                const float* __restrict__ dout, bf16_t* __restrict__ dF, float* __restrict__ dhw,
                float* __restrict__ dhb, int64_t P, int C)
{
    __shared__ float red[33][256];
    const int t = threadIdx.x;
    float sw[32];
    for (int c = 0; c < 32; ++c) sw[c] = 0.0f;
    float sb = 0.0f;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + t; p < P;
         p += (int64_t)gridDim.x * blockDim.x) {
        const float g = dout[p];
        sb += g;
        for (int c = 0; c < C; ++c) {
            const float f = bf2f(F[p * C + c]);
            sw[c] = fmaf(g, f, sw[c]);
            dF[p * C + c] = f2bf(f > 0.0f ? g * hw[c] : 0.0f);
        }
    }
    for (int c = 0; c < 32; ++c) red[c][t] = sw[c];
    red[32][t] = sb;
    __syncthreads();
    if (t <= 32 && (t < C || t == 32)) {
        float s = 0.0f;
        for (int k = 0; k < 256; ++k) s += red[t][k];
        if (t == 32) atomicAdd(dhb, s);
        else atomicAdd(&dhw[t], s);
    }
}

// head_bwd_kernel for C == 32 with 16-byte aligned F and dF: four lanes share a pixel, one
// 8-channel group each, so a wave moves 1 KB of F and of dF per access.  Every dF element
// is computed as above; the channel sums are added over the lanes of equal group with
// shuffles, then over the block's four waves in LDS, one atomic per block and output.
__global__ void __launch_bounds__(256)
head_bwd_vec_kernel(const bf16_t* __restrict__ F, const float* __restrict__ hw,
                    const float* __restrict__ dout, bf16_t* __restrict__ dF,
                    float* __restrict__ dhw, float* __restrict__ dhb, int64_t P)
{
    __shared__ float red[4][33];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, c8 = t & 3;
    float w[8], sw[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        w[e] = hw[8 * c8 + e];
        sw[e] = 0.0f;
    }
    float sb = 0.0f;
    const int64_t n = 4 * P;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + t; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float g = dout[i >> 2];
        if (c8 == 0) sb += g;
        float f[8], o[8];
        unpack8(*(const uint4*)(F + 8 * i), f);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            sw[e] = fmaf(g, f[e], sw[e]);
            o[e] = f[e] > 0.0f ? g * w[e] : 0.0f;
        }
        *(uint4*)(dF + 8 * i) = pack8(o);
    }
#pragma unroll
    for (int off = 4; off < 64; off <<= 1) {
#pragma unroll
        for (int e = 0; e < 8; ++e) sw[e] += __shfl_xor(sw[e], off);
        sb += __shfl_xor(sb, off);
    }
    if (lane < 4) {
#pragma unroll
        for (int e = 0; e < 8; ++e) red[wave][8 * lane + e] = sw[e];
        if (lane == 0) red[wave][32] = sb;
    }
    __syncthreads();
    if (t <= 32) {
        const float s = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
        atomicAdd(t == 32 ? dhb : &dhw[t], s);
    }
}

// d = out - target; grad = 2 d / P; partial sums of d^2 per block (FP64, fixed order).
__global__ void __launch_bounds__(256)
mse_kernel(const float* __restrict__ out, const float* __restrict__ target,
           float* __restrict__ grad, double* __restrict__ partials, int64_t P)
{
    __shared__ double red[256];
    double s = 0.0;
    const float scale = 2.0f / (float)P;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P;
         p += (int64_t)gridDim.x * blockDim.x) {
        const float d = out[p] - target[p];
        grad[p] = scale * d;
        s += (double)d * (double)d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// out = scale * sum of n partials, one block of 256 threads: thread t adds entries t,
// t + 256, ... in order, then a fixed tree in LDS (the same order on every run).
__global__ void __launch_bounds__(256)
sum_partials_kernel(const double* __restrict__ partials, int n, double scale,
                    double* __restrict__ out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) s += partials[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * scale;
}

// Adam with bias correction on FP32 master weights; refreshes the BF16 copy (may be null).
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v,
                            bf16_t* __restrict__ wbf, int64_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const float gk = g[k];
        const float mk = b1 * m[k] + (1.0f - b1) * gk;
        const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
        m[k] = mk;
        v[k] = vk;
        const float wk = w[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
        w[k] = wk;
        if (wbf) wbf[k] = f2bf(wk);
    }
}

__global__ void to_bf16_kernel(const float* __restrict__ x, bf16_t* __restrict__ y, int64_t n)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        y[k] = f2bf(x[k]);
}

unsigned grid_for(int64_t n, int block = 256)
{
    const int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 65536 ? (g > 0 ? g : 1) : 65536);
}

}  // namespace
}  // namespace itomo

using namespace itomo;

extern "C" {

int itomo_unet_conv3x3(const void* x, int n, int h, int w, int cin, const void* wt,
                       const float* bias, int cout, const void* mask, int relu, void* y,
                       void* stream)
{
    if (!x || !wt || !y || n < 1 || h < 1 || w < 1 || cin < 1 || cout < 1)
        return set_error(ITOMO_EINVAL, "unet_conv3x3: bad argument");
    const int64_t P = (int64_t)n * h * w;
    if (P > ((int64_t)1 << 31) / 8 || (int64_t)9 * cin > (1 << 20))
        return set_error(ITOMO_EINVAL, "unet_conv3x3: too large");
    const bool vec = cin % 8 == 0 && ((uintptr_t)x % 16) == 0 && ((uintptr_t)wt % 16) == 0;
    hipStream_t st = (hipStream_t)stream;
    if (cin == 1 && cout % 8 == 0 && cout <= 64 && ((uintptr_t)y % 16) == 0 &&
        ((uintptr_t)mask % 16) == 0) {
        conv3x3_c1_kernel<<<(unsigned)((P + 255) / 256), 256, 0, st>>>(
            (const bf16_t*)x, (const bf16_t*)wt, bias, (const bf16_t*)mask, (bf16_t*)y, n, h, w,
            cout, relu);
        return check_hip(hipGetLastError(), "unet_conv3x3 launch");
    }
    if (vec && cin % kConvChunk == 0 && cout >= 32) {
        // 16 x 16-pixel tiles; 64 output channels per block, or 32 where that is all
        const bool wide = cout > 32;
        const int nb = wide ? 64 : 32, th = 16;
        const int co_blocks = (cout + nb - 1) / nb;
        if ((int64_t)n * co_blocks > 65535) return set_error(ITOMO_EINVAL, "unet_conv3x3: too large");
        dim3 g((unsigned)((w + 15) / 16), (unsigned)((h + th - 1) / th), (unsigned)(n * co_blocks));
        // 16-byte output stores need whole 8-channel pieces, aligned
        const int vec_out = cout % 8 == 0 && ((uintptr_t)y % 16) == 0 &&
                            ((uintptr_t)mask % 16) == 0;
        if (wide)
            conv3x3_tiled_kernel<4, 4><<<g, 256, 0, st>>>((const bf16_t*)x, (const bf16_t*)wt, bias,
                                                          (const bf16_t*)mask, (bf16_t*)y, h, w,
                                                          cin, cout, relu, co_blocks, vec_out);
        else
            conv3x3_tiled_kernel<4, 2><<<g, 256, 0, st>>>((const bf16_t*)x, (const bf16_t*)wt, bias,
                                                          (const bf16_t*)mask, (bf16_t*)y, h, w,
                                                          cin, cout, relu, co_blocks, vec_out);
        return check_hip(hipGetLastError(), "unet_conv3x3 launch");
    }
    dim3 grid((unsigned)((P + 127) / 128), (unsigned)((cout + 31) / 32));
    if (vec)
        conv3x3_kernel<true><<<grid, 256, 0, st>>>((const bf16_t*)x, (const bf16_t*)wt, bias,
                                                   (const bf16_t*)mask, (bf16_t*)y, n, h, w,
                                                   cin, cout, relu);
    else
        conv3x3_kernel<false><<<grid, 256, 0, st>>>((const bf16_t*)x, (const bf16_t*)wt, bias,
                                                    (const bf16_t*)mask, (bf16_t*)y, n, h, w,
                                                    cin, cout, relu);
    return check_hip(hipGetLastError(), "unet_conv3x3 launch");
}

int itomo_unet_flip(const void* wt, int cin, int cout, void* wf, void* stream)
{
    if (!wt || !wf || cin < 1 || cout < 1) return set_error(ITOMO_EINVAL, "unet_flip: bad argument");
    const int64_t n = (int64_t)cout * 9 * cin;
    flip_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>((const bf16_t*)wt, (bf16_t*)wf,
                                                               cin, cout);
    return check_hip(hipGetLastError(), "unet_flip launch");
}

int itomo_unet_wgrad(const void* x, const void* dy, int n, int h, int w, int cin, int cout,
                     float* dw, float* db, void* stream)
{
    if (!x || !dy || !dw || n < 1 || h < 1 || w < 1 || cin < 1 || cout < 1 || cout > 1024)
        return set_error(ITOMO_EINVAL, "unet_wgrad: bad argument");
    const int64_t P = (int64_t)n * h * w;
    hipStream_t st = (hipStream_t)stream;
    if (cin == 1 && cout == 32 && ((uintptr_t)dy % 16) == 0) {
        // about 2048 blocks, at least 4 row segments each
        const int64_t nseg = (int64_t)n * h * ((w + 63) / 64);
        const int64_t g = std::max<int64_t>(1, std::min<int64_t>(2048, (nseg + 3) / 4));
        wgrad_c1_kernel<<<(unsigned)g, 256, 0, st>>>((const bf16_t*)x, (const bf16_t*)dy, dw, db,
                                                     n, h, w);
        return check_hip(hipGetLastError(), "unet_wgrad launch");
    }
    if (cin % 32 == 0 && cout % 32 == 0 && ((uintptr_t)x % 16) == 0 && ((uintptr_t)dy % 16) == 0) {
        // 64 output channels per block where they divide, else 32; about 2048 blocks
        const bool wide = cout % 64 == 0;
        const int mb = wide ? 64 : 32;
        const int ci_chunks = cin / 32;
        const int gxw = (cout / mb) * ci_chunks;
        const int64_t ntiles = (int64_t)n * ((h + 7) / 8) * ((w + 15) / 16);
        // at least 4 pixel tiles per block, so that the atomic adds stay few
        int64_t gz = std::min<int64_t>(2048 / gxw, (ntiles + 3) / 4);
        gz = std::max<int64_t>(1, std::min<int64_t>(gz, 65535));
        if (wide)
            wgrad_tiled_kernel<2><<<dim3((unsigned)gxw, 1, (unsigned)gz), 256, 0, st>>>(
                (const bf16_t*)x, (const bf16_t*)dy, dw, db, n, h, w, cin, cout, ci_chunks);
        else
            wgrad_tiled_kernel<1><<<dim3((unsigned)gxw, 1, (unsigned)gz), 256, 0, st>>>(
                (const bf16_t*)x, (const bf16_t*)dy, dw, db, n, h, w, cin, cout, ci_chunks);
        return check_hip(hipGetLastError(), "unet_wgrad launch");
    } else {
    const int ncol = 9 * cin;
    const unsigned gx = (unsigned)((ncol + 63) / 64), gy = (unsigned)((cout + 63) / 64);
    // split the pixels so that the grid holds ~2048 blocks, in multiples of 32 pixels
    int64_t nz = 2048 / ((int64_t)gx * gy);
    nz = nz < 1 ? 1 : (nz > 65535 ? 65535 : nz);
    int64_t chunk = (P + nz - 1) / nz;
    chunk = ((chunk + 31) / 32) * 32;
    nz = (P + chunk - 1) / chunk;
    wgrad_kernel<<<dim3(gx, gy, (unsigned)nz), 256, 0, st>>>((const bf16_t*)x, (const bf16_t*)dy,
                                                             dw, n, h, w, cin, cout, chunk);
    }
    int rc = check_hip(hipGetLastError(), "unet_wgrad launch");
    if (rc || !db) return rc;
    const int64_t bchunk = 1024;
    bias_grad_kernel<<<(unsigned)((P + bchunk - 1) / bchunk), cout, 0, st>>>(
        (const bf16_t*)dy, db, P, cout, bchunk);
    return check_hip(hipGetLastError(), "unet_bias_grad launch");
}

int itomo_unet_pool_fwd(const void* x, int n, int h, int w, int c, void* y, void* idx,
                        void* stream)
{
    if (!x || !y || !idx || n < 1 || h < 2 || w < 2 || c < 1 || (h & 1) || (w & 1))
        return set_error(ITOMO_EINVAL, "unet_pool_fwd: bad argument");
    if (c % 8 == 0 && (int64_t)n * (h / 2) <= 65535 && ((uintptr_t)x % 16) == 0 &&
        ((uintptr_t)y % 16) == 0 && ((uintptr_t)idx % 8) == 0) {
        const int t = (w / 2) * (c / 8);
        pool_fwd_vec_kernel<<<dim3((unsigned)((t + 255) / 256), (unsigned)(n * (h / 2))), 256, 0,
                              (hipStream_t)stream>>>((const bf16_t*)x, (bf16_t*)y,
                                                     (unsigned char*)idx, h, w, c);
        return check_hip(hipGetLastError(), "unet_pool_fwd launch");
    }
    const int64_t no = (int64_t)n * (h / 2) * (w / 2) * c;
    pool_fwd_kernel<<<grid_for(no), 256, 0, (hipStream_t)stream>>>(
        (const bf16_t*)x, (bf16_t*)y, (unsigned char*)idx, n, h, w, c);
    return check_hip(hipGetLastError(), "unet_pool_fwd launch");
}

int itomo_unet_pool_bwd(const void* dy, const void* idx, const void* skip, const void* xm,
                        int n, int h, int w, int c, void* dx, void* stream)
{
    if (!dy || !idx || !xm || !dx || n < 1 || h < 2 || w < 2 || c < 1 || (h & 1) || (w & 1))
        return set_error(ITOMO_EINVAL, "unet_pool_bwd: bad argument");
    if (c % 8 == 0 && (int64_t)n * h <= 65535 && ((uintptr_t)dy % 16) == 0 &&
        ((uintptr_t)idx % 8) == 0 && ((uintptr_t)skip % 16) == 0 && ((uintptr_t)xm % 16) == 0 &&
        ((uintptr_t)dx % 16) == 0) {
        const int t = w * (c / 8);
        pool_bwd_vec_kernel<<<dim3((unsigned)((t + 255) / 256), (unsigned)(n * h)), 256, 0,
                              (hipStream_t)stream>>>(
            (const bf16_t*)dy, (const unsigned char*)idx, (const bf16_t*)skip,
            (const bf16_t*)xm, (bf16_t*)dx, h, w, c);
        return check_hip(hipGetLastError(), "unet_pool_bwd launch");
    }
    const int64_t ni = (int64_t)n * h * w * c;
    pool_bwd_kernel<<<grid_for(ni), 256, 0, (hipStream_t)stream>>>(
        (const bf16_t*)dy, (const unsigned char*)idx, (const bf16_t*)skip, (const bf16_t*)xm,
        (bf16_t*)dx, n, h, w, c);
    return check_hip(hipGetLastError(), "unet_pool_bwd launch");
}

int itomo_unet_upcat_fwd(const void* a, const void* b, int n, int h, int w, int ca, int cb,
                         void* y, void* stream)
{
    if (!a || !b || !y || n < 1 || h < 2 || w < 2 || ca < 1 || cb < 1 || (h & 1) || (w & 1))
        return set_error(ITOMO_EINVAL, "unet_upcat_fwd: bad argument");
    if (ca % 8 == 0 && cb % 8 == 0 && (int64_t)n * h <= 65535 && ((uintptr_t)a % 16) == 0 &&
        ((uintptr_t)b % 16) == 0 && ((uintptr_t)y % 16) == 0) {
        const int t = w * ((ca + cb) / 8);
        upcat_fwd_vec_kernel<<<dim3((unsigned)((t + 255) / 256), (unsigned)(n * h)), 256, 0,
                               (hipStream_t)stream>>>((const bf16_t*)a, (const bf16_t*)b,
                                                      (bf16_t*)y, h, w, ca, cb);
        return check_hip(hipGetLastError(), "unet_upcat_fwd launch");
    }
    const int64_t no = (int64_t)n * h * w * (ca + cb);
    upcat_fwd_kernel<<<grid_for(no), 256, 0, (hipStream_t)stream>>>(
        (const bf16_t*)a, (const bf16_t*)b, (bf16_t*)y, n, h, w, ca, cb);
    return check_hip(hipGetLastError(), "unet_upcat_fwd launch");
}

int itomo_unet_upcat_bwd(const void* dy, const void* a, int n, int h, int w, int ca, int cb,
                         void* da, void* db, void* stream)
{
    if (!dy || !a || !da || !db || n < 1 || h < 2 || w < 2 || ca < 1 || cb < 1 || (h & 1) ||
        (w & 1))
        return set_error(ITOMO_EINVAL, "unet_upcat_bwd: bad argument");
    if (ca % 8 == 0 && cb % 8 == 0 && (int64_t)n * (h / 2) + (int64_t)n * h <= 65535 &&
        ((uintptr_t)dy % 16) == 0 && ((uintptr_t)a % 16) == 0 && ((uintptr_t)da % 16) == 0 &&
        ((uintptr_t)db % 16) == 0) {
        const int t = std::max((w / 2) * (ca / 8), w * (cb / 8));
        upcat_bwd_vec_kernel<<<dim3((unsigned)((t + 255) / 256), (unsigned)(n * (h / 2) + n * h)),
                               256, 0, (hipStream_t)stream>>>(
            (const bf16_t*)dy, (const bf16_t*)a, (bf16_t*)da, (bf16_t*)db, n, h, w, ca, cb);
        return check_hip(hipGetLastError(), "unet_upcat_bwd launch");
    }
    const int64_t nt = (int64_t)n * (h / 2) * (w / 2) * ca + (int64_t)n * h * w * cb;
    upcat_bwd_kernel<<<grid_for(nt), 256, 0, (hipStream_t)stream>>>(
        (const bf16_t*)dy, (const bf16_t*)a, (bf16_t*)da, (bf16_t*)db, n, h, w, ca, cb);
    return check_hip(hipGetLastError(), "unet_upcat_bwd launch");
}

int itomo_unet_head_fwd(const void* f, const float* hw, const float* hb, const float* img,
                        int64_t npixels, int c, float* out, void* stream)
{
    if (!f || !hw || !hb || !img || !out || npixels < 1 || c < 1 || c > 32)
        return set_error(ITOMO_EINVAL, "unet_head_fwd: bad argument");
    if (c == 32 && ((uintptr_t)f % 16) == 0) {
        head_fwd_vec_kernel<<<grid_for(npixels), 256, 0, (hipStream_t)stream>>>(
            (const bf16_t*)f, hw, hb, img, out, npixels);
        return check_hip(hipGetLastError(), "unet_head_fwd launch");
    }
    head_fwd_kernel<<<grid_for(npixels), 256, 0, (hipStream_t)stream>>>(
        (const bf16_t*)f, hw, hb, img, out, npixels, c);
    return check_hip(hipGetLastError(), "unet_head_fwd launch");
}

int itomo_unet_head_bwd(const void* f, const float* hw, const float* dout, int64_t npixels,
                        int c, void* df, float* dhw, float* dhb, void* stream)
{
    if (!f || !hw || !dout || !df || !dhw || !dhb || npixels < 1 || c < 1 || c > 32)
        return set_error(ITOMO_EINVAL, "unet_head_bwd: bad argument");
    if (c == 32 && ((uintptr_t)f % 16) == 0 && ((uintptr_t)df % 16) == 0) {
        const unsigned g = grid_for(4 * npixels) < 1024 ? grid_for(4 * npixels) : 1024;
        head_bwd_vec_kernel<<<g, 256, 0, (hipStream_t)stream>>>(
            (const bf16_t*)f, hw, dout, (bf16_t*)df, dhw, dhb, npixels);
        return check_hip(hipGetLastError(), "unet_head_bwd launch");
    }
    const unsigned g = grid_for(npixels) < 1024 ? grid_for(npixels) : 1024;
    head_bwd_kernel<<<g, 256, 0, (hipStream_t)stream>>>((const bf16_t*)f, hw, dout,
                                                         (bf16_t*)df, dhw, dhb, npixels, c);
    return check_hip(hipGetLastError(), "unet_head_bwd launch");
}

int itomo_unet_mse(const float* out, const float* target, int64_t npixels, float* grad,
                   double* loss, double* work, void* stream)
{
    if (!out || !target || !grad || !loss || !work || npixels < 1)
        return set_error(ITOMO_EINVAL, "unet_mse: bad argument");
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = grid_for(npixels) < ITOMO_UNET_MSE_WORK ? grid_for(npixels)
                                                               : ITOMO_UNET_MSE_WORK;
    mse_kernel<<<g, 256, 0, st>>>(out, target, grad, work, npixels);
    int rc = check_hip(hipGetLastError(), "unet_mse launch");
    if (rc) return rc;
    sum_partials_kernel<<<1, 256, 0, st>>>(work, (int)g, 1.0 / (double)npixels, loss);
    return check_hip(hipGetLastError(), "unet_mse sum launch");
}

int itomo_unet_adam(float* w, const float* g, float* m, float* v, void* wbf, int64_t n,
                    float lr, float beta1, float beta2, float eps, int step, void* stream)
{
    if (!w || !g || !m || !v || n < 1 || step < 1)
        return set_error(ITOMO_EINVAL, "unet_adam: bad argument");
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    adam_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(w, g, m, v, (bf16_t*)wbf, n, lr,
                                                               beta1, beta2, eps, c1, c2);
    return check_hip(hipGetLastError(), "unet_adam launch");
}

int itomo_unet_to_bf16(const float* x, int64_t n, void* y, void* stream)
{
    if (!x || !y || n < 1) return set_error(ITOMO_EINVAL, "unet_to_bf16: bad argument");
    to_bf16_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(x, (bf16_t*)y, n);
    return check_hip(hipGetLastError(), "unet_to_bf16 launch");
}

}  // extern "C"
