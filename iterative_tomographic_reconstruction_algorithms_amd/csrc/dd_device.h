// Device code of the distance-driven projector pair, shared by the per-plan kernels
// (dd.hip) and the persistent batch kernels (batch.hip).
#pragma once

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

// Start L* of the segment a sweep line drops (radon_dd.c:176-177).  The reference
// merges the pixel keys P_0..P_N with the bin keys D_0..D_N (on equal keys the pixel
// key first) and stops at the first key with index N, before that segment is used;
// L* is the key merged just before it:
//   P_N <= D_N (stop at P_N):  L* = max(P_{N-1}, largest D_j <  P_N)
//   D_N <  P_N (stop at D_N):  L* = max(D_{N-1}, largest P_k <= D_N)
// Every other segment ends at or before L*, so clipping each segment's right end to L*
// removes exactly the dropped one -- also when P_k == D_N for some k < N, where the
// dropped segment has zero length and the one ending at P_k is kept.
template <int ARITH>
__device__ __forceinline__ float dd_line_end(const DdView& g, int npix, int nh, int i)
{
    const float K = key_K<ARITH>(g, line_coord(i, nh));
    const float PN = __fsub_rn((float)(npix - nh), K);
    const float DN = key_D<ARITH>(g, npix - nh);
    const float ninf = __int_as_float(0xff800000);
    // binary search for the last index in [0, npix) whose key precedes the stop key;
    // both key sequences are non-decreasing in their index
    int lo = -1, hi = npix;
    if (PN <= DN) {
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (key_D<ARITH>(g, mid - nh) < PN) lo = mid; else hi = mid;
        }
        const float Dj = lo >= 0 ? key_D<ARITH>(g, lo - nh) : ninf;
        return fmaxf(__fsub_rn((float)(npix - 1 - nh), K), Dj);
    }
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (__fsub_rn((float)(mid - nh), K) <= DN) lo = mid; else hi = mid;
    }
    const float Pk = lo >= 0 ? __fsub_rn((float)(lo - nh), K) : ninf;
    return fmaxf(key_D<ARITH>(g, npix - 1 - nh), Pk);
}

}  // namespace

// ---- back projection ----------------------------------------------------------------
//
// Block = 64x64 pixel tile, 256 threads, each thread a 4x4 micro-tile (rows 4*ty.., columns
// 4*tx..).  For a row sweep a thread walks its 4 rows, 4 consecutive pixels each; for a
// column sweep its 4 columns.  Along one line the bin walk is incremental: the first
// pixel's bin is found once, then each next pixel continues from the previous pixel's
// last bin.  Views are processed in chunks of kBackViews: the sinogram values the tile
// can touch in each view (at most kBackFoot bins) are staged in LDS, with zeros for
// bins outside [0, npix), so the walk needs no bounds checks.  The accumulation order
// per pixel (views in list order, bins increasing) is the reference's.
constexpr int kBackTile = kDdBackTile;
constexpr int kBackViews = 32;
constexpr int kBackFoot = 100;  // >= 64*(|sin|+|cos|) + 5 for every view

template <int ARITH>
__device__ __forceinline__ void staged_range(const DdView& g, int npix, int nh, int r0,
                                             int q0, int* b_lo, int* cnt)
{
    int lo, hi;
    if (g.kind == DD_ROWSUM) {
        lo = r0;
        hi = r0 + kBackTile - 1;
    } else if (g.kind == DD_COLSUM) {
        lo = q0;
        hi = q0 + kBackTile - 1;
    } else if (g.kind == DD_SWEEP_ROWS || g.kind == DD_SWEEP_COLS) {
        // lines are rows (r) or columns (q); pixel boundaries along the line
        const int l0 = (g.kind == DD_SWEEP_ROWS) ? r0 : q0;
        const int p0 = (g.kind == DD_SWEEP_ROWS) ? q0 : r0;
        const float ka = (float)(l0 - nh) + 0.5f, kb = ka + (float)(kBackTile - 1);
        const float K1 = ka * g.t, K2 = kb * g.t;
        const float kmin = (float)(p0 - nh) - fmaxf(K1, K2);
        const float kmax = (float)(p0 + kBackTile - nh) - fminf(K1, K2);
        float f1 = floorf(kmin * g.est), f2 = floorf(kmax * g.est);
        f1 = fminf(fmaxf(f1, (float)(-nh - 4)), (float)(npix - nh + 4));
        f2 = fminf(fmaxf(f2, (float)(-nh - 4)), (float)(npix - nh + 4));
        lo = (int)f1 + nh - 2;
        hi = (int)f2 + nh + 2;
    } else {
        lo = 0;
        hi = -1;
    }
    lo = max(lo, -1);
    hi = min(hi, npix);
    *b_lo = lo;
    *cnt = max(0, min(hi - lo + 1, kBackFoot));
}

// One line of 4 pixels with along-line indices p0..p0+3.  S points at LDS such that
// S[d] is bin d of this view (zero outside [0, npix)); Ls is the line's L*.
template <int ARITH>
__device__ __forceinline__ void walk4(const DdView& g, int npix, int nh, float K, float Ls,
                                      int p0, const float* S, float& a0, float& a1,
                                      float& a2, float& a3)
{
    float x = (float)(p0 - nh);
    float lo = __fsub_rn(x, K);
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
    // invariant from here on: Dlo <= lo <= Dhi
#define ITOMO_PIX(acc)                                                               \
    {                                                                                \
        x = __fadd_rn(x, 1.0f);                                                      \
        const float hi = __fsub_rn(x, K);                                            \
        const float hc = fminf(hi, Ls);                                              \
        {                                                                            \
            const float R = fminf(hc, Dhi), L = fmaxf(lo, Dlo);                      \
            if (R > L) acc = accum<ARITH>(__fsub_rn(R, L), S[d], acc);               \
        }                                                                            \
        while (d < npix && Dhi < hi) {                                               \
            ++d;                                                                     \
            Dlo = Dhi;                                                               \
            Dhi = key_D<ARITH>(g, d + 1 - nh);                                       \
            const float R = fminf(hc, Dhi);                                          \
            if (R > Dlo) acc = accum<ARITH>(__fsub_rn(R, Dlo), S[d], acc);           \
        }                                                                            \
        lo = hi;                                                                     \
    }
    ITOMO_PIX(a0)
    ITOMO_PIX(a1)
    ITOMO_PIX(a2)
    ITOMO_PIX(a3)
#undef ITOMO_PIX
}

// One 64x64 tile (tile column bx, tile row by) of slice z; blk/nbx index its FP64
// partials.  Shared by the per-plan kernel and the persistent batch kernel.
template <int ARITH, int EPI, bool ONES>
__device__ __forceinline__ void dd_back_tile(const DdView* __restrict__ geom,
                                             const float* __restrict__ lstar, int nang, int npix,
                                             const float* __restrict__ sino,
                                             float* __restrict__ out, int64_t sino_stride,
                                             int64_t img_stride, int accumulate,
                                             const EpiArgs& e, int bx, int by, int z, int blk)
{
    __shared__ float sS[kBackViews][kBackFoot];
    __shared__ int sLo[kBackViews];
    __shared__ int sCnt[kBackViews];
    __shared__ float sLs[kBackViews][kBackTile];   // L* of the tile's lines per view

    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int r0 = by * kBackTile, q0 = bx * kBackTile;
    const int rr = r0 + 4 * ty, cc = q0 + 4 * tx;  // micro-tile origin
    sino += (int64_t)z * sino_stride;
    out += (int64_t)z * img_stride;
    const int nh = (int)(npix * 0.5);

    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = rr + i, q = cc + j;
            acc[i][j] = (EPI == EPI_STORE && accumulate && r < npix && q < npix)
                            ? out[(int64_t)r * npix + q]
                            : 0.0f;
        }

    for (int v0 = 0; v0 < nang; v0 += kBackViews) {
        const int nv = min(kBackViews, nang - v0);
        __syncthreads();  // previous chunk fully consumed
        if (tid < nv) {
            int lo, cnt;
            staged_range<ARITH>(geom[v0 + tid], npix, nh, r0, q0, &lo, &cnt);
            sLo[tid] = lo;
            sCnt[tid] = cnt;
        }
        __syncthreads();
        for (int k = tid; k < nv * kBackFoot; k += 256) {
            const int v = k / kBackFoot, j = k - v * kBackFoot;
            const int d = sLo[v] + j;
            float val = 0.0f;
            if (j < sCnt[v] && d >= 0 && d < npix)
                val = ONES ? 1.0f : sino[(int64_t)geom[v0 + v].row * npix + d];
            sS[v][j] = val;
        }
        for (int k = tid; k < nv * kBackTile; k += 256) {
            const int v = k / kBackTile, l = k - v * kBackTile;
            const int line = (geom[v0 + v].kind == DD_SWEEP_COLS ? q0 : r0) + l;
            sLs[v][l] = line < npix ? lstar[(int64_t)(v0 + v) * npix + line] : 0.0f;
        }
        __syncthreads();
        for (int v = 0; v < nv; ++v) {
            const DdView g = geom[v0 + v];
            const float* S = &sS[v][0] - sLo[v];
            if (g.kind == DD_SWEEP_ROWS) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float K = key_K<ARITH>(g, line_coord(rr + i, nh));
                    walk4<ARITH>(g, npix, nh, K, sLs[v][4 * ty + i], cc, S, acc[i][0],
                                 acc[i][1], acc[i][2], acc[i][3]);
                }
            } else if (g.kind == DD_SWEEP_COLS) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float K = key_K<ARITH>(g, line_coord(cc + j, nh));
                    walk4<ARITH>(g, npix, nh, K, sLs[v][4 * tx + j], rr, S, acc[0][j],
                                 acc[1][j], acc[2][j], acc[3][j]);
                }
            } else if (g.kind == DD_ROWSUM) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float sv = (rr + i < npix) ? S[rr + i] : 0.0f;
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = __fadd_rn(acc[i][j], sv);
                }
            } else if (g.kind == DD_COLSUM) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float sv = (cc + j < npix) ? S[cc + j] : 0.0f;
#pragma unroll
                    for (int i = 0; i < 4; ++i) acc[i][j] = __fadd_rn(acc[i][j], sv);
                }
            }
        }
    }

    double p1 = 0.0, p2 = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = rr + i, q = cc + j;
            if (r < npix && q < npix) {
                const int64_t o = (int64_t)r * npix + q;
                double q1 = 0.0, q2 = 0.0;
                if (epi_updates_x(EPI))
                    out[o] = epi_back<EPI>(acc[i][j], out[o], o, z, e, &q1, &q2);
                else
                    out[o] = epi_back<EPI>(acc[i][j], 0.0f, o, z, e, &q1, &q2);
                p1 += q1;
                p2 += q2;
            }
        }
    if (epi_has_partials(EPI) && e.partials)
        block_partials3(0.0, p1, p2,
                        e.partials + ((int64_t)z * e.blocks_per_slice + blk) * 3);
}


__device__ __forceinline__ __amdgpu_buffer_rsrc_t image_rsrc(const float* base, int npix)
{
    // byte size; npix <= 32766 keeps it below 2^32
    const unsigned bytes = (unsigned)npix * (unsigned)npix * 4u;
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, 0x00020000);
}

// image[idx] of one slice; indices outside [0, npix*npix) read 0.
__device__ __forceinline__ float load_px(__amdgpu_buffer_rsrc_t rs, int idx)
{
    // the builtin returns the 32 bits as an integer: reinterpret, do not convert
    return __int_as_float(
        (int)__builtin_amdgcn_raw_buffer_load_b32(rs, (int)((unsigned)idx * 4u), 0, 0));
}

// One candidate pixel of a bin [Dlo, Dhi] on a line with offset K: the pixel spans
// [Plo, x+1-K].  Only the first candidate can start left of Dlo.  Dhc is the bin's right
// end clipped to the line's L*.  Advances x and Plo.
template <int ARITH>
__device__ __forceinline__ float cand_step(bool first, bool in_range, float& x, float& Plo,
                                           float K, float Dlo, float Dhc, float val, float acc)
{
    x = __fadd_rn(x, 1.0f);
    const float Phi = __fsub_rn(x, K);
    const float R = (Phi < Dhc) ? Phi : Dhc;
    const float L = first ? ((Plo > Dlo) ? Plo : Dlo) : Plo;
    const bool ok = in_range && R > L;
    const float nacc = accum<ARITH>(__fsub_rn(R, L), val, acc);
    Plo = Phi;
    return ok ? nacc : acc;
}

// Forward projection: one thread per (bin, view blockIdx.y, slice blockIdx.z).
// Lines are image rows for row sweeps and image columns for column sweeps; column
// sweeps (and row sums) read the transposed image so that the lanes of a wave, which
// hold consecutive bins, read consecutive addresses.
// Bins [bx*kDdFwdBlock, +kDdFwdBlock) of view `view`, slice z; blk indexes the partials.
template <int ARITH, int EPI, bool ONES>
__device__ __forceinline__ void dd_fwd_block(const DdView* __restrict__ geom,
                                             const float* __restrict__ lstar, int nang,
                                             int npix, const float* __restrict__ image,
                                             const float* __restrict__ imageT,
                                             float* __restrict__ out, int64_t img_stride,
                                             int64_t sino_stride, int accumulate,
                                             const EpiArgs& e, int bx, int view, int z, int blk)
{
    const int d = bx * kDdFwdBlock + threadIdx.x;
    const bool active = d < npix;
    const DdView g = geom[view];
    image += (int64_t)z * img_stride;
    imageT += (int64_t)z * img_stride;
    const int64_t soff = (int64_t)z * sino_stride;
    const int64_t o = (int64_t)g.row * npix + d;
    const int nh = (int)(npix * 0.5);
    double p0 = 0.0;

    if (active) {
        float acc = (EPI == EPI_STORE && accumulate) ? out[soff + o] : 0.0f;
        if (g.kind == DD_ROWSUM) {
            // sum_j img[d][j], in j order, read as column d of the transpose
            const float* p = imageT + d;
            for (int j = 0; j < npix; ++j)
                acc = __fadd_rn(acc, ONES ? 1.0f : p[(int64_t)j * npix]);
        } else if (g.kind == DD_COLSUM) {
            const float* p = image + d;
            for (int j = 0; j < npix; ++j)
                acc = __fadd_rn(acc, ONES ? 1.0f : p[(int64_t)j * npix]);
        } else if (g.kind == DD_SWEEP_ROWS || g.kind == DD_SWEEP_COLS) {
            const __amdgpu_buffer_rsrc_t rs =
                image_rsrc((g.kind == DD_SWEEP_ROWS) ? image : imageT, npix);
            const float Dlo = key_D<ARITH>(g, d - nh);
            const float Dhi = key_D<ARITH>(g, d + 1 - nh);
            const float* ls = lstar + (int64_t)view * npix;
            const float x0 = (float)(-nh), xN = (float)(npix - nh);
            // lines whose pixel span [x0 - K, xN - K] can meet [Dlo, Dhi]:
            // x0 - Dhi < K_i < xN - Dlo with K_i ~ (i - nh + 0.5) * t; two lines of margin,
            // the exact test below decides.
            double ua = ((double)x0 - (double)Dhi) / (double)g.t;
            double ub = ((double)xN - (double)Dlo) / (double)g.t;
            if (ua > ub) { const double tmp = ua; ua = ub; ub = tmp; }
            ua = fmin(fmax(ua + nh - 0.5, -4.0), (double)npix + 4.0);
            ub = fmin(fmax(ub + nh - 0.5, -4.0), (double)npix + 4.0);
            const int i_lo = max(0, (int)floor(ua) - 2);
            const int i_hi = min(npix - 1, (int)ceil(ub) + 2);
            const int nc = g.ncand;
            for (int i = i_lo; i <= i_hi; ++i) {
                const float K = key_K<ARITH>(g, line_coord(i, nh));
                const float P0 = __fsub_rn(x0, K), PN = __fsub_rn(xN, K);
                if (!(PN > Dlo && P0 < Dhi)) continue;
                const float Dhc = fminf(Dhi, ls[i]);
                // the pixel whose span holds Dlo: estimate, then one exact correction
                float fe = floorf(__fadd_rn(Dlo, K));
                fe = fminf(fmaxf(fe, (float)(-nh - 1)), (float)(npix - nh));
                int q = (int)fe + nh;  // in [-1, npix]
                {
                    const float Pe = __fsub_rn((float)(q - nh), K);
                    const float Pe1 = __fsub_rn((float)(q + 1 - nh), K);
                    q += (Dlo >= Pe1) ? 1 : ((Dlo < Pe) ? -1 : 0);
                    // a bin starting before pixel 0 first meets pixel 0
                    q = min(max(q, 0), npix - 1);
                }
                // candidates q .. q+nc-1; loads go through a buffer descriptor whose
                // range check returns 0 outside the image, so no index clamping
                const int lbase = i * npix;
                float x = (float)(q - nh);
                float Plo = __fsub_rn(x, K);
                if (nc <= 3) {
                    float v[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) v[k] = ONES ? 1.0f : load_px(rs, lbase + q + k);
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        acc = cand_step<ARITH>(k == 0, q + k < npix, x, Plo, K, Dlo, Dhc, v[k], acc);
                } else {
                    for (int k0 = 0; k0 < nc; k0 += 4) {
                        float v[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            v[k] = ONES ? 1.0f : load_px(rs, lbase + q + k0 + k);
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            acc = cand_step<ARITH>(k0 + k == 0, (k0 + k < nc) && (q + k0 + k < npix),
                                                   x, Plo, K, Dlo, Dhc, v[k], acc);
                    }
                }
            }
        }
        out[soff + o] = epi_forward<EPI>(acc, o, soff, z, e, &p0);
    }
    if (epi_has_partials(EPI) && e.partials)
        block_partials3(p0, 0.0, 0.0,
                        e.partials + ((int64_t)z * e.blocks_per_slice + blk) * 3);
}

// out[c][r] = in[r][c] for one 32x32 tile of one npix x npix slice, through LDS
// (block 32 x 8 threads).
__device__ __forceinline__ void transpose_tile(const float* __restrict__ in,
                                               float* __restrict__ out, int npix,
                                               int64_t stride, int bx, int by, int z)
{
    __shared__ float tile[32][33];
    in += (int64_t)z * stride;
    out += (int64_t)z * stride;
    const int c0 = bx * 32, r0 = by * 32;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + threadIdx.x;
        if (r < npix && c < npix) tile[k][threadIdx.x] = in[(int64_t)r * npix + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = c0 + k, c = r0 + threadIdx.x;  // output row = input column
        if (r < npix && c < npix) out[(int64_t)r * npix + c] = tile[threadIdx.x][k];
    }
}

}  // namespace itomo
