// Device code of the distance-driven projector pair, shared by the per-plan kernels
// (dd.hip) and the persistent batch kernels (batch.hip).
#pragma once

#include "epilogue.h"

namespace itomo {

namespace {

// Bin boundary D for j = d - nh, with j given as an integer-valued float.
template <int ARITH>
__device__ __forceinline__ float key_Df(const DdView& g, float x)
{
    if (ARITH == ITOMO_ARITH_FAST) return __fmul_rn(x, g.rinv);
    return __fdiv_rn(x, g.b);
}

template <int ARITH>
__device__ __forceinline__ float key_D(const DdView& g, int j)
{
    return key_Df<ARITH>(g, (float)j);
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
// Block = T x T pixel tile, (T/P)^2 threads, each thread a P x P micro-tile (rows
// P*ty.., columns P*tx..).  For a row sweep a thread walks its P rows, P consecutive
// pixels each; for a column sweep its P columns.  Along one line the bin walk is
// incremental: the first pixel's bin is found once, then each next pixel continues from
// the previous pixel's last bin.  Views are processed in chunks of kBackViews: the
// sinogram values the tile can touch in each view (at most back_foot(T) bins) are staged
// in LDS, with zeros for bins outside [0, npix), so the walk needs no bounds checks.  The
// accumulation order per pixel (views in list order, bins increasing) is the reference's
// for every (T, P), so all tilings give bit-identical images.
constexpr int kBackViews = 32;

// Staged bins per view: the tile's footprint spans T*(|sin|+|cos|) <= T*sqrt(2) bins,
// plus 7 for the floors and the +-2 margins of staged_range.
template <int T>
constexpr int back_foot() { return (T * 1449) / 1024 + 10; }

template <int ARITH, int T>
__device__ __forceinline__ void staged_range(const DdView& g, int npix, int nh, int r0,
                                             int q0, int* b_lo, int* cnt)
{
    int lo, hi;
    if (g.kind == DD_ROWSUM) {
        lo = r0;
        hi = r0 + T - 1;
    } else if (g.kind == DD_COLSUM) {
        lo = q0;
        hi = q0 + T - 1;
    } else if (g.kind == DD_SWEEP_ROWS || g.kind == DD_SWEEP_COLS) {
        // lines are rows (r) or columns (q); pixel boundaries along the line
        const int l0 = (g.kind == DD_SWEEP_ROWS) ? r0 : q0;
        const int p0 = (g.kind == DD_SWEEP_ROWS) ? q0 : r0;
        const float ka = (float)(l0 - nh) + 0.5f, kb = ka + (float)(T - 1);
        const float K1 = ka * g.t, K2 = kb * g.t;
        const float kmin = (float)(p0 - nh) - fmaxf(K1, K2);
        const float kmax = (float)(p0 + T - nh) - fminf(K1, K2);
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
    *cnt = max(0, min(hi - lo + 1, back_foot<T>()));
}

// One line of P pixels with along-line indices p0..p0+P-1, accumulating into
// a[0], a[STRIDE], ..., a[(P-1)*STRIDE].  S points at LDS such that S[d] is bin d of
// this view (zero outside [0, npix)); Ls is the line's L*.
//
// Bins are at least as wide as pixels (1/|sin| or 1/|cos| >= 1), so a pixel crosses at
// most one bin boundary, except when the two widths agree to rounding (views flagged
// `multi`, the MULTI instance): there a loop picks up the further crossings.  Bins are
// visited in increasing order either way.  Past the last bin the weights vanish by the
// L* clip (Ls <= D[N]), so the walk needs no bound there.
//
// The walk keeps the bin that holds the pixel's left end lo (D_lo <= lo < D_hi), so
// every overlap starts at lo or at a crossed key; the one exception, a left end below
// bin -1, reads a staged zero.  A weight max(R - L, 0) of zero leaves the sum unchanged,
// so no overlap needs a test.
template <int ARITH, int P, int STRIDE, bool MULTI>
__device__ __forceinline__ void walk_line(const DdView& g, int npix, int nh, float K, float Ls,
                                          int p0, const float* S, float* a)
{
    float x = (float)(p0 - nh);
    float lo = __fsub_rn(x, K);
    // (index - nh) of the bin holding lo (D_j <= lo < D_j+1), then the clamp to
    // [-1, npix] (bins -1 and npix read staged zeros).  In index units the estimate
    // lo*est is off from lo/(bin width) by a relative 2^-22 (est * rinv = 1 + 2^-23 at
    // most) and the keys by a relative 2^-24; with |j| <= 2*16384 that stays below 0.01,
    // so floor(lo*est + 1/2) is j or j + 1 and one test against D_fd settles it.
    float fd = floorf(__fadd_rn(__fmul_rn(lo, g.est), 0.5f));
    fd = (lo < key_Df<ARITH>(g, fd)) ? __fsub_rn(fd, 1.0f) : fd;
    fd = __builtin_amdgcn_fmed3f(fd, (float)(-nh - 1), (float)(npix - nh));
    // fr = (index - nh) of the current bin's right key Dhi; the bin is (int)fr + nh - 1
    // (fr is an integer-valued float, the conversion exact)
    float fr = __fadd_rn(fd, 1.0f);
    float Dhi = key_Df<ARITH>(g, fr);
    const int dbase = nh - 1;
#define ITOMO_PIX(acc)                                                                  \
    {                                                                                   \
        x = __fadd_rn(x, 1.0f);                                                         \
        const float hi = __fsub_rn(x, K);                                               \
        const float hc = fminf(hi, Ls);                                                 \
        const int d = (int)fr + dbase;                                                  \
        const float s0 = S[d], s1 = S[d + 1];                                           \
        acc = accum<ARITH>(fmaxf(__fsub_rn(fminf(hc, Dhi), lo), 0.0f), s0, acc);        \
        const float fr2 = __fadd_rn(fr, 1.0f);                                          \
        const float Dhi2 = key_Df<ARITH>(g, fr2);                                       \
        /* one crossing at most: the span ends in (Dhi, Dhi2] when it crosses */        \
        const float R2 = MULTI ? fminf(hc, Dhi2) : hc;                                  \
        acc = accum<ARITH>(fmaxf(__fsub_rn(R2, Dhi), 0.0f), s1, acc);                   \
        const bool cross = Dhi < hi;                                                    \
        fr = cross ? fr2 : fr;                                                          \
        Dhi = cross ? Dhi2 : Dhi;                                                       \
        if (MULTI) {                                                                    \
            int dm = d + 1;                                                             \
            while (cross && dm < npix && Dhi < hi) {                                    \
                ++dm;                                                                   \
                const float Dlo = Dhi;                                                  \
                fr = __fadd_rn(fr, 1.0f);                                               \
                Dhi = key_Df<ARITH>(g, fr);                                             \
                acc = accum<ARITH>(fmaxf(__fsub_rn(fminf(hc, Dhi), Dlo), 0.0f), S[dm],  \
                                   acc);                                                \
            }                                                                           \
        }                                                                               \
        lo = hi;                                                                        \
    }
#pragma unroll
    for (int k = 0; k < P; ++k) ITOMO_PIX(a[k * STRIDE])
#undef ITOMO_PIX
}

// One T x T tile (tile column bx, tile row by) of slice z; blk indexes its FP64
// partials.  Shared by the per-plan kernel and the persistent batch kernel.
template <int ARITH, int EPI, bool ONES, int T, int P>
__device__ __forceinline__ void dd_back_tile(const DdView* __restrict__ geom,
                                             const float* __restrict__ lstar, int nang, int npix,
                                             const float* __restrict__ sino,
                                             float* __restrict__ out, int64_t sino_stride,
                                             int64_t img_stride, int accumulate,
                                             const EpiArgs& e, int bx, int by, int z, int blk)
{
    constexpr int TP = T / P, NT = TP * TP, FOOT = back_foot<T>();
    __shared__ float sS[kBackViews][FOOT];
    __shared__ int sLo[kBackViews];
    __shared__ int sCnt[kBackViews];
    __shared__ float sLs[kBackViews][T];   // L* of the tile's lines per view

    const int tid = threadIdx.x;
    // Threads -> micro-tiles: each wave takes an 8 x 8 block of micro-tiles (2 x 2 waves
    // per 16 x 16 block), so the 32 lanes of a half-wave cover a 32 x 16-pixel patch
    // (P = 4) whose projection onto the detector spans at most 28|c| + 12|s| < 31 bins:
    // the staged bins they read in one step lie in distinct LDS banks.  (Rows of 16
    // micro-tiles per half-wave span up to 60 bins and conflicted.)
    static_assert(TP == 8 || TP == 16, "micro-tile grid");
    const int lane = tid & 63, wv = tid >> 6;
    const int tx = TP == 16 ? ((wv & 1) << 3) + (lane & 7) : tid % TP;
    const int ty = TP == 16 ? ((wv >> 1) << 3) + (lane >> 3) : tid / TP;
    const int r0 = by * T, q0 = bx * T;
    const int rr = r0 + P * ty, cc = q0 + P * tx;  // micro-tile origin
    sino += (int64_t)z * sino_stride;
    out += (int64_t)z * img_stride;
    const int nh = (int)(npix * 0.5);

    float acc[P][P];
#pragma unroll
    for (int i = 0; i < P; ++i)
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int r = rr + i, q = cc + j;
            acc[i][j] = (EPI == EPI_STORE && accumulate && r < npix && q < npix)
                            ? out[(int64_t)r * npix + q]
                            : 0.0f;
        }
    // With small micro-tiles the epilogue inputs (current iterate and weight) are
    // loaded up front, so their latency hides behind the view loop.
    constexpr bool PRE = P <= 2 && (epi_updates_x(EPI) || epi_back_uses_w(EPI));
    float pre_x[P][P], pre_w[P][P];
#pragma unroll
    for (int i = 0; i < P; ++i)
#pragma unroll
        for (int j = 0; j < P; ++j) {
            pre_x[i][j] = 0.0f;
            pre_w[i][j] = 0.0f;
            const int r = rr + i, q = cc + j;
            if (PRE && r < npix && q < npix) {
                const int64_t o = (int64_t)r * npix + q;
                if (epi_updates_x(EPI)) pre_x[i][j] = out[o];
                if (epi_back_uses_w(EPI)) pre_w[i][j] = e.w[(int64_t)z * e.w_stride + o];
            }
        }

    for (int v0 = 0; v0 < nang; v0 += kBackViews) {
        const int nv = min(kBackViews, nang - v0);
        __syncthreads();  // previous chunk fully consumed
        if (tid < nv) {
            int lo, cnt;
            staged_range<ARITH, T>(geom[v0 + tid], npix, nh, r0, q0, &lo, &cnt);
            sLo[tid] = lo;
            sCnt[tid] = cnt;
        }
        __syncthreads();
        for (int k = tid; k < nv * FOOT; k += NT) {
            const int v = k / FOOT, j = k - v * FOOT;
            const int d = sLo[v] + j;
            float val = 0.0f;
            if (j < sCnt[v] && d >= 0 && d < npix)
                val = ONES ? 1.0f : sino[(int64_t)geom[v0 + v].row * npix + d];
            sS[v][j] = val;
        }
        for (int k = tid; k < nv * T; k += NT) {
            const int v = k / T, l = k - v * T;
            const int line = (geom[v0 + v].kind == DD_SWEEP_COLS ? q0 : r0) + l;
            sLs[v][l] = line < npix ? lstar[(int64_t)(v0 + v) * npix + line] : 0.0f;
        }
        __syncthreads();
        for (int v = 0; v < nv; ++v) {
            const DdView g = geom[v0 + v];
            const float* S = &sS[v][0] - sLo[v];
            if (g.kind == DD_SWEEP_ROWS) {
                if (g.multi) {
#pragma unroll
                    for (int i = 0; i < P; ++i)
                        walk_line<ARITH, P, 1, true>(g, npix, nh,
                                                     key_K<ARITH>(g, line_coord(rr + i, nh)),
                                                     sLs[v][P * ty + i], cc, S, &acc[i][0]);
                } else {
#pragma unroll
                    for (int i = 0; i < P; ++i)
                        walk_line<ARITH, P, 1, false>(g, npix, nh,
                                                      key_K<ARITH>(g, line_coord(rr + i, nh)),
                                                      sLs[v][P * ty + i], cc, S, &acc[i][0]);
                }
            } else if (g.kind == DD_SWEEP_COLS) {
                if (g.multi) {
#pragma unroll
                    for (int j = 0; j < P; ++j)
                        walk_line<ARITH, P, P, true>(g, npix, nh,
                                                     key_K<ARITH>(g, line_coord(cc + j, nh)),
                                                     sLs[v][P * tx + j], rr, S, &acc[0][j]);
                } else {
#pragma unroll
                    for (int j = 0; j < P; ++j)
                        walk_line<ARITH, P, P, false>(g, npix, nh,
                                                      key_K<ARITH>(g, line_coord(cc + j, nh)),
                                                      sLs[v][P * tx + j], rr, S, &acc[0][j]);
                }
            } else if (g.kind == DD_ROWSUM) {
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    const float sv = (rr + i < npix) ? S[rr + i] : 0.0f;
#pragma unroll
                    for (int j = 0; j < P; ++j) acc[i][j] = __fadd_rn(acc[i][j], sv);
                }
            } else if (g.kind == DD_COLSUM) {
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    const float sv = (cc + j < npix) ? S[cc + j] : 0.0f;
#pragma unroll
                    for (int i = 0; i < P; ++i) acc[i][j] = __fadd_rn(acc[i][j], sv);
                }
            }
        }
    }

    double p1 = 0.0, p2 = 0.0;
#pragma unroll
    for (int i = 0; i < P; ++i)
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int r = rr + i, q = cc + j;
            if (r < npix && q < npix) {
                const int64_t o = (int64_t)r * npix + q;
                double q1 = 0.0, q2 = 0.0;
                if (PRE)
                    out[o] = epi_back_v<EPI>(acc[i][j], pre_x[i][j], pre_w[i][j], e, &q1, &q2);
                else if (epi_updates_x(EPI))
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

// image[idx .. idx+3] of one slice through the buffer descriptor (out of range: 0).
// idx needs no 16-byte alignment.
__device__ __forceinline__ float4 load_px4(__amdgpu_buffer_rsrc_t rs, int idx)
{
    typedef int int4v __attribute__((ext_vector_type(4)));
    const int4v v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)((unsigned)idx * 4u), 0, 0);
    return make_float4(__int_as_float(v[0]), __int_as_float(v[1]), __int_as_float(v[2]),
                       __int_as_float(v[3]));
}

// v[0 .. NC) = image[idx .. idx+NC) of one slice (total pixels; out of range: 0), idx >= 0,
// with 16-byte loads: a bin's candidates on one line are consecutive pixels, so one load
// instruction fetches four of them.  A run that reaches the end of the slice is read
// element by element.
template <int NC>
__device__ __forceinline__ void load_run(__amdgpu_buffer_rsrc_t rs, int idx, int total,
                                         float* v)
{
#pragma unroll
    for (int k4 = 0; k4 < NC; k4 += 4) {
        const int i = idx + k4;
        float4 f;
        if (i + 3 < total) {
            f = load_px4(rs, i);
        } else {
            f.x = load_px(rs, i);
            f.y = load_px(rs, i + 1);
            f.z = load_px(rs, i + 2);
            f.w = load_px(rs, i + 3);
        }
        v[k4] = f.x;
        if (k4 + 1 < NC) v[k4 + 1] = f.y;
        if (k4 + 2 < NC) v[k4 + 2] = f.z;
        if (k4 + 3 < NC) v[k4 + 3] = f.w;
    }
}

// One candidate pixel of a bin [Dlo, Dhi] on a line with offset K: the pixel spans
// [x-K, x+1-K].  Dhc is the bin's right end clipped to the line's L*.  Advances x.
// The first candidate (Rp = x-K, its left end, on entry) can start left of Dlo and gets
// the weight max(min(P_hi, Dhc) - max(P_lo, Dlo), 0).  Every later one gets
// R_k - R_{k-1} with R_k = min(P_hi(k), Dhc), carried in Rp: its left end is the
// previous candidate's right end P_hi(k-1), and where P_hi(k-1) < Dhc that is the same
// subtraction as max(min(P_hi(k), Dhc) - P_hi(k-1), 0), while past the bin's end both
// give +0 (R_k = R_{k-1} = Dhc) -- so no clamp is needed.  A weight of +0 leaves the
// sum unchanged (every candidate value is a finite image value or a zero).
template <int ARITH>
__device__ __forceinline__ float cand_step(bool first, float& x, float& Rp, float K,
                                           float Dlo, float Dhc, float val, float acc)
{
    x = __fadd_rn(x, 1.0f);
    const float R = fminf(__fsub_rn(x, K), Dhc);
    const float w = first ? fmaxf(__fsub_rn(R, fmaxf(Rp, Dlo)), 0.0f) : __fsub_rn(R, Rp);
    Rp = R;
    return accum<ARITH>(w, val, acc);
}

// (index - nh) of the pixel whose span [P_q, P_q+1) holds Dlo on a line with offset K,
// clamped into [0, npix): a floor estimate, then one exact correction against the keys.
__device__ __forceinline__ float first_pixel(float Dlo, float K, int nh, int npix)
{
    // the estimate needs no clamp of its own: far outside the line the correction may
    // be inexact, but the final clamp then lands on the end pixel either way
    float fe = floorf(__fadd_rn(Dlo, K));
    const float fe1 = __fadd_rn(fe, 1.0f);
    const float Pe = __fsub_rn(fe, K);
    const float Pe1 = __fsub_rn(fe1, K);
    // selects, not branches: at most one of the two corrections applies (Pe <= Pe1)
    const float fm1 = __fsub_rn(fe, 1.0f);
    fe = (Dlo >= Pe1) ? fe1 : fe;
    fe = (Dlo < Pe) ? fm1 : fe;
    return __builtin_amdgcn_fmed3f(fe, (float)(-nh), (float)(npix - 1 - nh));
}

// first_pixel with one correction instead of two, for views whose keys all lie below 2^20
// in magnitude (every view with at most 32 candidates per bin and line, the staged and
// the grouped paths: bins narrower than 30.75 pixels and npix <= 32766 bound |D| and |K|
// by 16384 * 30.75 < 2^19; the chunked path for wider views keeps first_pixel).  There every
// rounding of a key or of the estimate is below 1/32, so with q the exact index
// (P_q <= Dlo < P_q+1) and Dlh = fl(Dlo + 1/2), rounded once per bin, floor(fl(Dlh + K)) is
// q or q + 1 and one test against P_fe settles it.
__device__ __forceinline__ float first_pixel_near(float Dlo, float Dlh, float K, int nh, int npix)
{
    float fe = floorf(__fadd_rn(Dlh, K));
    fe = (Dlo < __fsub_rn(fe, K)) ? __fsub_rn(fe, 1.0f) : fe;
    return __builtin_amdgcn_fmed3f(fe, (float)(-nh), (float)(npix - 1 - nh));
}

constexpr int kFwdLines = 4;   // lines whose gathers are in flight together

// Most candidate pixels per bin and line of a view in the main forward launch; wider
// views run in a launch of their own and read their candidates with 16-byte loads.
constexpr int kFwdMainMaxCand = 8;

// Gather-form forward of one bin over lines [i_lo, i_hi] for views with at most NC
// candidate pixels per bin and line, GL lines per group (see dd_fwd_block).
template <int ARITH, bool ONES, int NC, int GL>
__device__ __forceinline__ float fwd_grouped(const __amdgpu_buffer_rsrc_t rs, const DdView& g,
                                             const float* __restrict__ ls, int i_lo, int i_hi,
                                             int npix, int nh, float Dlo, float Dhi,
                                             bool last_bin, float acc)
{
    const float xNm1 = (float)(npix - 1 - nh);
    const float Dlh = __fadd_rn(Dlo, 0.5f);
    float u = line_coord(i_lo, nh);
    int lbase = i_lo * npix;
    int i = i_lo;
    for (; i + GL - 1 <= i_hi; i += GL) {
        float K[GL], xq[GL], v[GL][NC];
#pragma unroll
        for (int l = 0; l < GL; ++l) {
            K[l] = key_K<ARITH>(g, u);
            u = __fadd_rn(u, 1.0f);
            xq[l] = first_pixel_near(Dlo, Dlh, K[l], nh, npix);
            const int q = (int)xq[l] + nh;
            if (ONES || NC <= kFwdMainMaxCand) {
#pragma unroll
                for (int k = 0; k < NC; ++k) v[l][k] = ONES ? 1.0f : load_px(rs, lbase + q + k);
            } else {
                load_run<NC>(rs, lbase + q, npix * npix, v[l]);
            }
            lbase += npix;
        }
#pragma unroll
        for (int l = 0; l < GL; ++l) {
            float Dhc = Dhi;
            if (last_bin || Dhi > __fsub_rn(xNm1, K[l])) Dhc = fminf(Dhi, ls[i + l]);
            float xk = xq[l], Rp = __fsub_rn(xq[l], K[l]);
#pragma unroll
            for (int k = 0; k < NC; ++k)
                acc = cand_step<ARITH>(k == 0, xk, Rp, K[l], Dlo, Dhc, v[l][k], acc);
        }
    }
    for (; i <= i_hi; ++i, u = __fadd_rn(u, 1.0f), lbase += npix) {
        const float K = key_K<ARITH>(g, u);
        const float x = first_pixel_near(Dlo, Dlh, K, nh, npix);
        const int q = (int)x + nh;
        float Dhc = Dhi;
        if (last_bin || Dhi > __fsub_rn(xNm1, K)) Dhc = fminf(Dhi, ls[i]);
        float v[NC];
        if (ONES || NC <= kFwdMainMaxCand) {
#pragma unroll
            for (int k = 0; k < NC; ++k) v[k] = ONES ? 1.0f : load_px(rs, lbase + q + k);
        } else {
            load_run<NC>(rs, lbase + q, npix * npix, v);
        }
        float xk = x, Rp = __fsub_rn(x, K);
#pragma unroll
        for (int k = 0; k < NC; ++k)
            acc = cand_step<ARITH>(k == 0, xk, Rp, K, Dlo, Dhc, v[k], acc);
    }
    return acc;
}

// LDS of the staged forward path: 16 floats per thread of the block (64 B bytes for B
// threads, 4 KB per wave), shared by every candidate count; the launch of the widest
// views, which runs alone, stages 32 floats per thread.
template <int B>
constexpr int fwd_stage_floats() { return 16 * B; }
template <int B>
constexpr int fwd_stage_floats_wide() { return 32 * B; }
constexpr int kFwdStagedMaxCand = 6;  // views with more candidates take the gather path
constexpr int kFwdStagedWideMaxCand = 32;
constexpr int kFwdStageMaxLines = 8;

// Pixels of the staged window beyond (NC-1)*B, so that it holds every candidate of a
// block (see fwd_staged): max(0, NC + 7 - B/4), rounded up to a multiple of 4.
template <int B, int NC>
constexpr int fwd_window_extra()
{
    return NC + 7 - B / 4 > 0 ? (NC + 7 - B / 4 + 3) / 4 * 4 : 0;
}

// Forward projection of one sweep view with at most NC candidate pixels per bin and line
// (bins narrower than w < NC - 1.25 pixels), for the B bins [bx*B, +B) of a block.
// Lines are processed in chunks of FL: for each line of a chunk the block first copies
// the window of W pixels its bins can touch into LDS with 16-byte loads (one contiguous
// row segment per line; the window starts at q0, the first bin's estimated first pixel
// minus 2, rounded down to a multiple of 4), then every thread reads its candidates from
// LDS.  A bin's first candidate lies less than B*w + 6 pixels past q0: less than
// (B-1)*w + 1 + 5 for a bin that starts on the line, and for the last bin up to one bin
// width more where it starts left of the line's first pixel and its first candidate is
// clamped up to that pixel.  With W = (NC-1)*B + max(0, NC + 7 - B/4) >= B*w + 6 + NC
// the window holds every candidate of every bin that meets the line (offsets are
// clamped into it all the same; a bin that misses the line weights all its candidates
// by zero).  Lines run over the union of the block's per-bin line ranges; lines that
// miss a bin add nothing to it.  Every thread of the block takes part (inactive
// threads, past the last bin, compute a copy of the last bin and store nothing).
// sImg holds FLOATS floats; sQ0, sQb and sL kFwdStageMaxLines entries.
template <int ARITH, int B, int NC, int FLOATS>
__device__ __forceinline__ float fwd_staged(const DdView& g, const float* __restrict__ ls,
                                            int npix, int nh, __amdgpu_buffer_rsrc_t rs,
                                            int bx, int d, float acc, float* sImg, int* sQ0,
                                            int* sQb, float* sL)
{
    constexpr int W = (NC - 1) * B + fwd_window_extra<B, NC>();
    constexpr int FL = FLOATS / W < kFwdStageMaxLines ? FLOATS / W : kFwdStageMaxLines;
    constexpr int N4 = FL * (W / 4);          // 16-byte pieces per chunk
    constexpr int PER = (N4 + B - 1) / B;     // pieces per thread
    static_assert(FL >= 1 && W % 4 == 0, "staged window shape");
    const int tid = threadIdx.x;
    const int dk = min(d, npix - 1);
    const float Dlo = key_D<ARITH>(g, dk - nh);
    const float Dhi = key_D<ARITH>(g, dk + 1 - nh);
    const float Dlh = __fadd_rn(Dlo, 0.5f);
    const int d0 = bx * B, d1 = min(d0 + B - 1, npix - 1);
    const float D0lo = key_D<ARITH>(g, d0 - nh);
    const float D1hi = key_D<ARITH>(g, d1 + 1 - nh);
    const float x0 = (float)(-nh), xN = (float)(npix - nh);
    double ua = ((double)x0 - (double)D1hi) / (double)g.t;
    double ub = ((double)xN - (double)D0lo) / (double)g.t;
    if (ua > ub) { const double tmp = ua; ua = ub; ub = tmp; }
    ua = fmin(fmax(ua + nh - 0.5, -4.0), (double)npix + 4.0);
    ub = fmin(fmax(ub + nh - 0.5, -4.0), (double)npix + 4.0);
    const int i_lo = max(0, (int)floor(ua) - 2);
    const int i_hi = min(npix - 1, (int)ceil(ub) + 2);
    float u = line_coord(i_lo, nh);
    for (int c0 = i_lo; c0 <= i_hi; c0 += FL) {
        const int ncl = min(FL, i_hi - c0 + 1);
        __syncthreads();  // the previous chunk is consumed
        if (tid < ncl) {
            const float K = key_K<ARITH>(g, line_coord(c0 + tid, nh));
            const float est = fminf(fmaxf(floorf(__fadd_rn(D0lo, K)), (float)(-nh - W)),
                                    (float)(npix - nh + W));
            const int q0 = ((int)est + nh - 2) & ~3;
            sQ0[tid] = q0;
            sQb[tid] = 4 * (nh - q0 + tid * W);   // byte offset of pixel 0 in sImg
            sL[tid] = ls[c0 + tid];
        }
        __syncthreads();
        // every thread loads its window pieces into registers first, then stores them:
        // the loads of one chunk are all in flight together
        float4 v[PER];
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int k = tid + r * B;
            const int l = k / (W / 4), j4 = k - l * (W / 4);
            v[r] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (k < N4 && l < ncl) {
                const int q = sQ0[l] + 4 * j4;
                // a window that starts left of the row reads zeros there (negative index
                // -> out of range); past the row's end it reads the next row: never
                // weighted
                const int idx = (c0 + l) * npix + q;
                if (q >= 0 && idx + 3 < npix * npix) {
                    v[r] = load_px4(rs, idx);
                } else if (q >= 0) {  // straddles the end of the slice: per-element checks
                    v[r].x = load_px(rs, idx);
                    v[r].y = load_px(rs, idx + 1);
                    v[r].z = load_px(rs, idx + 2);
                    v[r].w = load_px(rs, idx + 3);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int k = tid + r * B;
            const int l = k / (W / 4), j4 = k - l * (W / 4);
            if (k < N4 && l < ncl) *(float4*)&sImg[l * W + 4 * j4] = v[r];
        }
        __syncthreads();
#pragma unroll
        for (int l = 0; l < FL; ++l) {
            if (l < ncl) {
                const float K = key_K<ARITH>(g, u);
                const float x = first_pixel_near(Dlo, Dlh, K, nh, npix);
                // byte offset of the first candidate in sImg, clamped into the line's window
                const int ob = min(max(((int)x << 2) + sQb[l], 4 * l * W), 4 * (l * W + W - NC));
                const float* wp = (const float*)((const char*)sImg + ob);
                // clipping every weight's right end to L* is the exact rule; it changes
                // only the weights that end at the line's dropped segment
                const float Dhc = fminf(Dhi, sL[l]);
                float xk = x, Rp = __fsub_rn(x, K);
#pragma unroll
                for (int k = 0; k < NC; ++k)
                    acc = cand_step<ARITH>(k == 0, xk, Rp, K, Dlo, Dhc, wp[k], acc);
            }
            u = __fadd_rn(u, 1.0f);
        }
    }
    return acc;
}

// Forward projection: one thread per (bin, view blockIdx.y, slice blockIdx.z).
// Lines are image rows for row sweeps and image columns for column sweeps; column
// sweeps (and row sums) read the transposed image so that the lanes of a wave, which
// hold consecutive bins, read consecutive addresses.
// Bins [bx*B, +B) of view `view`, slice z, one per thread; blk indexes the partials.
// Which views a forward launch handles (block-uniform): all of them, the views with at
// most kFwdMainMaxCand candidate pixels per bin and line (kPathMain), or the wider ones
// (kPathWide).  The wide views' larger staging windows and deep load groups need LDS and
// registers; compiled into a kernel of their own they leave the main kernel's occupancy
// alone.
constexpr int kPathAll = 0, kPathMain = 1, kPathWide = 2;

// Whether the grouped gather with NC candidates is compiled into launch PATH.
constexpr bool fwd_in_path(int path, int nc)
{
    return path == kPathAll || ((path == kPathMain) == (nc <= kFwdMainMaxCand));
}

template <int ARITH, int EPI, bool ONES, int B, int PATH = kPathAll>
__device__ __forceinline__ void dd_fwd_block(const DdView* __restrict__ geom,
                                             const float* __restrict__ lstar, int nang,
                                             int npix, const float* __restrict__ image,
                                             const float* __restrict__ imageT,
                                             float* __restrict__ out, int64_t img_stride,
                                             int64_t sino_stride, int accumulate,
                                             const EpiArgs& e, int bx, int view, int z)
{
    const int d = bx * B + threadIdx.x;
    const bool active = d < npix;
    const DdView g = geom[view];
    image += (int64_t)z * img_stride;
    imageT += (int64_t)z * img_stride;
    const int64_t soff = (int64_t)z * sino_stride;
    const int64_t o = (int64_t)g.row * npix + d;
    const int nh = (int)(npix * 0.5);
    double p0 = 0.0;

    const bool sweep = g.kind == DD_SWEEP_ROWS || g.kind == DD_SWEEP_COLS;
    const bool staged =
        !ONES && sweep &&
        (PATH == kPathWide ? g.ncand <= kFwdStagedWideMaxCand : g.ncand <= kFwdStagedMaxCand);
    if (staged) {
        // block-uniform: every thread joins the LDS-staged path; the epilogue inputs
        // are loaded first so that their latency hides behind the sweep
        constexpr int FLOATS =
            PATH == kPathWide ? fwd_stage_floats_wide<B>() : fwd_stage_floats<B>();
        __shared__ __attribute__((aligned(16))) float sImg[FLOATS];
        __shared__ int sQ0[kFwdStageMaxLines], sQb[kFwdStageMaxLines];
        __shared__ float sL[kFwdStageMaxLines];
        const EpiFwdIn pre = active ? epi_fwd_load<EPI>(o, soff, z, e) : EpiFwdIn{0.0f, 0.0f};
        const __amdgpu_buffer_rsrc_t rs =
            image_rsrc((g.kind == DD_SWEEP_ROWS) ? image : imageT, npix);
        float acc = (EPI == EPI_STORE && accumulate && active) ? out[soff + o] : 0.0f;
        const float* ls = lstar + (int64_t)view * npix;
        const int nc = g.ncand;
        if constexpr (PATH == kPathWide) {
            // the launch of views with more than 8 candidates; up to 32 are staged (one
            // or two lines per chunk), wider views take the gather path
            static_assert(kFwdStagedWideMaxCand == 32, "staged candidate counts");
            if (nc <= 12)
                acc = fwd_staged<ARITH, B, 12, FLOATS>(g, ls, npix, nh, rs, bx, d, acc, sImg,
                                                       sQ0, sQb, sL);
            else if (nc <= 16)
                acc = fwd_staged<ARITH, B, 16, FLOATS>(g, ls, npix, nh, rs, bx, d, acc, sImg,
                                                       sQ0, sQb, sL);
            else if (nc <= 24)
                acc = fwd_staged<ARITH, B, 24, FLOATS>(g, ls, npix, nh, rs, bx, d, acc, sImg,
                                                       sQ0, sQb, sL);
            else
                acc = fwd_staged<ARITH, B, 32, FLOATS>(g, ls, npix, nh, rs, bx, d, acc, sImg,
                                                       sQ0, sQb, sL);
        } else {
            // 3, 4 or up to kFwdStagedMaxCand candidates (staging 7-8 measured slower than
            // the gather path)
            static_assert(kFwdStagedMaxCand == 6, "staged candidate counts");
            if (nc <= 3)
                acc = fwd_staged<ARITH, B, 3, FLOATS>(g, ls, npix, nh, rs, bx, d, acc, sImg,
                                                      sQ0, sQb, sL);
            else if (nc == 4)
                acc = fwd_staged<ARITH, B, 4, FLOATS>(g, ls, npix, nh, rs, bx, d, acc, sImg,
                                                      sQ0, sQb, sL);
            else
                acc = fwd_staged<ARITH, B, 6, FLOATS>(g, ls, npix, nh, rs, bx, d, acc, sImg,
                                                      sQ0, sQb, sL);
        }
        if (active) out[soff + o] = epi_forward<EPI>(acc, pre, &p0);
    } else if (active) {
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
            // x0 - Dhi < K_i < xN - Dlo with K_i ~ (i - nh + 0.5) * t, two lines of margin
            // (margin lines that miss the bin add nothing).
            double ua = ((double)x0 - (double)Dhi) / (double)g.t;
            double ub = ((double)xN - (double)Dlo) / (double)g.t;
            if (ua > ub) { const double tmp = ua; ua = ub; ub = tmp; }
            ua = fmin(fmax(ua + nh - 0.5, -4.0), (double)npix + 4.0);
            ub = fmin(fmax(ub + nh - 0.5, -4.0), (double)npix + 4.0);
            const int i_lo = max(0, (int)floor(ua) - 2);
            const int i_hi = min(npix - 1, (int)ceil(ub) + 2);
            // Lines outside the bin, and candidates past the bin or past the line's end,
            // need no test: the first pixel is clamped into [0, npix), and every weight's
            // right end is clipped to L* <= P[N], so such candidates have R <= L.
            // L* >= min(P[N-1], D[N-1]) on every line, so the clip (and the table read)
            // is needed only where Dhi exceeds that bound: the line's last pixel or the
            // last bin.
            const int nc = g.ncand;
            const bool last_bin = d == npix - 1;
            const float xNm1 = (float)(npix - 1 - nh);
            float u = line_coord(i_lo, nh);
            int lbase = i_lo * npix;
            // Up to 16 candidates: lines in groups, all their loads issued before the first
            // is consumed, so a wave keeps several gathers in flight; the lines are then
            // accumulated in increasing order.  Candidates past the bin add nothing.
            // Instances whose candidate counts the staged path takes in this launch are
            // left out (they would only add registers).
            constexpr int SMAX = ONES ? 0
                                 : (PATH == kPathWide ? kFwdStagedWideMaxCand
                                                      : kFwdStagedMaxCand);
            if (fwd_in_path(PATH, 3) && 3 > SMAX && nc <= 3)
                acc = fwd_grouped<ARITH, ONES, 3, kFwdLines>(rs, g, ls, i_lo, i_hi, npix, nh, Dlo,
                                                           Dhi, last_bin, acc);
            else if (fwd_in_path(PATH, 4) && 4 > SMAX && nc == 4)
                acc = fwd_grouped<ARITH, ONES, 4, 6>(rs, g, ls, i_lo, i_hi, npix, nh, Dlo, Dhi,
                                                     last_bin, acc);
            else if (fwd_in_path(PATH, 6) && 6 > SMAX && nc <= 6)
                acc = fwd_grouped<ARITH, ONES, 6, 4>(rs, g, ls, i_lo, i_hi, npix, nh, Dlo, Dhi,
                                                     last_bin, acc);
            else if (fwd_in_path(PATH, 8) && 8 > SMAX && nc <= 8)
                acc = fwd_grouped<ARITH, ONES, 8, 3>(rs, g, ls, i_lo, i_hi, npix, nh, Dlo, Dhi,
                                                     last_bin, acc);
            else if (fwd_in_path(PATH, 12) && 12 > SMAX && nc <= 12)
                acc = fwd_grouped<ARITH, ONES, 12, 2>(rs, g, ls, i_lo, i_hi, npix, nh, Dlo, Dhi,
                                                      last_bin, acc);
            else if (fwd_in_path(PATH, 16) && 16 > SMAX && nc <= 16)
                acc = fwd_grouped<ARITH, ONES, 16, 2>(rs, g, ls, i_lo, i_hi, npix, nh, Dlo, Dhi,
                                                      last_bin, acc);
            else {
                for (int i = i_lo; i <= i_hi; ++i, u = __fadd_rn(u, 1.0f), lbase += npix) {
                    const float K = key_K<ARITH>(g, u);
                    const float x = first_pixel(Dlo, K, nh, npix);
                    const int q = (int)x + nh;
                    float Dhc = Dhi;
                    if (last_bin || Dhi > __fsub_rn(xNm1, K)) Dhc = fminf(Dhi, ls[i]);
                    float xk = x, Rp = __fsub_rn(x, K);
                    for (int k0 = 0; k0 < nc; k0 += 4) {
                        float v[4] = {1.0f, 1.0f, 1.0f, 1.0f};
                        if (!ONES) load_run<4>(rs, lbase + q + k0, npix * npix, v);
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            acc = cand_step<ARITH>(k0 + k == 0, xk, Rp, K, Dlo, Dhc, v[k], acc);
                    }
                }
            }
        }
        out[soff + o] = epi_forward<EPI>(acc, o, soff, z, e, &p0);
    }
    // one FP64 slot per 64-bin chunk of the view (the layout of dd_fwd_slots); a wave
    // past the last bin has no chunk and writes nothing
    const int chunk = bx * (B / 64) + (int)(threadIdx.x >> 6);
    if (epi_has_partials(EPI) && e.partials && chunk * 64 < npix)
        wave_partials3(p0, 0.0, 0.0,
                       e.partials + ((int64_t)z * e.blocks_per_slice +
                                     (int64_t)view * ((npix + 63) / 64) + chunk) * 3);
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
