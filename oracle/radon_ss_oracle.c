/*
 * CPU restatement of the reference's slant-stacking projector.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP slant-stacking
 * projector is compared against; nothing in the product library links it or calls it.
 *
 * Restated from (paths relative to the reference repository):
 *   common/operators/radon_module/pymodule_radon_slant_stacking/radon_ss.c:43-196
 *
 * Geometry.  Bin k = t + nh (t = -nh .. nh-1) of view v holds the line x c + y s = t
 * (s, c = sin, cos of the view); the result goes to sinogram row nang-1-v
 * (radon_ss.c:111).  ray_limits (radon_ss.c:43-88) clips the line to the square
 * [-nh, nh-1]^2: the x range is the middle two of {x(y=-nh), x(y=nh-1), -nh, nh-1}
 * sorted, the y range likewise (for |s| < 1e-5 or |c| < 1e-5 fixed ranges).  Views with
 * |s| <= sqrt(1/2) step y over [round(ymin), round(ymax)) and sample x = (t - y s)/c on
 * image row y + nh; the others step x and sample y = (t - x c)/s on column x + nh
 * (radon_ss.c:120-192).  Each sample adds 1/|c| (or 1/|s|) times the image value at the
 * nearest pixel (method 0) or at the two neighbours with linear weights (method 1; the
 * reference checks both indices against [0, npix)).  Image rows are not flipped.
 *
 * Out-of-square samples.  For a line that misses the square (|t| close to nh; the
 * square [-nh, nh-1]^2 is not centred) the middle two of the sorted limits lie outside
 * [-nh, nh-1], and the reference then steps rows (or columns) outside the image: its
 * flat index (y + nh)*npix + x wraps into other image rows or leaves the array
 * (out-of-bounds reads in the forward projection, writes in the back projection).
 * method 0 also checks no sampled index.  Here every sample whose stepped or sampled
 * index falls outside [0, npix) is skipped; this equals the reference on every bin
 * whose samples all fall inside the square.  oracle_ss_edge_bins() marks the other
 * bins and reports the flat-index extent the reference reaches, so that the reference
 * can be run on a guard-padded image for comparison.
 *
 * Two float32 arithmetic forms are restated, selected by `arith`:
 *   0 (IEEE): the C as written (float divisions, double 1/|c|, libm round/floor);
 *   1 (FAST): the C as gcc 11 -O3 -ffast-math compiles it (the reference's flags,
 *     create_module.py:14), read from the assembly of that build: sincosf; limits as
 *     fma(nh, s, t)/c etc.; x = fma(-y, s, t) * (1/c); round(v) as
 *     trunc(v + copysign(0.49999997f, v)) in float; the method-1 weight 1 - w as
 *     (1 - x) + floor(x); every update as one double FMA (forward: (w/|c|) * img + sino;
 *     back: w * (sino/|c|) + img).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stddef.h>

static const double kS45 = 0.70710678118654757;   /* radon_ss.c:12 */
static const double kEps = 1e-5;                   /* radon_ss.c:11 */

static int round_f_fast(float v)
{
    return (int)truncf(v + copysignf(0.49999997f, v));
}

/* Middle two of {p, q, lo, hi} (lo <= hi) in ascending order. */
static void middle_two(float p, float q, float lo, float hi, float *a, float *b)
{
    const float mn = p < q ? p : q, mx = p < q ? q : p;
    const float m1 = mn > lo ? mn : lo, m2 = mx < hi ? mx : hi;
    *a = m1 < m2 ? m1 : m2;       /* second smallest */
    *b = m1 < m2 ? m2 : m1;       /* third smallest */
}

/* Step range [*k0, *k1) of bin t (radon_ss.c:43-88 and the loop bounds 121/159). */
static void ss_range(int t, float s, float c, int nh, int ybranch, int fast, int *k0, int *k1)
{
    float lim[4];
    const float tf = (float)t, nhf = (float)nh, nh1 = (float)(nh - 1), mnh = (float)(-nh);
    if (fabs((double)s) < kEps) {
        lim[0] = 0.0f; lim[1] = 0.0f; lim[2] = mnh; lim[3] = nh1;
    } else if (fabs((double)c) < kEps) {
        lim[0] = mnh; lim[1] = nh1; lim[2] = 0.0f; lim[3] = 0.0f;
    } else {
        float x1, x2, y1, y2;
        if (fast) {
            x1 = fmaf(nhf, s, tf) / c;
            x2 = fmaf(-s, nh1, tf) / c;
            y1 = fmaf(nhf, c, tf) / s;
            y2 = fmaf(-c, nh1, tf) / s;
        } else {
            x1 = tf / c + (nhf * s) / c;
            x2 = tf / c - (nh1 * s) / c;
            y1 = tf / s + (nhf * c) / s;
            y2 = tf / s - (nh1 * c) / s;
        }
        middle_two(x1, x2, mnh, nh1, &lim[0], &lim[1]);
        middle_two(y1, y2, mnh, nh1, &lim[2], &lim[3]);
    }
    const float lo = ybranch ? lim[2] : lim[0], hi = ybranch ? lim[3] : lim[1];
    *k0 = fast ? round_f_fast(lo) : (int)round((double)lo);
    *k1 = fast ? round_f_fast(hi) : (int)round((double)hi);
}

int oracle_radon_ss(float *image, int npix, const float *angles, int nang, int oper,
                    int method, float *sino, int arith)
{
    const int nh = (int)(npix * 0.5);
    const int fast = arith == 1;
    for (int v = 0; v < nang; ++v) {
        const float theta = angles[v];
        float s, c;
        if (fast) {
            sincosf(theta, &s, &c);
        } else {
            s = (float)sin((double)theta);
            c = (float)cos((double)theta);
        }
        const int ybranch = !(kS45 < fabs((double)s));
        /* the stepped coordinate u, the sampled one q: q = (t - u a) / b */
        const float a = ybranch ? s : c, b = ybranch ? c : s;
        const float rb = 1.0f / b;
        const double rab = 1.0 / fabs((double)b);
        float *row = sino + (size_t)(nang - 1 - v) * npix;
        for (int t = -nh; t < nh; ++t) {
            float *bin = row + (t + nh);
            int u0, u1;
            ss_range(t, s, c, nh, ybranch, fast, &u0, &u1);
            for (int u = u0; u < u1; ++u) {
                if (u + nh < 0 || u + nh >= npix) continue;   /* outside the square */
                const float qf = fast ? fmaf(-(float)u, a, (float)t) * rb
                                      : (float)t / b - ((float)u * a) / b;
                /* pixel (sampled index p, stepped index u + nh) */
#define PIX(p) (ybranch ? image + (size_t)(u + nh) * npix + (p) \
                        : image + (size_t)(p) * npix + (u + nh))
                if (method == 0) {
                    const int p = (fast ? round_f_fast(qf) : (int)round((double)qf)) + nh;
                    if (p < 0 || p >= npix) continue;     /* outside the square */
                    float *px = PIX(p);
                    if (oper == 0)
                        *bin = fast ? (float)fma((double)*px, rab, (double)*bin)
                                    : (float)((double)*bin + rab * (double)*px);
                    else
                        *px = fast ? (float)fma((double)*bin, rab, (double)*px)
                                   : (float)((double)*px + rab * (double)*bin);
                } else {
                    const float fl = fast ? floorf(qf) : (float)floor((double)qf);
                    const float w = fast ? qf - fl : (float)((double)qf - floor((double)qf));
                    const float wl = fast ? (1.0f - qf) + fl : 1.0f - w;
                    const int p = (int)fl + nh;
                    if (p >= 0 && p <= npix - 1) {
                        float *px = PIX(p);
                        if (oper == 0)
                            *bin = fast ? (float)fma((double)wl * rab, (double)*px, (double)*bin)
                                        : (float)((double)*bin + rab * (double)wl * (double)*px);
                        else
                            *px = fast ? (float)fma((double)wl, (double)*bin * rab, (double)*px)
                                       : (float)((double)*px + rab * (double)wl * (double)*bin);
                    }
                    if (p + 1 >= 0 && p + 1 <= npix - 1) {
                        float *px = PIX(p + 1);
                        if (oper == 0)
                            *bin = fast ? (float)fma((double)w * rab, (double)*px, (double)*bin)
                                        : (float)((double)*bin + rab * (double)w * (double)*px);
                        else
                            *px = fast ? (float)fma((double)w, (double)*bin * rab, (double)*px)
                                       : (float)((double)*px + rab * (double)w * (double)*bin);
                    }
                }
#undef PIX
            }
        }
    }
    return 0;
}

/* Marks flags[row*npix + bin] = 1 for the bins where the reference steps or samples
 * outside the square (see above) and returns in ext[0], ext[1] the smallest and largest
 * flat image index the reference reaches over all bins (for sizing guard rows). */
int oracle_ss_edge_bins(int npix, const float *angles, int nang, int method, int arith,
                        unsigned char *flags, long long *ext)
{
    const int nh = (int)(npix * 0.5);
    const int fast = arith == 1;
    long long lo = 0, hi = 0;
    for (int v = 0; v < nang; ++v) {
        const float theta = angles[v];
        float s, c;
        if (fast) {
            sincosf(theta, &s, &c);
        } else {
            s = (float)sin((double)theta);
            c = (float)cos((double)theta);
        }
        const int ybranch = !(kS45 < fabs((double)s));
        const float a = ybranch ? s : c, b = ybranch ? c : s;
        const float rb = 1.0f / b;
        for (int t = -nh; t < nh; ++t) {
            unsigned char *fl = flags + (size_t)(nang - 1 - v) * npix + (t + nh);
            *fl = 0;
            int u0, u1;
            ss_range(t, s, c, nh, ybranch, fast, &u0, &u1);
            for (int u = u0; u < u1; ++u) {
                const float qf = fast ? fmaf(-(float)u, a, (float)t) * rb
                                      : (float)t / b - ((float)u * a) / b;
                int p0, p1;
                if (method == 0) {
                    p0 = p1 = (fast ? round_f_fast(qf) : (int)round((double)qf)) + nh;
                } else {
                    p0 = (int)(fast ? floorf(qf) : (float)floor((double)qf)) + nh;
                    p1 = p0 + 1;
                    /* the reference checks these against [0, npix) */
                    if (p0 < 0) p0 = p1;
                    if (p1 > npix - 1) p1 = p0;
                    if (p0 < 0 || p0 > npix - 1) continue;
                }
                const int su = u + nh;
                if (su < 0 || su >= npix || p0 < 0 || p1 >= npix) *fl = 1;
                for (int k = 0; k < 2; ++k) {
                    const int p = k ? p1 : p0;
                    const long long f = ybranch ? (long long)su * npix + p
                                                : (long long)p * npix + su;
                    if (f < lo) lo = f;
                    if (f > hi) hi = f;
                }
            }
        }
    }
    ext[0] = lo;
    ext[1] = hi;
    return 0;
}

/* The same sampling in double precision (IEEE forms, double limits). */
int oracle_radon_ss_f64(double *image, int npix, const double *angles, int nang, int oper,
                        int method, double *sino)
{
    const int nh = (int)(npix * 0.5);
    for (int v = 0; v < nang; ++v) {
        const double theta = angles[v], s = sin(theta), c = cos(theta);
        const int ybranch = !(kS45 < fabs(s));
        const double a = ybranch ? s : c, b = ybranch ? c : s, rab = 1.0 / fabs(b);
        double *row = sino + (size_t)(nang - 1 - v) * npix;
        for (int t = -nh; t < nh; ++t) {
            double lo, hi;
            if (fabs(s) < kEps) {
                lo = ybranch ? -nh : 0; hi = ybranch ? nh - 1 : 0;
            } else if (fabs(c) < kEps) {
                lo = ybranch ? 0 : -nh; hi = ybranch ? 0 : nh - 1;
            } else {
                /* ends of the line on the two walls u = -nh and u = nh-1 of the stepped
                   coordinate's partner, clipped to [-nh, nh-1] */
                const double e1 = (t + nh * b) / a, e2 = (t - (nh - 1) * b) / a;
                const double mn = fmin(e1, e2), mx = fmax(e1, e2);
                const double m1 = fmax(mn, -nh), m2 = fmin(mx, nh - 1);
                lo = fmin(m1, m2); hi = fmax(m1, m2);
            }
            const int u0 = (int)round(lo), u1 = (int)round(hi);
            for (int u = u0; u < u1; ++u) {
                if (u + nh < 0 || u + nh >= npix) continue;
                const double qf = (t - u * a) / b;
#define PIXD(p) (ybranch ? image + (size_t)(u + nh) * npix + (p) \
                         : image + (size_t)(p) * npix + (u + nh))
                if (method == 0) {
                    const int p = (int)round(qf) + nh;
                    if (p < 0 || p >= npix) continue;
                    if (oper == 0) row[t + nh] += rab * *PIXD(p);
                    else *PIXD(p) += rab * row[t + nh];
                } else {
                    const double fl = floor(qf), w = qf - fl;
                    const int p = (int)fl + nh;
                    if (p >= 0 && p <= npix - 1) {
                        if (oper == 0) row[t + nh] += rab * (1 - w) * *PIXD(p);
                        else *PIXD(p) += rab * (1 - w) * row[t + nh];
                    }
                    if (p + 1 >= 0 && p + 1 <= npix - 1) {
                        if (oper == 0) row[t + nh] += rab * w * *PIXD(p + 1);
                        else *PIXD(p + 1) += rab * w * row[t + nh];
                    }
                }
#undef PIXD
            }
        }
    }
    return 0;
}
