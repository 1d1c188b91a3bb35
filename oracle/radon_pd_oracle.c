/*
 * CPU restatement of the reference's pixel-driven projector.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP pixel-driven
 * projector is compared against; nothing in the product library links it or calls it.
 *
 * Restated from (paths relative to the reference repository):
 *   common/operators/radon_module/pymodule_radon_pixel_driven/radon_pd.c:12-90
 *
 * Like the reference C it accumulates into its output array (the Cython wrapper hands
 * it a freshly zeroed one, radon_pixel_driven.pyx:38,66).  View v writes sinogram row v
 * (no reversal); pixel (i, j) of the loops is image[(npix-1-i)*npix + j].  Each pixel
 * is sampled at 4 points, of which only two are distinct: the reference indexes
 * delta[2k+1] for both coordinates (radon_pd.c:40-41), i.e. offsets (-1/4,-1/4) twice and
 * (+1/4,+1/4) twice.
 *
 * method 0 (nearest bin): the reference indexes sino[v*npix + l] with l = round(nh - 0.5
 * - t) and no range check (radon_pd.c:58-63).  Corner pixels have l < 0 or l >= npix, so
 * their samples land in the neighbouring sinogram row (flat index v*npix + l), or outside
 * the array for the first and last rows.  This restatement keeps the in-array part of
 * that behaviour exactly (forward: added to the neighbouring row; back: read from it) and
 * skips flat indices outside [0, nang*npix), where the reference's behaviour is
 * undefined.  method 1 (linear) keeps the reference's own check 0 < bin < npix (bin 0
 * never receives anything, radon_pd.c:74-83).
 *
 * Two float32 arithmetic forms are restated, selected by `arith`:
 *   0 (IEEE): the C as written (float products, double where the C promotes, libm
 *     round/floor/ceil, no contraction);
 *   1 (FAST): the C as gcc 11 -O3 -ffast-math compiles it (the reference's flags,
 *     create_module.py:14), read from the assembly of that build: sincosf; x0 and y0 as
 *     j + (0.5f - nh); t = |fma(-c, y, x*s)|; x*s/c as (x*s)*(1/c); round(v) as
 *     trunc(v + copysign(0.49999999999999994, v)); floor/ceil in float; the linear
 *     weight of the lower bin as |(t - nh) + uf|; the accumulations as double FMAs.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stddef.h>

static const double kPi = 3.141592653589793;   /* radon_pd.c:10 */
static const float kDelta[8] = {-0.25f, -0.25f, 0.25f, -0.25f,
                                -0.25f, 0.25f, 0.25f, 0.25f};   /* radon_pd.c:12-15 */

static double round_fast(double v)
{
    return trunc(v + copysign(0.49999999999999994, v));
}

/* Keep or negate t (radon_pd.c:44-53), as the compiled conditions evaluate it. */
static float sign_t(float t, float theta, float x, float y, float q)
{
    const double th = (double)theta;
    if (th < kPi / 2 && y > q) return t;
    if (th > kPi / 2 && y > q) return t;
    if (theta == 0.0f && y > 0.0f) return t;
    if (th == kPi / 2 && x < 0.0f) return t;
    return -t;
}

int oracle_radon_pd(float *image, int npix, const float *angles, int nang, int oper,
                    int method, float *sino, int arith)
{
    const int nh = (int)(npix * 0.5);
    const int fast = arith == 1;
    const float nhf = (float)nh;
    for (int v = 0; v < nang; ++v) {
        const float theta = angles[v];
        float s, c;
        if (fast) {
            sincosf(theta, &s, &c);
        } else {
            s = (float)sin((double)theta);
            c = (float)cos((double)theta);
        }
        const float rc = 1.0f / c;
        float *row = sino + (size_t)v * npix;
        for (int i = 0; i < npix; ++i) {
            for (int j = 0; j < npix; ++j) {
                float *pix = image + (size_t)(npix - 1 - i) * npix + j;
                float x0, y0;
                if (fast) {
                    x0 = (float)j + (0.5f - nhf);
                    y0 = (float)i + (0.5f - nhf);
                } else {
                    x0 = (float)((double)((float)j - nhf) + 0.5);
                    y0 = (float)((double)((float)i - nhf) + 0.5);
                }
                for (int k = 0; k < 4; ++k) {
                    const float d = kDelta[2 * k + 1];
                    float x, y, t, q;
                    if (fast) {
                        x = d + x0;
                        y = d + y0;
                        const float xs = x * s;
                        t = fabsf(fmaf(-c, y, xs));
                        q = xs * rc;
                    } else {
                        x = x0 + d;
                        y = y0 + d;
                        t = fabsf(x * s - y * c);
                        q = (x * s) / c;
                    }
                    t = sign_t(t, theta, x, y, q);
                    if (method == 0) {
                        const double vl = ((double)nh - 0.5) - (double)t;
                        const int l = (int)(fast ? round_fast(vl) : round(vl));
                        const ptrdiff_t f = (ptrdiff_t)v * npix + l;
                        if (f < 0 || f >= (ptrdiff_t)nang * npix) continue;   /* undefined */
                        if (oper == 0)
                            sino[f] = fast ? (float)fma((double)*pix, 0.25, (double)sino[f])
                                           : (float)((double)sino[f] + 0.25 * (double)*pix);
                        else
                            *pix = fast ? (float)fma((double)sino[f], 0.25, (double)*pix)
                                        : (float)((double)*pix + 0.25 * (double)sino[f]);
                    } else {
                        const float t1 = nhf - t;
                        float lf, uf, wl, wu;
                        int l, u;
                        if (fast) {
                            lf = floorf(t1);
                            uf = ceilf(t1);
                            l = (int)round_fast((double)lf - 0.5);
                            u = (int)round_fast((double)uf - 0.5);
                            wl = fabsf((t - nhf) + uf);
                            wu = fabsf(t1 - lf);
                        } else {
                            lf = (float)floor((double)t1);
                            uf = (float)ceil((double)t1);
                            l = (int)round((double)lf - 0.5);
                            u = (int)round((double)uf - 0.5);
                            wl = fabsf(uf - t1);
                            wu = fabsf(t1 - lf);
                        }
                        if (l > 0 && l < npix) {
                            if (oper == 0)
                                row[l] = fast ? (float)fma((double)wl, 0.25 * (double)*pix, (double)row[l])
                                              : (float)((double)row[l] + 0.25 * (double)*pix * (double)wl);
                            else
                                *pix = fast ? (float)fma((double)row[l], (double)wl * 0.25, (double)*pix)
                                            : (float)((double)*pix + 0.25 * (double)row[l] * (double)wl);
                        }
                        if (u > 0 && u < npix) {
                            if (oper == 0)
                                row[u] = fast ? (float)fma((double)wu, 0.25 * (double)*pix, (double)row[u])
                                              : (float)((double)row[u] + 0.25 * (double)*pix * (double)wu);
                            else
                                *pix = fast ? (float)fma((double)row[u], (double)wu * 0.25, (double)*pix)
                                            : (float)((double)*pix + 0.25 * (double)row[u] * (double)wu);
                        }
                    }
                }
            }
        }
    }
    return 0;
}

/* The same sampling and bin rule in double precision (no fast-math forms). */
int oracle_radon_pd_f64(double *image, int npix, const double *angles, int nang, int oper,
                        int method, double *sino)
{
    const int nh = (int)(npix * 0.5);
    for (int v = 0; v < nang; ++v) {
        const double theta = angles[v], s = sin(theta), c = cos(theta);
        double *row = sino + (size_t)v * npix;
        for (int i = 0; i < npix; ++i)
            for (int j = 0; j < npix; ++j) {
                double *pix = image + (size_t)(npix - 1 - i) * npix + j;
                const double x0 = j - nh + 0.5, y0 = i - nh + 0.5;
                for (int k = 0; k < 4; ++k) {
                    const double d = kDelta[2 * k + 1], x = x0 + d, y = y0 + d;
                    double t = fabs(x * s - y * c);
                    const double q = x * s / c;
                    if (!((theta < kPi / 2 && y > q) || (theta > kPi / 2 && y > q) ||
                          (theta == 0 && y > 0) || (theta == kPi / 2 && x < 0)))
                        t = -t;
                    if (method == 0) {
                        const ptrdiff_t f = (ptrdiff_t)v * npix + (int)round(nh - 0.5 - t);
                        if (f < 0 || f >= (ptrdiff_t)nang * npix) continue;
                        if (oper == 0) sino[f] += 0.25 * *pix;
                        else *pix += 0.25 * sino[f];
                    } else {
                        const double t1 = nh - t, lf = floor(t1), uf = ceil(t1);
                        const int l = (int)round(lf - 0.5), u = (int)round(uf - 0.5);
                        if (l > 0 && l < npix) {
                            if (oper == 0) row[l] += 0.25 * *pix * fabs(uf - t1);
                            else *pix += 0.25 * row[l] * fabs(uf - t1);
                        }
                        if (u > 0 && u < npix) {
                            if (oper == 0) row[u] += 0.25 * *pix * fabs(t1 - lf);
                            else *pix += 0.25 * row[u] * fabs(t1 - lf);
                        }
                    }
                }
            }
    }
    return 0;
}
