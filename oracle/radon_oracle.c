/*
 * CPU restatement of the reference's parallel-beam radon projectors.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP projectors are
 * compared against; nothing in the product library links it or calls it.
 *
 * Restated from (paths relative to the reference repository):
 *   common/operators/radon_module/pymodule_radon_distance_driven/radon_dd.c:54-285
 *   common/operators/radon_module/pymodule_radon_ray_driven/radon_rd.c:15-189
 *
 * Every function accumulates into its output array, exactly like the reference C
 * (the Cython wrappers hand it freshly zeroed arrays: radon_distance_driven.pyx:29,56,
 * radon_ray_driven.pyx:28,54).
 *
 * Two float32 arithmetic forms are restated, selected by `arith`:
 *   ORACLE_IEEE (0): the C as written, evaluated with IEEE single precision and no
 *                    contraction (what `gcc -O2 -ffp-contract=off` produces).
 *   ORACLE_FAST (1): the C as gcc 11 -O3 -ffast-math rewrites it (the reference's own
 *                    flags, radon_module/<name>/create_module.py:14): sincosf for the
 *                    trigonometry, a/b -> a*(1/b) with the reciprocal hoisted,
 *                    a/c + b/c -> (a+b)*(1/c), and fused multiply-add accumulation.
 * The *_f64 functions restate the same algorithms in double precision; they are the
 * "FP64 oracle" that float32 results are measured against.
 *
 * The distance-driven sweep is restated as a merge of the two boundary sequences
 * (pixel boundaries and detector boundaries, each already sorted) instead of the
 * reference's quicksort of the concatenation; segments, their (pixel, bin) labels and
 * the dropped final segment are the same.  The two orders differ only in how exactly
 * equal keys are ordered, which changes nothing but zero-length segments.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stddef.h>
#include <stdlib.h>

#define ORACLE_IEEE 0
#define ORACLE_FAST 1

static const double kPi = 3.141592653589793;   /* radon_dd.c:13, radon_rd.c:10 */
static const double kEps = 1.0e-7;              /* radon_dd.c:14, radon_rd.c:11 */

static void trig_f32(float theta, int arith, float *s, float *c)
{
    if (arith == ORACLE_FAST) {
        sincosf(theta, s, c);
    } else {
        *s = (float)sin((double)theta);
        *c = (float)cos((double)theta);
    }
}

/* ------------------------------------------------------------------------------ */
/* Distance-driven                                                                 */
/* ------------------------------------------------------------------------------ */

/*
 * One sweep line (radon_dd.c:125-189).  P[0..n] are the pixel boundaries along the
 * line, D[0..n] the detector-bin boundaries; both increasing.  Pixel q spans
 * [P[q], P[q+1]] and lives at img[q*img_stride]; bin d spans [D[d], D[d+1]] and lives
 * at row[d].  Every segment between consecutive keys that lies inside one pixel and
 * one bin transfers weight (key - previous key).  The sweep stops at the first key
 * whose index is n, before that segment is used (radon_dd.c:176-177).
 */
static void dd_sweep_f32(const float *P, const float *D, int n, float *img,
                         ptrdiff_t img_stride, float *row, int oper, int use_fma)
{
    int jp = 0, jd = 0;          /* next unread pixel / detector boundary */
    int cp = -1, cd = -1;        /* pixel / bin currently open (-1: none yet) */
    float prev = 0.0f;
    int first = 1;
    for (;;) {
        float key;
        int idx, is_det;
        if (P[jp] <= D[jd]) { key = P[jp]; idx = jp; is_det = 0; }
        else                { key = D[jd]; idx = jd; is_det = 1; }
        if (!first) {
            if (idx == n)
                break;
            if (cp >= 0 && cd >= 0) {
                float w = key - prev;
                float *pix = img + (ptrdiff_t)cp * img_stride;
                if (oper == 0)
                    row[cd] = use_fma ? fmaf(w, *pix, row[cd]) : row[cd] + w * *pix;
                else
                    *pix = use_fma ? fmaf(w, row[cd], *pix) : *pix + w * row[cd];
            }
        }
        if (is_det) { cd = idx; ++jd; } else { cp = idx; ++jp; }
        prev = key;
        first = 0;
    }
}

static void dd_sweep_f64(const double *P, const double *D, int n, double *img,
                         ptrdiff_t img_stride, double *row, int oper)
{
    int jp = 0, jd = 0, cp = -1, cd = -1, first = 1;
    double prev = 0.0;
    for (;;) {
        double key;
        int idx, is_det;
        if (P[jp] <= D[jd]) { key = P[jp]; idx = jp; is_det = 0; }
        else                { key = D[jd]; idx = jd; is_det = 1; }
        if (!first) {
            if (idx == n)
                break;
            if (cp >= 0 && cd >= 0) {
                double w = key - prev;
                double *pix = img + (ptrdiff_t)cp * img_stride;
                if (oper == 0) row[cd] += w * *pix;
                else           *pix += w * row[cd];
            }
        }
        if (is_det) { cd = idx; ++jd; } else { cp = idx; ++jp; }
        prev = key;
        first = 0;
    }
}

/*
 * radon_dd.c:54-285.  `angles` are the radians the C receives, i.e. already
 * fftshift-ed by the wrapper (radon_distance_driven.pyx:27,54).  View v is written to
 * sinogram row nang-1-v (radon_dd.c:79).  Returns 0, or -1 on allocation failure.
 */
int oracle_radon_dd(float *image, int npix, const float *angles, int nang, int oper,
                    float *sino, int arith)
{
    const int nh = (int)(npix * 0.5);
    const int fast = (arith == ORACLE_FAST);
    float *P = (float *)malloc(sizeof(float) * (size_t)(npix + 1));
    float *D = (float *)malloc(sizeof(float) * (size_t)(npix + 1));
    if (!P || !D) { free(P); free(D); return -1; }

    for (int v = 0; v < nang; ++v) {
        const float theta = angles[v];
        float s, c;
        trig_f32(theta, arith, &s, &c);
        float *row = sino + (size_t)(nang - 1 - v) * (size_t)npix;

        if (fabs((double)s) < kEps) {                        /* radon_dd.c:75-84 */
            for (int i = 0; i < npix; ++i)
                for (int j = 0; j < npix; ++j) {
                    if (oper == 0) row[i] += image[(size_t)i * npix + j];
                    else           image[(size_t)i * npix + j] += row[i];
                }
        } else if (fabs((double)c) < kEps) {                 /* radon_dd.c:87-96 */
            for (int i = 0; i < npix; ++i)
                for (int j = 0; j < npix; ++j) {
                    if (oper == 0) row[i] += image[(size_t)j * npix + i];
                    else           image[(size_t)j * npix + i] += row[i];
                }
        } else if ((double)theta > kPi / 2) {                /* radon_dd.c:99-191 */
            /* sweep along image row i; pixel keys x - y*c/s, bin keys x/s */
            const float rs = 1.0f / s, t = c * rs;
            for (int i = 0; i < npix; ++i) {
                const float y = (float)((double)(i - nh) + 0.5);
                const float K = fast ? y * t : (y * c) / s;
                for (int j = 0; j <= npix; ++j) {
                    const float x = (float)(j - nh);
                    P[j] = x - K;
                    D[j] = fast ? x * rs : x / s;
                }
                dd_sweep_f32(P, D, npix, image + (size_t)i * npix, 1, row, oper, fast);
            }
        } else if ((double)theta < kPi / 2.0 && theta != 0.0f) {  /* radon_dd.c:194-283 */
            /* sweep along image column i; pixel keys y - x*s/c, bin keys y/c */
            const float rc = 1.0f / c, t = s * rc;
            for (int i = 0; i < npix; ++i) {
                const float x = (float)((double)(i - nh) + 0.5);
                const float K = fast ? x * t : (x * s) / c;
                for (int j = 0; j <= npix; ++j) {
                    const float y = (float)(j - nh);
                    P[j] = y - K;
                    D[j] = fast ? y * rc : y / c;
                }
                dd_sweep_f32(P, D, npix, image + i, npix, row, oper, fast);
            }
        }
    }
    free(P);
    free(D);
    return 0;
}

/* Same algorithm in double precision, angles taken as given (no float rounding). */
int oracle_radon_dd_f64(double *image, int npix, const double *angles, int nang,
                        int oper, double *sino)
{
    const int nh = (int)(npix * 0.5);
    double *P = (double *)malloc(sizeof(double) * (size_t)(npix + 1));
    double *D = (double *)malloc(sizeof(double) * (size_t)(npix + 1));
    if (!P || !D) { free(P); free(D); return -1; }

    for (int v = 0; v < nang; ++v) {
        const double theta = angles[v];
        const double s = sin(theta), c = cos(theta);
        double *row = sino + (size_t)(nang - 1 - v) * (size_t)npix;
        if (fabs(s) < kEps) {
            for (int i = 0; i < npix; ++i)
                for (int j = 0; j < npix; ++j) {
                    if (oper == 0) row[i] += image[(size_t)i * npix + j];
                    else           image[(size_t)i * npix + j] += row[i];
                }
        } else if (fabs(c) < kEps) {
            for (int i = 0; i < npix; ++i)
                for (int j = 0; j < npix; ++j) {
                    if (oper == 0) row[i] += image[(size_t)j * npix + i];
                    else           image[(size_t)j * npix + i] += row[i];
                }
        } else if (theta > kPi / 2) {
            for (int i = 0; i < npix; ++i) {
                const double y = (double)(i - nh) + 0.5;
                for (int j = 0; j <= npix; ++j) {
                    const double x = (double)(j - nh);
                    P[j] = x - y * c / s;
                    D[j] = x / s;
                }
                dd_sweep_f64(P, D, npix, image + (size_t)i * npix, 1, row, oper);
            }
        } else if (theta < kPi / 2.0 && theta != 0.0) {
            for (int i = 0; i < npix; ++i) {
                const double x = (double)(i - nh) + 0.5;
                for (int j = 0; j <= npix; ++j) {
                    const double y = (double)(j - nh);
                    P[j] = y - x * s / c;
                    D[j] = y / c;
                }
                dd_sweep_f64(P, D, npix, image + i, npix, row, oper);
            }
        }
    }
    free(P);
    free(D);
    return 0;
}

/* ------------------------------------------------------------------------------ */
/* Ray-driven                                                                      */
/* ------------------------------------------------------------------------------ */

/*
 * Siddon-style traversal of one sub-ray from p1 to p2 (radon_rd.c:15-71).  The segment
 * walked since the previous step is credited to the cell just entered; the walk takes
 * one last step past alpha = 1.  Cells with row or column index 0 are skipped
 * (radon_rd.c:62).  `alpha` is uninitialised in the reference (radon_rd.c:17,40); gcc
 * materialises it as 0, which is restated here.
 */
static float round_f32(float v, int fast)
{
    if (fast)   /* gcc -ffast-math inline round: trunc(v + copysign(0.49999997f, v)) */
        return truncf(v + copysignf(0.49999997f, v));
    return (float)round((double)v);
}

static void siddon_f32(float *sino_bin, float *image, int npix, const float p1[2],
                       const float p2[2], int oper, int fast)
{
    const int nh = (int)(npix * 0.5);
    const float d0 = p2[0] - p1[0], d1 = p2[1] - p1[1];
    const float incr = fast ? sqrtf(fmaf(d0, d0, d1 * d1))
                            : (float)sqrt((double)(d0 * d0 + d1 * d1));
    int ii = (int)round_f32(p1[0], fast);
    int jj = (int)round_f32(p1[1], fast);
    const float dx = (p2[0] > p1[0]) ? 1.0f : -1.0f;
    const float dy = (p2[1] > p1[1]) ? 1.0f : -1.0f;
    float alpha = 0.0f, alpha_old = 0.0f;

    while (alpha <= 1) {
        const float ay = (float)(((double)jj + 0.5 * (double)dy - (double)p1[1]) / (double)d1);
        const float ax = (float)(((double)ii + 0.5 * (double)dx - (double)p1[0]) / (double)d0);
        if (ax < ay) {
            alpha = ax;
            ii = (int)((float)ii + dx);
        } else if (ax > ay) {
            alpha = ay;
            jj = (int)((float)jj + dy);
        } else if (ax == ay) {
            alpha = ay;
            ii = (int)((float)ii + dx);
            jj = (int)((float)jj + dy);
        }
        const float l = (alpha - alpha_old) * incr;
        const int i = ii + nh, j = jj + nh;
        if (i > 0 && i < npix && j > 0 && j < npix) {
            float *pix = image + (size_t)i * npix + j;
            if (oper == 0) {
                *sino_bin = fast ? (float)fma(0.2 * (double)l, (double)*pix, (double)*sino_bin)
                                 : (float)((double)*sino_bin + 0.2 * (double)l * (double)*pix);
            } else {
                *pix = fast ? (float)fma(0.2 * (double)*sino_bin, (double)l, (double)*pix)
                            : (float)((double)*pix + 0.2 * (double)l * (double)*sino_bin);
            }
        }
        alpha_old = alpha;
    }
}

/* Entry/exit points of sub-ray i on the square [-nh, nh]^2 (radon_rd.c:123-179). */
static void rd_endpoints_f32(int i, int npix, float theta, float s, float c, int fast,
                             float p1[2], float p2[2])
{
    const int nh = (int)(npix * 0.5);
    float t;
    if (fast)
        t = (float)fabs(fma((double)i, 1.0 / 6.0, -(double)nh));
    else
        t = (float)fabs(i / 6.0 - nh);
    if (i >= 3 * npix && (double)theta < kPi / 2) t = -t;
    if (i < 3 * npix && (double)theta > kPi / 2) t = -t;

    float x1, y1, x2, y2;
    const float nhf = (float)nh, mnhf = (float)(-nh);
    if ((double)theta < kPi / 2) {
        if (fast) {
            const float rc = 1.0f / c, rs = 1.0f / s;
            y1 = (t + mnhf * s) * rc;
            x1 = (mnhf * c - t) * rs;
            y2 = (t + nhf * s) * rc;
            x2 = (nhf * c - t) * rs;
        } else {
            y1 = -nh * s / c + t / c;
            x1 = -nh * c / s - t / s;
            y2 = nh * s / c + t / c;
            x2 = nh * c / s - t / s;
        }
        if (y1 >= -nh && y1 <= nh) { p1[0] = mnhf; p1[1] = y1; }
        else                       { p1[0] = x1;   p1[1] = mnhf; }
        if (y2 >= -nh && y2 <= nh) { p2[0] = nhf;  p2[1] = y2; }
        else                       { p2[0] = x2;   p2[1] = nhf; }
    } else {
        if (fast) {
            const float rc = 1.0f / c, rs = 1.0f / s;
            y1 = (nhf * s - t) * rc;
            x1 = (t + mnhf * c) * rs;
            y2 = (mnhf * s - t) * rc;
            x2 = (t + nhf * c) * rs;
        } else {
            y1 = nh * s / c - t / c;
            x1 = -nh * c / s + t / s;
            y2 = -nh * s / c - t / c;
            x2 = nh * c / s + t / s;
        }
        if (y1 >= -nh && y1 <= nh) { p1[0] = nhf;  p1[1] = y1; }
        else                       { p1[0] = x1;   p1[1] = mnhf; }
        if (y2 >= -nh && y2 <= nh) { p2[0] = mnhf; p2[1] = y2; }
        else                       { p2[0] = x2;   p2[1] = nhf; }
    }
}

/* radon_rd.c:76-189.  No angle permutation: view v is sinogram row v. */
int oracle_radon_rd(float *image, int npix, const float *angles, int nang, int oper,
                    float *sino, int arith)
{
    const int fast = (arith == ORACLE_FAST);
    for (int v = 0; v < nang; ++v) {
        const float theta = angles[v];
        float s, c;
        trig_f32(theta, arith, &s, &c);
        float *row = sino + (size_t)v * npix;
        if (fabs((double)s) < kEps) {                        /* radon_rd.c:96-105 */
            for (int i = 1; i < npix; ++i)
                for (int j = 0; j < npix; ++j) {
                    if (oper == 0) row[npix - 1 - i] += image[(size_t)j * npix + i];
                    else           image[(size_t)j * npix + i] += row[npix - 1 - i];
                }
        } else if (fabs((double)c) < kEps) {                 /* radon_rd.c:108-117 */
            for (int i = 0; i < npix; ++i)
                for (int j = 0; j < npix; ++j) {
                    if (oper == 0) row[i] += image[(size_t)i * npix + j];
                    else           image[(size_t)i * npix + j] += row[i];
                }
        } else {                                             /* radon_rd.c:120-187 */
            for (int i = 0; i < 6 * npix; ++i) {
                if (i % 6 == 0)
                    continue;
                float p1[2], p2[2];
                rd_endpoints_f32(i, npix, theta, s, c, fast, p1, p2);
                siddon_f32(row + i / 6, image, npix, p1, p2, oper, fast);
            }
        }
    }
    return 0;
}

/* Double-precision restatement (same traversal and quirks, no float rounding). */
static void siddon_f64(double *sino_bin, double *image, int npix, const double p1[2],
                       const double p2[2], int oper)
{
    const int nh = (int)(npix * 0.5);
    const double d0 = p2[0] - p1[0], d1 = p2[1] - p1[1];
    const double incr = sqrt(d0 * d0 + d1 * d1);
    int ii = (int)round(p1[0]), jj = (int)round(p1[1]);
    const int dx = (p2[0] > p1[0]) ? 1 : -1;
    const int dy = (p2[1] > p1[1]) ? 1 : -1;
    double alpha = 0.0, alpha_old = 0.0;
    while (alpha <= 1) {
        const double ay = (jj + 0.5 * dy - p1[1]) / d1;
        const double ax = (ii + 0.5 * dx - p1[0]) / d0;
        if (ax < ay)      { alpha = ax; ii += dx; }
        else if (ax > ay) { alpha = ay; jj += dy; }
        else if (ax == ay){ alpha = ay; ii += dx; jj += dy; }
        const double l = (alpha - alpha_old) * incr;
        const int i = ii + nh, j = jj + nh;
        if (i > 0 && i < npix && j > 0 && j < npix) {
            double *pix = image + (size_t)i * npix + j;
            if (oper == 0) *sino_bin += 0.2 * l * *pix;
            else           *pix += 0.2 * l * *sino_bin;
        }
        alpha_old = alpha;
    }
}

int oracle_radon_rd_f64(double *image, int npix, const double *angles, int nang,
                        int oper, double *sino)
{
    const int nh = (int)(npix * 0.5);
    for (int v = 0; v < nang; ++v) {
        const double theta = angles[v];
        const double s = sin(theta), c = cos(theta);
        double *row = sino + (size_t)v * npix;
        if (fabs(s) < kEps) {
            for (int i = 1; i < npix; ++i)
                for (int j = 0; j < npix; ++j) {
                    if (oper == 0) row[npix - 1 - i] += image[(size_t)j * npix + i];
                    else           image[(size_t)j * npix + i] += row[npix - 1 - i];
                }
        } else if (fabs(c) < kEps) {
            for (int i = 0; i < npix; ++i)
                for (int j = 0; j < npix; ++j) {
                    if (oper == 0) row[i] += image[(size_t)i * npix + j];
                    else           image[(size_t)i * npix + j] += row[i];
                }
        } else {
            for (int i = 0; i < 6 * npix; ++i) {
                if (i % 6 == 0)
                    continue;
                double t = fabs(i / 6.0 - nh);
                if (i >= 3 * npix && theta < kPi / 2) t = -t;
                if (i < 3 * npix && theta > kPi / 2) t = -t;
                double p1[2], p2[2], x1, y1, x2, y2;
                if (theta < kPi / 2) {
                    y1 = -nh * s / c + t / c;  x1 = -nh * c / s - t / s;
                    y2 = nh * s / c + t / c;   x2 = nh * c / s - t / s;
                    if (y1 >= -nh && y1 <= nh) { p1[0] = -nh; p1[1] = y1; } else { p1[0] = x1; p1[1] = -nh; }
                    if (y2 >= -nh && y2 <= nh) { p2[0] = nh;  p2[1] = y2; } else { p2[0] = x2; p2[1] = nh; }
                } else {
                    y1 = nh * s / c - t / c;   x1 = -nh * c / s + t / s;
                    y2 = -nh * s / c - t / c;  x2 = nh * c / s + t / s;
                    if (y1 >= -nh && y1 <= nh) { p1[0] = nh;  p1[1] = y1; } else { p1[0] = x1; p1[1] = -nh; }
                    if (y2 >= -nh && y2 <= nh) { p2[0] = -nh; p2[1] = y2; } else { p2[0] = x2; p2[1] = nh; }
                }
                siddon_f64(row + i / 6, image, npix, p1, p2, oper);
            }
        }
    }
    return 0;
}
