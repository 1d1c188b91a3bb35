/*
 * CPU restatement of the reference's SPS penalty terms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP penalty kernels are
 * compared against; nothing in the product library links it or calls it.
 *
 * Restated from (paths relative to the reference repository):
 *   common/operators/regularization_module/huber.c:50-118, tikhonov.c:150-195,
 *   haar.c:228-275, penalty_weights.c:7-15, penalty.pyx:16-106
 *
 * For every pixel (i, j) of an nr x nc image I and its 3x3 neighbours (k1, k2) inside the
 * image, in the order k1 = -1..1, k2 = -1..1, with d = I[i][j] - I[i+k1][j+k2] and the
 * neighbour weight w = 1/sqrt(2) (diagonal), 1 (edge), 4 (centre, where d = 0):
 *   huber num:    R += w * d / (1 + |d / delta|)
 *   huber den:    R += 2w / (1 + |d / delta|)
 *   tikhonov num: R += w * 2d          (den: the constant 2, penalty.pyx:73)
 *   haar num:     R += w * |d|         (den: 0, penalty.pyx:104)
 * Output arrays start at zero (the wrapper's np.zeros).
 *
 * Two float32 forms, selected by `arith`:
 *   0 (IEEE): the C as written (float d / delta, the huber potential's 1 + |.| and quotient
 *     in double, rounded to float on return; R += w * pot as float multiply and add);
 *   1 (FAST): the module as its -O3 -ffast-math build computes it (create_penalty_module.py
 *     flags), read from the assembly of that build and confirmed entry by entry against it
 *     (the quotient d / delta stays a division in every code path): every update is a
 *     float FMA, R = fma(w, pot, R) (huber den: fma(w + w, pot, R); tikhonov:
 *     fma(w + w, d, R); haar: fma(|d|, w, R)).
 */
#include <math.h>
#include <stddef.h>

enum { PEN_HUBER = 0, PEN_TIKHONOV = 1, PEN_HAAR = 2 };

static float weight(int k)
{
    if (k == 4) return 4.0f;
    if (k == 0 || k == 2 || k == 6 || k == 8) return (float)(1.0 / sqrt(2.0));
    return 1.0f;
}

int oracle_penalty(const float *img, int nr, int nc, int kind, float delta, int arith,
                   float *r_num, float *r_den)
{
    const int fast = arith == 1;
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < nc; ++j) {
            float rn = 0.0f, rd = 0.0f;
            const float c = img[(size_t)i * nc + j];
            for (int k = 0, k1 = -1; k1 <= 1; ++k1)
                for (int k2 = -1; k2 <= 1; ++k2, ++k) {
                    if (i + k1 < 0 || i + k1 > nr - 1 || j + k2 < 0 || j + k2 > nc - 1) continue;
                    const float nb = img[(size_t)(i + k1) * nc + j + k2];
                    const float d = c - nb;
                    const float w = weight(k);
                    if (kind == PEN_HUBER) {
                        const float q = fabsf(d / delta);
                        const double s = (double)q + 1.0;
                        const float pn = (float)((double)d / s);
                        const float pd = (float)(1.0 / s);
                        if (fast) {
                            rn = fmaf(w, pn, rn);
                            rd = fmaf(w + w, pd, rd);
                        } else {
                            rn = rn + w * pn;
                            rd = rd + (2.0f * w) * pd;
                        }
                    } else if (kind == PEN_TIKHONOV) {
                        rn = fast ? fmaf(w + w, d, rn) : rn + w * (2.0f * d);
                    } else {
                        const float a = d >= 0.0f ? d : nb - c;
                        rn = fast ? fmaf(a, w, rn) : rn + w * a;
                    }
                }
            r_num[(size_t)i * nc + j] = rn;
            r_den[(size_t)i * nc + j] = kind == PEN_HUBER ? rd : (kind == PEN_TIKHONOV ? 2.0f : 0.0f);
        }
    return 0;
}
