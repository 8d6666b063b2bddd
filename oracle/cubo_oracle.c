/*
 * cubo_oracle.c — test infrastructure only: the CPU restatement of CuboKit's .cpu hot path that every
 * GPU result is checked against, and the CPU column of the benchmark. It is never linked into the
 * product library (libcubo_hip.so) and never called on a product path.
 *
 * Each function names the reference file:line it restates (paths relative to the CuboKit repository).
 * The code is written from the definitions, in C, with Double arithmetic.
 *
 * Pinning (which check fixes which function) is listed in tests/test_oracle_pins.py and
 * tests/test_oracle_math.py; the functions without a reference golden vector (mask stack, bilinear shift,
 * adjoints, DPC step) are pinned by brute force, closed forms, finite differences and dot-product tests.
 */
#include "cubo_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------------------------------- */
/* Detector masks: Sources/CuboVirtual/DetectorGeometry.swift:43-98 (py4DSTEM make_detector).      */
/* ---------------------------------------------------------------------------------------------- */

int oc_mask_point(int qy, int qx, int row, int col, float *out) {
    if (qy <= 0 || qx <= 0) return OC_INVALID_SHAPE;
    memset(out, 0, sizeof(float) * (size_t)qy * (size_t)qx);
    if (row < 0 || row >= qy || col < 0 || col >= qx) return OC_POINT_OUTSIDE;   /* :48-52 */
    out[(size_t)row * qx + col] = 1.0f;
    return OC_OK;
}

/* d^2 = (i - cr)^2 + (j - cc)^2 in Double, row term first (DetectorGeometry.swift:88-98). */
static void fill_pred(float *out, int qy, int qx, double cr, double cc, double lo2, double hi2, int annulus) {
    for (int i = 0; i < qy; ++i) {
        double di = (double)i - cr;
        double di2 = di * di;
        for (int j = 0; j < qx; ++j) {
            double dj = (double)j - cc;
            double d2 = di2 + dj * dj;
            int sel = annulus ? (d2 > lo2 && d2 < hi2) : (d2 < hi2);
            out[(size_t)i * qx + j] = sel ? 1.0f : 0.0f;
        }
    }
}

int oc_mask_circle(int qy, int qx, double cr, double cc, double r, float *out) {
    if (qy <= 0 || qx <= 0) return OC_INVALID_SHAPE;
    if (!isfinite(cr) || !isfinite(cc) || !isfinite(r)) return OC_NON_FINITE;     /* :54 */
    if (r < 0) return OC_NEGATIVE_RADIUS;                                          /* :55 */
    fill_pred(out, qy, qx, cr, cc, 0.0, r * r, 0);
    return OC_OK;
}

int oc_mask_annulus(int qy, int qx, double cr, double cc, double inner, double outer, float *out) {
    if (qy <= 0 || qx <= 0) return OC_INVALID_SHAPE;
    if (!isfinite(cr) || !isfinite(cc) || !isfinite(inner) || !isfinite(outer)) return OC_NON_FINITE; /* :59 */
    if (inner < 0) return OC_NEGATIVE_RADIUS;                                                          /* :62 */
    if (!(outer > inner)) return OC_INNER_NOT_BELOW_OUTER;                                             /* :63 */
    fill_pred(out, qy, qx, cr, cc, inner * inner, outer * outer, 1);
    return OC_OK;
}

int oc_mask_rectangle(int qy, int qx, double r0, double r1, double c0, double c1, float *out) {
    if (qy <= 0 || qx <= 0) return OC_INVALID_SHAPE;
    double b[4] = {r0, r1, c0, c1};
    for (int k = 0; k < 4; ++k) {
        if (!isfinite(b[k])) return OC_NON_FINITE;                                  /* :68 */
        b[k] = nearbyint(b[k]);                      /* round half to even in the default rounding mode */
        if (b[k] < 0) return OC_NEGATIVE_RECT_BOUND;                                /* :70 */
    }
    memset(out, 0, sizeof(float) * (size_t)qy * (size_t)qx);
    long a0 = (long)fmin(b[0], (double)qy), a1 = (long)fmin(b[1], (double)qy);      /* :72-73 clip */
    long e0 = (long)fmin(b[2], (double)qx), e1 = (long)fmin(b[3], (double)qx);
    if (a0 < a1 && e0 < e1)
        for (long i = a0; i < a1; ++i)
            for (long j = e0; j < e1; ++j) out[i * qx + j] = 1.0f;
    return OC_OK;
}

/* ---------------------------------------------------------------------------------------------- */
/* Virtual image: Sources/CuboVirtual/CPUReducers.swift:13-39 — widen to Double, dot with the Double  */
/* mask over every pixel, per position.                                                             */
/* ---------------------------------------------------------------------------------------------- */

void oc_virtual_image(const float *cube, int64_t npos, int64_t ndet, const float *mask, double *out) {
    for (int64_t p = 0; p < npos; ++p) {
        const float *x = cube + p * ndet;
        double s = 0.0;
        for (int64_t i = 0; i < ndet; ++i) s += (double)x[i] * (double)mask[i];
        out[p] = s;
    }
}

/* ---------------------------------------------------------------------------------------------- */
/* Mean and max: CPUReducers.swift:46-97 and CompensatedSum.swift:134-161 — per pixel TwoSum in     */
/* scan order seeded at -0, value() keeps a non-finite sum and a -0 sum; max in Float32 seeded at   */
/* -inf, NaN wins (vDSP_vmax, NaNMaxAndPolarThresholdTests.swift:12-26); among equal values the     */
/* later one is kept (Swift.max(a, b) = b >= a ? b : a, MetalReducers.swift:482).                     */
/* ---------------------------------------------------------------------------------------------- */

static inline float max_step(float m, float x) {
    if (isnan(m) || isnan(x)) return NAN;
    return x >= m ? x : m;
}

static inline double comp_value(double sum, double corr) {
    if (!isfinite(sum)) return sum;
    if (sum == 0.0 && signbit(sum) && corr == 0.0) return sum;
    return sum + corr;
}

static void mean_max_block(const float *cube, int64_t npos, int64_t ndet, int64_t lo, int64_t hi,
                           double *sum, double *corr, float *mx) {
    for (int64_t p = 0; p < npos; ++p) {
        const float *x = cube + p * ndet;
        for (int64_t i = lo; i < hi; ++i) {
            double v = (double)x[i];
            double a = sum[i];
            double t = a + v;
            double bb = t - a;
            corr[i] += (a - (t - bb)) + (v - bb);
            sum[i] = t;
            mx[i] = max_step(mx[i], x[i]);
        }
    }
}

void oc_mean_max(const float *cube, int64_t npos, int64_t ndet, double *mean, double *max) {
    double *sum = (double *)malloc(sizeof(double) * ndet);
    double *corr = (double *)malloc(sizeof(double) * ndet);
    float *mx = (float *)malloc(sizeof(float) * ndet);
    for (int64_t i = 0; i < ndet; ++i) { sum[i] = -0.0; corr[i] = -0.0; mx[i] = -INFINITY; }
    mean_max_block(cube, npos, ndet, 0, ndet, sum, corr, mx);
    for (int64_t i = 0; i < ndet; ++i) {
        mean[i] = comp_value(sum[i], corr[i]) / (double)npos;
        max[i] = (double)mx[i];
    }
    free(sum); free(corr); free(mx);
}

/* ---------------------------------------------------------------------------------------------- */
/* Centre of mass: Sources/CuboVirtual/CenterOfMass.swift:51-91 — w, w*i, w*j formed in Double,      */
/* Neumaier sums in pixel order, a non-finite step carries no correction.                             */
/* ---------------------------------------------------------------------------------------------- */

typedef struct { double sum, corr; } neumaier;
static inline void nm_add(neumaier *n, double x) {
    double t = n->sum + x;
    if (isfinite(t)) n->corr += fabs(n->sum) >= fabs(x) ? (n->sum - t) + x : (x - t) + n->sum;
    n->sum = t;
}
static inline double nm_value(const neumaier *n) { return n->sum + n->corr; }

void oc_com_sums(const float *cube, int64_t npos, int qy, int qx, const float *weights,
                 double *s, double *mr, double *mc) {
    int64_t ndet = (int64_t)qy * qx;
    double *w = (double *)malloc(sizeof(double) * ndet);
    double *wi = (double *)malloc(sizeof(double) * ndet);
    double *wj = (double *)malloc(sizeof(double) * ndet);
    for (int64_t k = 0; k < ndet; ++k) {
        w[k] = weights ? (double)weights[k] : 1.0;
        wi[k] = w[k] * (double)(k / qx);
        wj[k] = w[k] * (double)(k % qx);
    }
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npos; ++p) {
        const float *x = cube + p * ndet;
        neumaier a = {0, 0}, b = {0, 0}, d = {0, 0};
        for (int64_t k = 0; k < ndet; ++k) {
            double v = (double)x[k];
            nm_add(&a, v * w[k]);
            nm_add(&b, v * wi[k]);
            nm_add(&d, v * wj[k]);
        }
        s[p] = nm_value(&a); mr[p] = nm_value(&b); mc[p] = nm_value(&d);
    }
    free(w); free(wi); free(wj);
}

/* ---------------------------------------------------------------------------------------------- */
/* Mask stack (no reference function: K virtual images plus the CoM weights are one contraction).    */
/* Y[p,k] = sum_i double(x[p,i]) * W[k,i] with a Neumaier sum, so the FP64 error stays ~1e-16.        */
/* ---------------------------------------------------------------------------------------------- */

void oc_mask_stack(const float *cube, int64_t npos, int64_t ndet, const double *w, int k, double *y) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npos; ++p) {
        const float *x = cube + p * ndet;
        for (int q = 0; q < k; ++q) {
            neumaier a = {0, 0};
            const double *wq = w + (int64_t)q * ndet;
            for (int64_t i = 0; i < ndet; ++i) nm_add(&a, (double)x[i] * wq[i]);
            y[p * k + q] = nm_value(&a);
        }
    }
}

/* dW[k,i] = sum_p dY[p,k] * x[p,i]: the adjoint of oc_mask_stack with respect to W. */
void oc_mask_stack_grad_w(const float *cube, int64_t npos, int64_t ndet, const double *dy, int k, double *dw) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ndet; ++i) {
        for (int q = 0; q < k; ++q) {
            neumaier a = {0, 0};
            for (int64_t p = 0; p < npos; ++p) nm_add(&a, dy[p * k + q] * (double)cube[p * ndet + i]);
            dw[(int64_t)q * ndet + i] = nm_value(&a);
        }
    }
}

/* ---------------------------------------------------------------------------------------------- */
/* Bilinear centre shift: Sources/CuboVirtual/ShiftedImaging.swift:43-73 and 109-122.                */
/* dr = origin_r - mean(origin_r), fr = floor(dr), wr = dr - fr; the mask is the blend of the four    */
/* periodic rolls by (fr, fc), (fr+1, fc), (fr, fc+1), (fr+1, fc+1);                                  */
/* roll: out[(i+sr)%qy, (j+sc)%qx] = a[i, j].                                                        */
/* ---------------------------------------------------------------------------------------------- */

static int wrap_shift(double v, int n) {                 /* ShiftedImaging.swift:109-113 */
    double m = fmod(v, (double)n);
    if (m < 0) m += (double)n;
    return ((int)m) % n;
}

static double mean_plain(const double *a, int64_t n) {   /* the origin map mean (OriginMap.meanRow) */
    return oc_pairwise_sum(a, n) / (double)n;
}

/* ---------------------------------------------------------------------------------------------- */
/* Selected-area and centre-shifted patterns.                                                        */
/*  - pattern over .mask / .weights without a shift: SelectedDiffraction.swift:110-179 — per pixel a */
/*    TwoSum over the picked positions in scan order from +0 (CompensatedSum.add: Double(x) * w      */
/*    formed first), value() / count or / Neumaier(weights); the max of Double(x) * w seeded -inf,    */
/*    NaN wins, of equal values the first kept (v > m ? v : m);                                       */
/*  - pattern with PatternShift.rounded: SelectedDiffraction.swift:225-271 — each picked pattern      */
/*    rolled by wrap(round_half_even(meanRow - row[k])), ..., Neumaier per pixel, first kept;          */
/*  - meanAndMax(of:shift: .rounded): ShiftedImaging.swift:125-167 — every position, Neumaier per     */
/*    pixel / scanCount, the max with Swift.max (the later of equal values kept).                     */
/* ---------------------------------------------------------------------------------------------- */

static inline double max_first(double m, double v) {
    if (isnan(m) || isnan(v)) return NAN;
    return v > m ? v : m;
}
static inline double max_later(double m, double v) {
    if (isnan(m) || isnan(v)) return NAN;
    return v >= m ? v : m;
}

int oc_pattern(const float *cube, int64_t npos, int qy, int qx, int stat, const uint8_t *take, const double *w,
               const double *orow, const double *ocol, int ties_later, int fourier, double *out) {
    const int64_t ndet = (int64_t)qy * qx;
    if (take && w) return OC_INVALID_PARAMETER;
    if (w)
        for (int64_t k = 0; k < npos; ++k)
            if (!isfinite(w[k])) return OC_NON_FINITE;
    int64_t selected = 0;
    for (int64_t k = 0; k < npos; ++k) selected += take ? (take[k] != 0) : 1;
    if (selected == 0) return OC_INVALID_PARAMETER;
    const int shifted = orow != NULL;
    double mr = 0, mc = 0;
    if (shifted) {
        for (int64_t k = 0; k < npos; ++k)
            if (!isfinite(orow[k]) || !isfinite(ocol[k])) return OC_NON_FINITE;
        mr = mean_plain(orow, npos);
        mc = mean_plain(ocol, npos);
    }
    double *sum = (double *)malloc(sizeof(double) * ndet), *corr = (double *)malloc(sizeof(double) * ndet);
    double *top = (double *)malloc(sizeof(double) * ndet), *moved = (double *)malloc(sizeof(double) * ndet);
    double *im = (double *)malloc(sizeof(double) * ndet), *fr = (double *)malloc(sizeof(double) * qy),
           *fc = (double *)malloc(sizeof(double) * qx);
    /* NumPy's fftfreq (ShiftedImaging.swift:169-172) */
    for (int i = 0; i < qy; ++i) fr[i] = (double)(i < (qy - 1) / 2 + 1 ? i : i - qy) * (1.0 / (double)qy);
    for (int j = 0; j < qx; ++j) fc[j] = (double)(j < (qx - 1) / 2 + 1 ? j : j - qx) * (1.0 / (double)qx);
    for (int64_t i = 0; i < ndet; ++i) { sum[i] = 0.0; corr[i] = 0.0; top[i] = -INFINITY; }
    neumaier total = {0, 0};
    for (int64_t k = 0; k < npos; ++k) {
        if (take && !take[k]) continue;
        const double wk = w ? w[k] : 1.0;
        nm_add(&total, wk);
        const float *x = cube + k * ndet;
        if (shifted && fourier) {
            /* get_shifted_ar: Re(IFFT2(FFT2(p) exp(-2 pi i (sc f_c + sr f_r)))) (SelectedDiffraction.swift:273-285) */
            const double sr = mr - orow[k], sc = mc - ocol[k];
            for (int64_t i = 0; i < ndet; ++i) { moved[i] = (double)x[i]; im[i] = 0.0; }
            oc_fft2d(moved, im, qy, qx, 0);
            for (int i = 0; i < qy; ++i)
                for (int j = 0; j < qx; ++j) {
                    const double angle = -2 * M_PI * (sc * fc[j] + sr * fr[i]);
                    const double c = cos(angle), s = sin(angle), a = moved[(int64_t)i * qx + j], b = im[(int64_t)i * qx + j];
                    moved[(int64_t)i * qx + j] = a * c - b * s;
                    im[(int64_t)i * qx + j] = a * s + b * c;
                }
            oc_fft2d(moved, im, qy, qx, 1);
        } else if (shifted) {
            /* np.roll: out[(i + sr) % qy, (j + sc) % qx] = p[i, j] */
            const int sr = wrap_shift(nearbyint(mr - orow[k]), qy), sc = wrap_shift(nearbyint(mc - ocol[k]), qx);
            for (int i = 0; i < qy; ++i)
                for (int j = 0; j < qx; ++j) moved[(int64_t)((i + sr) % qy) * qx + (j + sc) % qx] = (double)x[(int64_t)i * qx + j];
        } else {
            for (int64_t i = 0; i < ndet; ++i) moved[i] = (double)x[i];
        }
        for (int64_t i = 0; i < ndet; ++i) {
            const double v = w ? moved[i] * wk : moved[i];
            if (stat == 0) {
                if (shifted) {               /* Neumaier (Neumaier.add, CenterOfMass.swift:86-90) */
                    neumaier n = {sum[i], corr[i]};
                    nm_add(&n, v);
                    sum[i] = n.sum; corr[i] = n.corr;
                } else {                     /* TwoSum (CompensatedSum.scalar, CompensatedSum.swift:146-152) */
                    const double a = sum[i], t = a + v, bb = t - a;
                    corr[i] += (a - (t - bb)) + (v - bb);
                    sum[i] = t;
                }
            } else {
                top[i] = ties_later ? max_later(top[i], v) : max_first(top[i], v);
            }
        }
    }
    int st = OC_OK;
    if (stat == 0) {
        const double norm = w ? nm_value(&total) : (double)selected;
        if (norm == 0) st = OC_ZERO_FIELD;
        for (int64_t i = 0; i < ndet && st == OC_OK; ++i) {
            const double v = shifted ? sum[i] + corr[i] : comp_value(sum[i], corr[i]);
            out[i] = v / norm;
        }
    } else {
        for (int64_t i = 0; i < ndet; ++i) out[i] = top[i];
    }
    free(sum); free(corr); free(top); free(moved); free(im); free(fr); free(fc);
    return st;
}


/* value of roll(mask, sr, sc) at (i, j) = mask[(i - sr) mod qy, (j - sc) mod qx] */
static inline double rolled(const double *m, int qy, int qx, int sr, int sc, int i, int j) {
    int a = i - sr; a %= qy; if (a < 0) a += qy;
    int b = j - sc; b %= qx; if (b < 0) b += qx;
    return m[(int64_t)a * qx + b];
}

typedef struct { int ir, ic; double wr, wc; } bil_shift;

static bil_shift bil_of(double dr, double dc, int qy, int qx) {
    bil_shift s;
    double fr = floor(dr), fc = floor(dc);
    s.wr = dr - fr; s.wc = dc - fc;
    s.ir = wrap_shift(fr, qy); s.ic = wrap_shift(fc, qx);
    return s;
}

int oc_vi_bilinear(const float *cube, int64_t npos, int qy, int qx, const double *mask,
                   const double *origin_r, const double *origin_c, double *out) {
    for (int64_t p = 0; p < npos; ++p)
        if (!isfinite(origin_r[p]) || !isfinite(origin_c[p])) return OC_NON_FINITE;
    int64_t ndet = (int64_t)qy * qx;
    double mr = mean_plain(origin_r, npos), mc = mean_plain(origin_c, npos);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npos; ++p) {
        bil_shift s = bil_of(origin_r[p] - mr, origin_c[p] - mc, qy, qx);
        double w00 = (1 - s.wr) * (1 - s.wc), w10 = s.wr * (1 - s.wc), w01 = (1 - s.wr) * s.wc, w11 = s.wr * s.wc;
        neumaier a = {0, 0};
        const float *x = cube + p * ndet;
        for (int i = 0; i < qy; ++i)
            for (int j = 0; j < qx; ++j) {
                double m = rolled(mask, qy, qx, s.ir, s.ic, i, j) * w00 +
                           rolled(mask, qy, qx, s.ir + 1, s.ic, i, j) * w10 +
                           rolled(mask, qy, qx, s.ir, s.ic + 1, i, j) * w01 +
                           rolled(mask, qy, qx, s.ir + 1, s.ic + 1, i, j) * w11;
                nm_add(&a, (double)x[(int64_t)i * qx + j] * m);
            }
        out[p] = nm_value(&a);
    }
    return OC_OK;
}

/* MaskShift.rounded (ShiftedImaging.swift:65-71): the mask rolled by wrap(round_half_even(origin - mean)) per
 * position, Neumaier over Double(x) * m in pixel order. */
int oc_vi_rounded(const float *cube, int64_t npos, int qy, int qx, const double *mask, const double *origin_r,
                  const double *origin_c, double *out) {
    for (int64_t p = 0; p < npos; ++p)
        if (!isfinite(origin_r[p]) || !isfinite(origin_c[p])) return OC_NON_FINITE;
    int64_t ndet = (int64_t)qy * qx;
    double mr = mean_plain(origin_r, npos), mc = mean_plain(origin_c, npos);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npos; ++p) {
        const int sr = wrap_shift(nearbyint(origin_r[p] - mr), qy), sc = wrap_shift(nearbyint(origin_c[p] - mc), qx);
        neumaier a = {0, 0};
        const float *x = cube + p * ndet;
        for (int i = 0; i < qy; ++i)
            for (int j = 0; j < qx; ++j)
                nm_add(&a, (double)x[(int64_t)i * qx + j] * rolled(mask, qy, qx, sr, sc, i, j));
        out[p] = nm_value(&a);
    }
    return OC_OK;
}

/* dW[u] = sum_p dY[p] * sum_corners w_corner * x[p, u + shift_corner]: the adjoint with respect to the
 * unshifted mask (x at pixel (i, j) meets mask[(i - sr) mod qy, (j - sc) mod qx]). */
void oc_vi_bilinear_grad_w(const float *cube, int64_t npos, int qy, int qx, const double *origin_r,
                           const double *origin_c, const double *dy, double *dw) {
    int64_t ndet = (int64_t)qy * qx;
    double mr = mean_plain(origin_r, npos), mc = mean_plain(origin_c, npos);
    for (int64_t u = 0; u < ndet; ++u) dw[u] = 0.0;
    for (int64_t p = 0; p < npos; ++p) {
        bil_shift s = bil_of(origin_r[p] - mr, origin_c[p] - mc, qy, qx);
        double wts[4] = {(1 - s.wr) * (1 - s.wc), s.wr * (1 - s.wc), (1 - s.wr) * s.wc, s.wr * s.wc};
        int sr[4] = {s.ir, s.ir + 1, s.ir, s.ir + 1}, sc[4] = {s.ic, s.ic, s.ic + 1, s.ic + 1};
        const float *x = cube + p * ndet;
        for (int c = 0; c < 4; ++c) {
            double g = dy[p] * wts[c];
            for (int a = 0; a < qy; ++a)
                for (int b = 0; b < qx; ++b) {
                    int i = (a + sr[c]) % qy, j = (b + sc[c]) % qx;
                    dw[(int64_t)a * qx + b] += g * (double)x[(int64_t)i * qx + j];
                }
        }
    }
}

/* Gradient with respect to the per-position origins, inside the bilinear cell (d wr / d dr = 1), then
 * through the mean: dr_p = o_p - mean(o), so dL/do_q = g_q - mean(g). */
int oc_vi_bilinear_grad_origin(const float *cube, int64_t npos, int qy, int qx, const double *mask,
                               const double *origin_r, const double *origin_c, const double *dy,
                               double *d_origin_r, double *d_origin_c) {
    int64_t ndet = (int64_t)qy * qx;
    double mr = mean_plain(origin_r, npos), mc = mean_plain(origin_c, npos);
    double *gr = (double *)malloc(sizeof(double) * npos), *gc = (double *)malloc(sizeof(double) * npos);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npos; ++p) {
        bil_shift s = bil_of(origin_r[p] - mr, origin_c[p] - mc, qy, qx);
        const float *x = cube + p * ndet;
        double c00 = 0, c10 = 0, c01 = 0, c11 = 0;
        for (int i = 0; i < qy; ++i)
            for (int j = 0; j < qx; ++j) {
                double v = (double)x[(int64_t)i * qx + j];
                c00 += v * rolled(mask, qy, qx, s.ir, s.ic, i, j);
                c10 += v * rolled(mask, qy, qx, s.ir + 1, s.ic, i, j);
                c01 += v * rolled(mask, qy, qx, s.ir, s.ic + 1, i, j);
                c11 += v * rolled(mask, qy, qx, s.ir + 1, s.ic + 1, i, j);
            }
        /* Y = (1-wr)(1-wc)c00 + wr(1-wc)c10 + (1-wr)wc c01 + wr wc c11 */
        gr[p] = dy[p] * ((1 - s.wc) * (c10 - c00) + s.wc * (c11 - c01));
        gc[p] = dy[p] * ((1 - s.wr) * (c01 - c00) + s.wr * (c11 - c10));
    }
    double mgr = 0, mgc = 0;
    for (int64_t p = 0; p < npos; ++p) { mgr += gr[p]; mgc += gc[p]; }
    mgr /= (double)npos; mgc /= (double)npos;
    for (int64_t p = 0; p < npos; ++p) { d_origin_r[p] = gr[p] - mgr; d_origin_c[p] = gc[p] - mgc; }
    free(gr); free(gc);
    return OC_OK;
}

/* ---------------------------------------------------------------------------------------------- */
/* NumPy pairwise sum: Sources/CuboCore/NumPySum.swift:23-48 (numpy pairwise_sum: < 8 in order;       */
/* <= 128 in eight interleaved partials combined ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)) then the rest;   */
/* larger: split at a multiple of 8 and recurse).                                                    */
/* ---------------------------------------------------------------------------------------------- */

double oc_pairwise_sum(const double *a, int64_t n) {
    if (n < 8) {
        double res = 0.0;
        for (int64_t i = 0; i < n; ++i) res += a[i];
        return res;
    } else if (n <= 128) {
        double r[8];
        for (int k = 0; k < 8; ++k) r[k] = a[k];
        int64_t i = 8;
        for (; i < n - (n % 8); i += 8)
            for (int k = 0; k < 8; ++k) r[k] += a[i + k];
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res += a[i];
        return res;
    } else {
        int64_t n2 = n / 2;
        n2 -= n2 % 8;
        return oc_pairwise_sum(a, n2) + oc_pairwise_sum(a + n2, n - n2);
    }
}

static double np_mean(const double *a, int64_t n) { return n ? oc_pairwise_sum(a, n) / (double)n : NAN; }

/* ---------------------------------------------------------------------------------------------- */
/* FFT in Double with NumPy's conventions (Sources/CuboCore/DoubleFFT.swift:119-195: forward           */
/* exp(-2 pi i jk/N) unnormalised, inverse exp(+) divided by N, rows then columns). Radix-2 for powers */
/* of two, an exact-index direct DFT otherwise (the oracle only meets small grids).                    */
/* ---------------------------------------------------------------------------------------------- */

static void fft1d(double *re, double *im, int n, int inverse, double *tr, double *ti);

/* Lengths above 2048 that are not powers of two: Bluestein's algorithm (the DFT as a chirp convolution, through
 * power-of-two transforms of length M >= 2n - 1): with jk = (j^2 + k^2 - (k - j)^2) / 2,
 * X_k = c_k sum_j (x_j c_j) conj(c_{k-j}), c_j = exp(-+i pi j^2 / n), the angles from j^2 mod 2n in integers. A plan
 * holds the chirp and the transform of the kernel conj(c) (placed at m and M - m) for one length and direction. */
typedef struct {
    int n, M;
    double *cr, *ci, *kr, *ki;
} bluestein_plan;

static int bluestein_needed(int n) { return n > 2048 && (n & (n - 1)) != 0; }

static void bluestein_make(bluestein_plan *p, int n, int inverse) {
    int M = 1;
    while (M < 2 * n - 1) M <<= 1;
    p->n = n;
    p->M = M;
    p->cr = (double *)malloc(sizeof(double) * (size_t)n * 2);
    p->ci = p->cr + n;
    p->kr = (double *)calloc((size_t)M * 2, sizeof(double));
    p->ki = p->kr + M;
    const double sign = inverse ? 1.0 : -1.0;
    for (int j = 0; j < n; ++j) {
        const long long q = ((long long)j * j) % (2LL * n);
        const double ang = sign * M_PI * (double)q / (double)n;
        p->cr[j] = cos(ang);
        p->ci[j] = sin(ang);
        p->kr[j] = p->cr[j];
        p->ki[j] = -p->ci[j];
        if (j > 0) { p->kr[M - j] = p->cr[j]; p->ki[M - j] = -p->ci[j]; }
    }
    double *t = (double *)malloc(sizeof(double) * (size_t)M * 2);
    fft1d(p->kr, p->ki, M, 0, t, t + M);
    free(t);
}

static void bluestein_free(bluestein_plan *p) {
    free(p->cr);
    free(p->kr);
}

/* work: 4 M doubles */
static void bluestein_row(const bluestein_plan *p, double *re, double *im, double *work) {
    const int n = p->n, M = p->M;
    double *ar = work, *ai = work + M, *t = work + 2 * M;
    for (int j = 0; j < M; ++j) {
        if (j < n) {
            ar[j] = re[j] * p->cr[j] - im[j] * p->ci[j];
            ai[j] = re[j] * p->ci[j] + im[j] * p->cr[j];
        } else {
            ar[j] = 0.0;
            ai[j] = 0.0;
        }
    }
    fft1d(ar, ai, M, 0, t, t + M);
    for (int m = 0; m < M; ++m) {
        const double xr = ar[m] * p->kr[m] - ai[m] * p->ki[m], xi = ar[m] * p->ki[m] + ai[m] * p->kr[m];
        ar[m] = xr;
        ai[m] = xi;
    }
    fft1d(ar, ai, M, 1, t, t + M);
    for (int k = 0; k < n; ++k) {
        const double xr = ar[k] / M, xi = ai[k] / M;
        re[k] = xr * p->cr[k] - xi * p->ci[k];
        im[k] = xr * p->ci[k] + xi * p->cr[k];
    }
}

static void bluestein(double *re, double *im, int n, int inverse) {
    bluestein_plan p;
    bluestein_make(&p, n, inverse);
    double *work = (double *)malloc(sizeof(double) * (size_t)p.M * 4);
    bluestein_row(&p, re, im, work);
    free(work);
    bluestein_free(&p);
}

static void fft1d(double *re, double *im, int n, int inverse, double *tr, double *ti) {
    if (n <= 1) return;
    double sign = inverse ? 1.0 : -1.0;
    if ((n & (n - 1)) != 0 && n > 2048) {
        bluestein(re, im, n, inverse);
        return;
    }
    if ((n & (n - 1)) == 0) {
        for (int i = 1, j = 0; i < n; ++i) {            /* bit reversal */
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { double t = re[i]; re[i] = re[j]; re[j] = t; t = im[i]; im[i] = im[j]; im[j] = t; }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int half = len >> 1;
            for (int k = 0; k < half; ++k) {
                double ang = sign * 2.0 * M_PI * (double)k / (double)len;
                double wr = cos(ang), wi = sin(ang);
                for (int s = 0; s < n; s += len) {
                    int a = s + k, b = a + half;
                    double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
                    re[b] = re[a] - xr; im[b] = im[a] - xi;
                    re[a] += xr; im[a] += xi;
                }
            }
        }
    } else {
        for (int k = 0; k < n; ++k) {
            double sr = 0, si = 0;
            for (int j = 0; j < n; ++j) {
                long idx = ((long)j * k) % n;
                double ang = sign * 2.0 * M_PI * (double)idx / (double)n;
                double c = cos(ang), s = sin(ang);
                sr += re[j] * c - im[j] * s;
                si += re[j] * s + im[j] * c;
            }
            tr[k] = sr; ti[k] = si;
        }
        memcpy(re, tr, sizeof(double) * n);
        memcpy(im, ti, sizeof(double) * n);
    }
}

/* Rows, then columns, each line independent of the others (so the threads do not change the result); the Bluestein
 * plans are made once per axis. */
void oc_fft2d(double *re, double *im, int rows, int cols, int inverse) {
    bluestein_plan pr, pc;
    const int br = bluestein_needed(rows), bc = bluestein_needed(cols);
    if (bc) bluestein_make(&pc, cols, inverse);
    if (br) bluestein_make(&pr, rows, inverse);
    const int m = rows > cols ? rows : cols;
    const int mw = 4 * ((bc ? pc.M : 0) > (br ? pr.M : 0) ? (bc ? pc.M : 0) : (br ? pr.M : 0)) + 4 * m;
#pragma omp parallel
    {
        double *tr = (double *)malloc(sizeof(double) * (size_t)mw);
        double *ti = tr + m, *cr = tr + 2 * m, *ci = tr + 3 * m, *work = tr + 4 * m;
#pragma omp for schedule(static)
        for (int r = 0; r < rows; ++r) {
            if (bc) bluestein_row(&pc, re + (int64_t)r * cols, im + (int64_t)r * cols, work);
            else fft1d(re + (int64_t)r * cols, im + (int64_t)r * cols, cols, inverse, tr, ti);
        }
#pragma omp for schedule(static)
        for (int c = 0; c < cols; ++c) {
            for (int r = 0; r < rows; ++r) { cr[r] = re[(int64_t)r * cols + c]; ci[r] = im[(int64_t)r * cols + c]; }
            if (br) bluestein_row(&pr, cr, ci, work);
            else fft1d(cr, ci, rows, inverse, tr, ti);
            for (int r = 0; r < rows; ++r) { re[(int64_t)r * cols + c] = cr[r]; im[(int64_t)r * cols + c] = ci[r]; }
        }
        free(tr);
    }
    if (inverse) {
        double s = 1.0 / ((double)rows * (double)cols);
        for (int64_t k = 0; k < (int64_t)rows * cols; ++k) { re[k] *= s; im[k] *= s; }
    }
    if (bc) bluestein_free(&pc);
    if (br) bluestein_free(&pr);
}

/* ---------------------------------------------------------------------------------------------- */
/* Plane fit and normalisation: Sources/CuboPhase/CenterOfMassReference.swift:27-84.                 */
/* ---------------------------------------------------------------------------------------------- */

int oc_plane_fit(const double *m, int R, int C, double tol, double *out) {
    int64_t N = (int64_t)R * C, nf = 0;
    double rs = 0, cs = 0;
    for (int64_t k = 0; k < N; ++k)
        if (isfinite(m[k])) { ++nf; rs += (double)(k / C); cs += (double)(k % C); }
    if (nf < 3) return OC_PLANE_DEGENERATE;
    double n = (double)nf, rMean = rs / n, cMean = cs / n;
    double suu = 0, suv = 0, svv = 0, sum = 0;
    for (int64_t k = 0; k < N; ++k) {
        if (!isfinite(m[k])) continue;
        double u = (double)(k / C) - rMean, v = (double)(k % C) - cMean;
        suu += u * u; suv += u * v; svv += v * v; sum += m[k];
    }
    double a = sum / n, sum_um = 0, sum_vm = 0;
    for (int64_t k = 0; k < N; ++k) {
        if (!isfinite(m[k])) continue;
        double u = (double)(k / C) - rMean, v = (double)(k % C) - cMean;
        sum_um += u * (m[k] - a); sum_vm += v * (m[k] - a);
    }
    double det = suu * svv - suv * suv;
    if (!(det > tol * fmax(suu * svv, 1.0))) return OC_PLANE_DEGENERATE;
    double b = (svv * sum_um - suv * sum_vm) / det;
    double c = (suu * sum_vm - suv * sum_um) / det;
    for (int64_t k = 0; k < N; ++k) out[k] = a + b * ((double)(k / C) - rMean) + c * ((double)(k % C) - cMean);
    return OC_OK;
}

static double clean_nan(double v) {
    if (isnan(v)) return 0.0;
    if (v == INFINITY) return DBL_MAX;
    if (v == -INFINITY) return -DBL_MAX;
    return v;
}

int oc_normalize(const double *mx, const double *my, int R, int C, int reference, double sampling, double tol,
                 const double *gx, const double *gy, double *fx, double *fy, double *nx, double *ny) {
    int64_t N = (int64_t)R * C;
    if (!(tol >= 0) || !isfinite(tol)) return OC_INVALID_PARAMETER;        /* CenterOfMassReference.swift:30-32 */
    if (reference == 1) {
        int st = oc_plane_fit(mx, R, C, tol, fx);
        if (st) return st;
        st = oc_plane_fit(my, R, C, tol, fy);
        if (st) return st;
    } else if (reference == 2) {
        double a = np_mean(mx, N), b = np_mean(my, N);
        for (int64_t k = 0; k < N; ++k) { fx[k] = a; fy[k] = b; }
    } else if (reference == 3) {                                             /* .given(field), :45-49 */
        for (int64_t k = 0; k < N; ++k) { fx[k] = gx[k]; fy[k] = gy[k]; }
    } else {
        for (int64_t k = 0; k < N; ++k) { fx[k] = 0; fy[k] = 0; }
    }
    for (int64_t k = 0; k < N; ++k) {
        nx[k] = clean_nan(mx[k] - fx[k]) * sampling;
        ny[k] = clean_nan(my[k] - fy[k]) * sampling;
    }
    return OC_OK;
}

/* ---------------------------------------------------------------------------------------------- */
/* R-Q rotation: Sources/CuboPhase/RotationSolver.swift:48-150.                                      */
/* ---------------------------------------------------------------------------------------------- */

void oc_rotate(const double *x, const double *y, int64_t n, double radians, int transpose, double *ox, double *oy) {
    double c = cos(radians), s = sin(radians);
    const double *a = transpose ? y : x, *b = transpose ? x : y;
    for (int64_t k = 0; k < n; ++k) { ox[k] = c * a[k] - s * b[k]; oy[k] = s * a[k] + c * b[k]; }
}

double oc_curl(const double *x, const double *y, int R, int C) {
    int64_t n = (int64_t)(R - 2) * (C - 2), k = 0;
    double *v = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int r = 1; r < R - 1; ++r)
        for (int c = 1; c < C - 1; ++c)
            v[k++] = fabs((y[(int64_t)(r + 1) * C + c] - y[(int64_t)(r - 1) * C + c]) -
                          (x[(int64_t)r * C + c + 1] - x[(int64_t)r * C + c - 1]));
    double m = np_mean(v, n);
    free(v);
    return m;
}

double oc_divergence(const double *x, const double *y, int R, int C) {
    int64_t n = (int64_t)(R - 2) * (C - 2), k = 0;
    double *v = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int r = 1; r < R - 1; ++r)
        for (int c = 1; c < C - 1; ++c)
            v[k++] = fabs((x[(int64_t)(r + 1) * C + c] - x[(int64_t)(r - 1) * C + c]) +
                          (y[(int64_t)r * C + c + 1] - y[(int64_t)r * C + c - 1]));
    double m = np_mean(v, n);
    free(v);
    return m;
}

static int usable(const double *nx, const double *ny, int R, int C) {
    if (R < 3 || C < 3) return OC_SCAN_TOO_SMALL;
    for (int64_t k = 0; k < (int64_t)R * C; ++k)
        if (!isfinite(nx[k]) || !isfinite(ny[k])) return OC_NON_FINITE;
    return OC_OK;
}

int oc_rotation_curve(const double *nx, const double *ny, int R, int C, const double *angles_deg, int nangles,
                      int transpose, int maximize, double *curve) {
    int st = usable(nx, ny, R, C);
    if (st) return st;
    int64_t N = (int64_t)R * C;
#pragma omp parallel for schedule(dynamic)
    for (int a = 0; a < nangles; ++a) {
        double *x = (double *)malloc(sizeof(double) * N * 2), *y = x + N;
        oc_rotate(nx, ny, N, angles_deg[a] * (M_PI / 180.0), transpose, x, y);
        curve[a] = maximize ? oc_divergence(x, y, R, C) : oc_curl(x, y, R, C);
        free(x);
    }
    return OC_OK;
}

/* RotationSolver.swift:48-90, with its three branches: a forced angle (both orientations scored there unless the
 * orientation is forced too), a forced orientation (one curve), or the full search (both curves). */
static int better(int maximize, double a, double b) { return maximize ? a > b : a < b; }

static int best_index(int maximize, const double *curve, int n) {
    int k = 0;
    for (int i = 1; i < n; ++i)
        if (better(maximize, curve[i], curve[k])) k = i;
    return k;
}

int oc_rotation_solve(const double *nx, const double *ny, int R, int C, const double *angles_deg, int nangles,
                      int maximize, const double *force_deg, int force_t, double *degrees, int *transposed,
                      int *n_curve, int *n_curve_t, double *curve, double *curve_t, double *cx, double *cy) {
    int st = usable(nx, ny, R, C);
    if (st) return st;
    *n_curve = 0;
    *n_curve_t = 0;
    if (force_deg) {
        if (!isfinite(*force_deg)) return OC_INVALID_PARAMETER;              /* invalidAngles, :67 */
        *degrees = *force_deg;
        if (force_t >= 0) {
            *transposed = force_t;
        } else {
            oc_rotation_curve(nx, ny, R, C, force_deg, 1, 0, maximize, curve);
            oc_rotation_curve(nx, ny, R, C, force_deg, 1, 1, maximize, curve_t);
            *n_curve = 1;
            *n_curve_t = 1;
            *transposed = better(maximize, curve_t[0], curve[0]);
        }
    } else {
        if (nangles < 1) return OC_INVALID_PARAMETER;
        for (int a = 0; a < nangles; ++a)
            if (!isfinite(angles_deg[a])) return OC_INVALID_PARAMETER;     /* :74-76 */
        if (force_t >= 0) {
            double *c = force_t ? curve_t : curve;
            oc_rotation_curve(nx, ny, R, C, angles_deg, nangles, force_t, maximize, c);
            *(force_t ? n_curve_t : n_curve) = nangles;
            *transposed = force_t;
            *degrees = angles_deg[best_index(maximize, c, nangles)];
        } else {
            oc_rotation_curve(nx, ny, R, C, angles_deg, nangles, 0, maximize, curve);
            oc_rotation_curve(nx, ny, R, C, angles_deg, nangles, 1, maximize, curve_t);
            *n_curve = nangles;
            *n_curve_t = nangles;
            int i = best_index(maximize, curve, nangles), j = best_index(maximize, curve_t, nangles);
            int t = better(maximize, curve_t[j], curve[i]);                 /* the plain orientation wins ties */
            *transposed = t;
            *degrees = angles_deg[t ? j : i];
        }
    }
    oc_rotate(nx, ny, (int64_t)R * C, *degrees * (M_PI / 180.0), *transposed, cx, cy);
    return OC_OK;
}

/* ---------------------------------------------------------------------------------------------- */
/* Iterative DPC: Sources/CuboPhase/DPC.swift:26-102 (py4DSTEM DPC.reconstruct, Double).              */
/* ---------------------------------------------------------------------------------------------- */

static void fftfreq(int n, double d, double *out) {
    double scale = 1.0 / ((double)n * d);
    int positive = (n - 1) / 2 + 1;
    for (int k = 0; k < n; ++k) out[k] = (double)(k < positive ? k : k - n) * scale;
}

typedef struct {
    int R, C, Rp, Cp;
    double sr, sc;
    double *gx, *gy;      /* gx = -0.25 kx kden, the operator is i*gx */
} dpc_ops;

static void dpc_ops_init(dpc_ops *o, int R, int C, int Rp, int Cp, double sr, double sc) {
    o->R = R; o->C = C; o->Rp = Rp; o->Cp = Cp; o->sr = sr; o->sc = sc;
    double *kx = (double *)malloc(sizeof(double) * (Rp + Cp)), *ky = kx + Rp;
    fftfreq(Rp, sr, kx);
    fftfreq(Cp, sc, ky);
    o->gx = (double *)malloc(sizeof(double) * (int64_t)Rp * Cp * 2);
    o->gy = o->gx + (int64_t)Rp * Cp;
    for (int r = 0; r < Rp; ++r)
        for (int c = 0; c < Cp; ++c) {
            double den = kx[r] * kx[r] + ky[c] * ky[c];
            if (r == 0 && c == 0) den = INFINITY;
            double kden = 1.0 / den;
            o->gx[(int64_t)r * Cp + c] = -0.25 * kx[r] * kden;
            o->gy[(int64_t)r * Cp + c] = -0.25 * ky[c] * kden;
        }
    free(kx);
}
static void dpc_ops_free(dpc_ops *o) { free(o->gx); }

/* The forward differences plus the field on the scan region (DPC.swift:59-69); returns the mean square. */
static double dpc_forward(const dpc_ops *o, const double *phi, const double *fx, const double *fy,
                          double *dx, double *dy, double *tmp) {
    int Rp = o->Rp, Cp = o->Cp;
    memset(dx, 0, sizeof(double) * (int64_t)Rp * Cp);
    memset(dy, 0, sizeof(double) * (int64_t)Rp * Cp);
    for (int r = 0; r < o->R; ++r)
        for (int c = 0; c < o->C; ++c) {
            double up = phi[(int64_t)((r - 1 + Rp) % Rp) * Cp + c], down = phi[(int64_t)((r + 1) % Rp) * Cp + c];
            double left = phi[(int64_t)r * Cp + (c - 1 + Cp) % Cp], right = phi[(int64_t)r * Cp + (c + 1) % Cp];
            int64_t k = (int64_t)r * Cp + c;
            dx[k] = (up - down) / (2 * o->sr) + fx[(int64_t)r * o->C + c];
            dy[k] = (left - right) / (2 * o->sc) + fy[(int64_t)r * o->C + c];
            tmp[(int64_t)r * o->C + c] = dx[k] * dx[k] + dy[k] * dy[k];
        }
    return np_mean(tmp, (int64_t)o->R * o->C);
}

/* update = Re(IFFT(FFT(dx) * (i gx) + FFT(dy) * (i gy))) (DPC.swift:71-79); dx, dy are overwritten. */
static void dpc_update(const dpc_ops *o, double *dx, double *dy, double *xi, double *yi, double *ur, double *ui) {
    int64_t n = (int64_t)o->Rp * o->Cp;
    memset(xi, 0, sizeof(double) * n);
    memset(yi, 0, sizeof(double) * n);
    oc_fft2d(dx, xi, o->Rp, o->Cp, 0);
    oc_fft2d(dy, yi, o->Rp, o->Cp, 0);
    for (int64_t k = 0; k < n; ++k) {
        ur[k] = -xi[k] * o->gx[k] + -yi[k] * o->gy[k];
        ui[k] = dx[k] * o->gx[k] + dy[k] * o->gy[k];
    }
    oc_fft2d(ur, ui, o->Rp, o->Cp, 1);
}

int oc_dpc_reconstruct(const double *fx, const double *fy, int R, int C, double sr, double sc, double padding,
                       int max_iter, double step, double stop, int backtrack, double *phase, double *errors,
                       int *accepted, int *n_iter, double *final_step) {
    if (!(sr > 0) || !(sc > 0) || !isfinite(sr) || !isfinite(sc)) return OC_INVALID_PARAMETER;
    if (!(padding >= 1) || !isfinite(padding) || max_iter < 0 || !(step > 0) || !isfinite(step))
        return OC_INVALID_PARAMETER;
    int64_t N = (int64_t)R * C;
    for (int64_t k = 0; k < N; ++k)
        if (!isfinite(fx[k]) || !isfinite(fy[k])) return OC_NON_FINITE;
    double *tmp = (double *)malloc(sizeof(double) * N);
    for (int64_t k = 0; k < N; ++k) tmp[k] = fx[k] * fx[k] + fy[k] * fy[k];
    double denom = np_mean(tmp, N);
    if (!(denom > 0)) { free(tmp); return OC_ZERO_FIELD; }
    int Rp = (int)nearbyint((double)R * padding), Cp = (int)nearbyint((double)C * padding);
    dpc_ops o;
    dpc_ops_init(&o, R, C, Rp, Cp, sr, sc);
    int64_t n = (int64_t)Rp * Cp;
    double *phi = (double *)calloc(n * 7, sizeof(double));
    double *dx = phi + n, *dy = dx + n, *xi = dy + n, *yi = xi + n, *ur = yi + n, *ui = ur + n;
    double error = INFINITY;
    int ne = 0, ni = 0;
    for (int it = 0; it < max_iter; ++it) {
        if (step < stop) break;
        double newError = dpc_forward(&o, phi, fx, fy, dx, dy, tmp) / denom;
        if (newError > error && backtrack) {
            step /= 2;
            accepted[ni++] = 0;
            continue;
        }
        error = newError;
        dpc_update(&o, dx, dy, xi, yi, ur, ui);
        for (int64_t k = 0; k < n; ++k) phi[k] += step * ur[k];
        errors[ne++] = error;
        accepted[ni++] = 1;
    }
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) phase[(int64_t)r * C + c] = phi[(int64_t)r * Cp + c];
    *n_iter = ni;
    *final_step = step;
    (void)ne;
    dpc_ops_free(&o);
    free(phi); free(tmp);
    return OC_OK;
}

/* One accepted DPC iteration without the backtracking decision: phi_out = phi + step * update, and the
 * error of phi (DPC.swift:59-82). */
int oc_dpc_step(const double *phi, const double *fx, const double *fy, int R, int C, int Rp, int Cp,
                double sr, double sc, double step, double *phi_out, double *error_out) {
    int64_t N = (int64_t)R * C, n = (int64_t)Rp * Cp;
    double *tmp = (double *)malloc(sizeof(double) * N);
    for (int64_t k = 0; k < N; ++k) tmp[k] = fx[k] * fx[k] + fy[k] * fy[k];
    double denom = np_mean(tmp, N);
    if (!(denom > 0)) { free(tmp); return OC_ZERO_FIELD; }
    dpc_ops o;
    dpc_ops_init(&o, R, C, Rp, Cp, sr, sc);
    double *buf = (double *)calloc(n * 6, sizeof(double));
    double *dx = buf, *dy = dx + n, *xi = dy + n, *yi = xi + n, *ur = yi + n, *ui = ur + n;
    *error_out = dpc_forward(&o, phi, fx, fy, dx, dy, tmp) / denom;
    dpc_update(&o, dx, dy, xi, yi, ur, ui);
    for (int64_t k = 0; k < n; ++k) phi_out[k] = phi[k] + step * ur[k];
    dpc_ops_free(&o);
    free(buf); free(tmp);
    return OC_OK;
}

/* A^T(g) for A(u) = Re(IFFT(FFT(u) * H)), H = i*gk: Re(IFFT(FFT(g) * conj(H))). out = A_x^T(g), A_y^T(g). */
static void dpc_update_adjoint(const dpc_ops *o, const double *g, double *ax, double *ay) {
    int64_t n = (int64_t)o->Rp * o->Cp;
    double *re = (double *)malloc(sizeof(double) * n * 4), *im = re + n, *r2 = im + n, *i2 = r2 + n;
    memcpy(re, g, sizeof(double) * n);
    memset(im, 0, sizeof(double) * n);
    oc_fft2d(re, im, o->Rp, o->Cp, 0);
    /* (a + ib) * conj(i g) = (a + ib)(-i g) = b g - i a g */
    for (int64_t k = 0; k < n; ++k) { r2[k] = im[k] * o->gx[k]; i2[k] = -re[k] * o->gx[k]; }
    oc_fft2d(r2, i2, o->Rp, o->Cp, 1);
    memcpy(ax, r2, sizeof(double) * n);
    for (int64_t k = 0; k < n; ++k) { r2[k] = im[k] * o->gy[k]; i2[k] = -re[k] * o->gy[k]; }
    oc_fft2d(r2, i2, o->Rp, o->Cp, 1);
    memcpy(ay, r2, sizeof(double) * n);
    free(re);
}

/* Vector-Jacobian product of oc_dpc_step: given dL/dphi_out and dL/derror, dL/dphi, dL/dfx, dL/dfy. */
int oc_dpc_step_vjp(const double *phi, const double *fx, const double *fy, int R, int C, int Rp, int Cp,
                    double sr, double sc, double step, const double *g_phi_out, double g_err,
                    double *g_phi, double *g_fx, double *g_fy) {
    int64_t N = (int64_t)R * C, n = (int64_t)Rp * Cp;
    double *tmp = (double *)malloc(sizeof(double) * N);
    for (int64_t k = 0; k < N; ++k) tmp[k] = fx[k] * fx[k] + fy[k] * fy[k];
    double denom = np_mean(tmp, N);
    if (!(denom > 0)) { free(tmp); return OC_ZERO_FIELD; }
    dpc_ops o;
    dpc_ops_init(&o, R, C, Rp, Cp, sr, sc);
    double *buf = (double *)calloc(n * 4, sizeof(double));
    double *dx = buf, *dy = dx + n, *gdx = dy + n, *gdy = gdx + n;
    double msq = dpc_forward(&o, phi, fx, fy, dx, dy, tmp);
    /* phi_out = phi + step*(A_x dx + A_y dy): d/d dx = step * A_x^T g */
    dpc_update_adjoint(&o, g_phi_out, gdx, gdy);
    for (int64_t k = 0; k < n; ++k) { gdx[k] *= step; gdy[k] *= step; }
    /* error = mean_scan(dx^2 + dy^2) / denom: d/d dx = 2 dx / (N denom) on the scan region */
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) {
            int64_t k = (int64_t)r * Cp + c;
            gdx[k] += g_err * 2.0 * dx[k] / ((double)N * denom);
            gdy[k] += g_err * 2.0 * dy[k] / ((double)N * denom);
        }
    /* dx is zero off the scan region, so only the scan entries of gdx, gdy flow back */
    for (int64_t k = 0; k < n; ++k) g_phi[k] = g_phi_out[k];
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) {
            int64_t k = (int64_t)r * Cp + c;
            double a = gdx[k] / (2 * sr), b = gdy[k] / (2 * sc);
            /* dx[r,c] = (phi[r-1,c] - phi[r+1,c]) / (2 sr) + fx; dy[r,c] = (phi[r,c-1] - phi[r,c+1]) / (2 sc) + fy */
            g_phi[(int64_t)((r - 1 + Rp) % Rp) * Cp + c] += a;
            g_phi[(int64_t)((r + 1) % Rp) * Cp + c] -= a;
            g_phi[(int64_t)r * Cp + (c - 1 + Cp) % Cp] += b;
            g_phi[(int64_t)r * Cp + (c + 1) % Cp] -= b;
            g_fx[(int64_t)r * C + c] = gdx[k];
            g_fy[(int64_t)r * C + c] = gdy[k];
        }
    /* the denominator depends on the field: error = msq / denom, denom = mean(fx^2 + fy^2) */
    double dd = -g_err * msq / (denom * denom);
    for (int64_t k = 0; k < N; ++k) {
        g_fx[k] += dd * 2.0 * fx[k] / (double)N;
        g_fy[k] += dd * 2.0 * fy[k] / (double)N;
    }
    dpc_ops_free(&o);
    free(buf); free(tmp);
    return OC_OK;
}

/* ---------------------------------------------------------------------------------------------- */
/* CPU baseline: the reference's .cpu algorithms over a uint16 cube, tile by tile converted to        */
/* Float32 (MappedCube.swift:150-164, ScanTiling.automatic CubeSource.swift:96-100), VI per position  */
/* serially (CPUReducers.swift:13-39), mean+max on 128-pixel blocks across threads (:46-97), CoM       */
/* parallel over positions (CenterOfMass.swift:51-79).                                              */
/* ---------------------------------------------------------------------------------------------- */

void oc_cpu_first_processing_u16(const uint16_t *cube, int64_t npos, int qy, int qx, int64_t rows_per_tile,
                                 int64_t positions_per_row, const float *bf, const float *adf, int threads,
                                 double *vi_bf, double *vi_adf, double *mean, double *max,
                                 double *s, double *mr, double *mc) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    int64_t ndet = (int64_t)qy * qx;
    int64_t tile_pos = rows_per_tile * positions_per_row;
    float *tile = (float *)malloc(sizeof(float) * tile_pos * ndet);
    double *sum = (double *)malloc(sizeof(double) * ndet), *corr = (double *)malloc(sizeof(double) * ndet);
    float *mx = (float *)malloc(sizeof(float) * ndet);
    double *pat = (double *)malloc(sizeof(double) * ndet);
    double *bfd = (double *)malloc(sizeof(double) * ndet), *adfd = (double *)malloc(sizeof(double) * ndet);
    for (int64_t i = 0; i < ndet; ++i) {
        sum[i] = -0.0; corr[i] = -0.0; mx[i] = -INFINITY; bfd[i] = bf[i]; adfd[i] = adf[i];
    }
    const int64_t block = 128, nblocks = (ndet + block - 1) / block;
    for (int64_t p0 = 0; p0 < npos; p0 += tile_pos) {
        int64_t np = npos - p0 < tile_pos ? npos - p0 : tile_pos;
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < np * ndet; ++k) tile[k] = (float)cube[p0 * ndet + k];
        for (int64_t p = 0; p < np; ++p) {                 /* VI: serial over positions, as the .cpu path */
            const float *x = tile + p * ndet;
            for (int64_t i = 0; i < ndet; ++i) pat[i] = (double)x[i];
            double a = 0, b = 0;
            for (int64_t i = 0; i < ndet; ++i) a += pat[i] * bfd[i];
            for (int64_t i = 0; i < ndet; ++i) b += pat[i] * adfd[i];
            vi_bf[p0 + p] = a; vi_adf[p0 + p] = b;
        }
#pragma omp parallel for schedule(dynamic)
        for (int64_t b = 0; b < nblocks; ++b) {
            int64_t lo = b * block, hi = lo + block < ndet ? lo + block : ndet;
            mean_max_block(tile, np, ndet, lo, hi, sum, corr, mx);
        }
        oc_com_sums(tile, np, qy, qx, NULL, s + p0, mr + p0, mc + p0);
    }
    for (int64_t i = 0; i < ndet; ++i) {
        mean[i] = comp_value(sum[i], corr[i]) / (double)npos;
        max[i] = (double)mx[i];
    }
    free(tile); free(sum); free(corr); free(mx); free(pat); free(bfd); free(adfd);
}

/* ---------------------------------------------------------------------------------------------- */
/* Origins.                                                                                          */
/*  - py4DSTEM's get_origin (OriginMeasurement.swift:52-95): scipy's gaussian_filter (sigma = r,      */
/*    truncate 4, mode nearest; weights exp(-0.5/sigma^2 x^2) over NumPy's pairwise sum, SciPyKernels */
/*    .swift:66-78; axis 0 then axis 1 in NI_Correlate1D's order, :128-150), the first argmax (the   */
/*    first NaN if any), the Neumaier centre of mass of p * [hypot < r * rscale];                    */
/*  - ProbeSize.estimate (ProbeSize.swift:25-58);                                                    */
/*  - the iterated-window estimator in Float32 (IteratedWindowOrigin.swift:48-70, 136-180).          */
/* ---------------------------------------------------------------------------------------------- */

int oc_gaussian_weights(double sigma, double *w, int max_len) {
    const int R = (int)(4.0 * sigma + 0.5);
    if (2 * R + 1 > max_len) return -1;
    const double sigma2 = sigma * sigma;
    for (int x = -R; x <= R; ++x) w[x + R] = exp(-0.5 / sigma2 * (double)(x * x));
    const double s = oc_pairwise_sum(w, 2 * R + 1);
    for (int k = 0; k <= 2 * R; ++k) w[k] = w[k] / s;
    return R;
}

static inline int nearest(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

void oc_gaussian_filter(const double *p, int qy, int qx, const double *w, int R, double *out) {
    double *a = (double *)malloc(sizeof(double) * qy * qx);
    const double centre = w[R];
    for (int i = 0; i < qy; ++i)
        for (int j = 0; j < qx; ++j) {
            double acc = p[i * qx + j] * centre;
            for (int t = R; t >= 1; --t) acc += (p[nearest(i - t, qy) * qx + j] + p[nearest(i + t, qy) * qx + j]) * w[R - t];
            a[i * qx + j] = acc;
        }
    for (int i = 0; i < qy; ++i)
        for (int j = 0; j < qx; ++j) {
            double acc = a[i * qx + j] * centre;
            for (int t = R; t >= 1; --t) acc += (a[i * qx + nearest(j - t, qx)] + a[i * qx + nearest(j + t, qx)]) * w[R - t];
            out[i * qx + j] = acc;
        }
    free(a);
}

int oc_origin_py4dstem(const float *cube, int64_t npos, int qy, int qx, double radius, double rscale,
                       const uint8_t *take, int fast_center, double *out_r, double *out_c) {
    if (!(radius > 0) || !isfinite(radius) || radius * radius < 2.2250738585072014e-308 || !(4.0 * radius + 0.5 < 1e6))
        return OC_INVALID_PARAMETER;
    if (!(rscale > 0) || !isfinite(rscale)) return OC_INVALID_PARAMETER;
    const int64_t n = (int64_t)qy * qx;
    const int len = 2 * (int)(4.0 * radius + 0.5) + 1;
    double *w = (double *)malloc(sizeof(double) * len);
    const int R = oc_gaussian_weights(radius, w, len);
    const double limit = radius * rscale;
#pragma omp parallel for schedule(dynamic)
    for (int64_t k = 0; k < npos; ++k) {
        if (take && !take[k]) { out_r[k] = NAN; out_c[k] = NAN; continue; }
        double *p = (double *)malloc(sizeof(double) * n), *g = (double *)malloc(sizeof(double) * n);
        for (int64_t i = 0; i < n; ++i) p[i] = (double)cube[k * n + i];
        oc_gaussian_filter(p, qy, qx, w, R, g);
        int64_t best = 0, nan = -1;
        for (int64_t i = 0; i < n; ++i) if (isnan(g[i])) { nan = i; break; }
        if (nan >= 0) best = nan; else for (int64_t i = 1; i < n; ++i) if (g[i] > g[best]) best = i;
        const double i0 = (double)(best / qx), j0 = (double)(best % qx);
        if (fast_center) { out_r[k] = i0; out_c[k] = j0; free(p); free(g); continue; }
        neumaier s = {0, 0}, mr = {0, 0}, mc = {0, 0};
        for (int64_t i = 0; i < n; ++i) {
            const double di = (double)(i / qx) - i0, dj = (double)(i % qx) - j0;
            const double v = p[i] * (hypot(di, dj) < limit ? 1.0 : 0.0);
            nm_add(&s, v); nm_add(&mr, (double)(i / qx) * v); nm_add(&mc, (double)(i % qx) * v);
        }
        out_r[k] = nm_value(&mr) / nm_value(&s);
        out_c[k] = nm_value(&mc) / nm_value(&s);
        free(p); free(g);
    }
    free(w);
    return OC_OK;
}

static int cmp_double(const void *a, const void *b) {
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int oc_probe_size(const double *v, int rows, int cols, double lower, double upper, int count, double out[4]) {
    if (count < 2) return OC_INVALID_PARAMETER;
    const int64_t n = (int64_t)rows * cols;
    double *th = (double *)malloc(sizeof(double) * count), *radii = (double *)malloc(sizeof(double) * count);
    double *grad = (double *)malloc(sizeof(double) * count), *sorted = (double *)malloc(sizeof(double) * count);
    const double step = (upper - lower) / (double)(count - 1);
    for (int k = 0; k < count; ++k) th[k] = (double)k * step + lower;
    th[count - 1] = upper;
    double maximum = -INFINITY;
    for (int64_t i = 0; i < n; ++i) maximum = (isnan(maximum) || isnan(v[i])) ? NAN : (v[i] > maximum ? v[i] : maximum);
    for (int k = 0; k < count; ++k) {
        const double limit = maximum * th[k];
        int64_t c = 0;
        for (int64_t i = 0; i < n; ++i) c += v[i] > limit;
        radii[k] = sqrt((double)c / M_PI);
    }
    grad[0] = radii[1] - radii[0];
    grad[count - 1] = radii[count - 1] - radii[count - 2];
    for (int k = 1; k < count - 1; ++k) grad[k] = (radii[k + 1] - radii[k - 1]) / 2.0;
    memcpy(sorted, grad, sizeof(double) * count);
    qsort(sorted, count, sizeof(double), cmp_double);
    const double median = count % 2 ? sorted[count / 2] : (sorted[count / 2 - 1] + sorted[count / 2]) / 2;
    double *tr = (double *)malloc(sizeof(double) * count), *tt = (double *)malloc(sizeof(double) * count);
    int nt = 0;
    for (int k = 0; k < count; ++k)
        if (grad[k] <= 0 && grad[k] >= 2 * median) { tr[nt] = radii[k]; tt[nt] = th[k]; ++nt; }
    const double radius = nt ? oc_pairwise_sum(tr, nt) / (double)nt : NAN;
    const double threshold = nt ? oc_pairwise_sum(tt, nt) / (double)nt : NAN;
    const double limit = maximum * threshold;
    double *m = (double *)malloc(sizeof(double) * n), *mr = (double *)malloc(sizeof(double) * n),
           *mc = (double *)malloc(sizeof(double) * n);
    for (int64_t i = 0; i < n; ++i) {
        m[i] = v[i] * (v[i] > limit ? 1.0 : 0.0);
        mr[i] = (double)(i / cols) * m[i];
        mc[i] = (double)(i % cols) * m[i];
    }
    const double total = oc_pairwise_sum(m, n);
    out[0] = radius;
    out[1] = oc_pairwise_sum(mr, n) / total;
    out[2] = oc_pairwise_sum(mc, n) / total;
    out[3] = threshold;
    free(th); free(radii); free(grad); free(sorted); free(tr); free(tt); free(m); free(mr); free(mc);
    return (isfinite(out[0]) && isfinite(out[1]) && isfinite(out[2])) ? OC_OK : OC_ZERO_FIELD;
}

static float iw_filled(const float *pat, int row, int col, int qy, int qx) {
    const float v = pat[row * qx + col];
    if (isfinite(v)) return v;
    float nb[8];
    int k = 0;
    for (int dr = -1; dr <= 1; ++dr)
        for (int dc = -1; dc <= 1; ++dc) {
            if (dr == 0 && dc == 0) continue;
            const int rr = row + dr, cc = col + dc;
            if (rr < 0 || rr >= qy || cc < 0 || cc >= qx) continue;
            const float w = pat[rr * qx + cc];
            if (!isfinite(w)) continue;
            int i = k++;
            while (i > 0 && nb[i - 1] > w) { nb[i] = nb[i - 1]; --i; }
            nb[i] = w;
        }
    if (k == 0) return 0.0f;
    return (k % 2 == 1) ? nb[k / 2] : 0.5f * (nb[k / 2 - 1] + nb[k / 2]);
}

void oc_origin_iterated(const float *cube, int64_t npos, int qy, int qx, float r, float rscale, int passes, float margin,
                        int binned, double *out_r, double *out_c) {
    const int64_t n = (int64_t)qy * qx;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < npos; ++k) {
        const float *pat = cube + k * n;
        const int bin = binned ? (roundf(r) > 1 ? (int)roundf(r) : 1) : 1;
        float best = -3.402823466e38f;
        float cr = 0.5f * (float)(qy - 1), cc = 0.5f * (float)(qx - 1);
        for (int by = 0; by < qy; by += bin) {
            const int ye = by + bin < qy ? by + bin : qy;
            for (int bx = 0; bx < qx; bx += bin) {
                const int xe = bx + bin < qx ? bx + bin : qx;
                float s = 0.0f;
                for (int y = by; y < ye; ++y)
                    for (int x = bx; x < xe; ++x) s += iw_filled(pat, y, x, qy, qx);
                if (s > best) { best = s; cc = 0.5f * (float)(bx + xe - 1); cr = 0.5f * (float)(by + ye - 1); }
            }
        }
        const float rs = r * rscale, rm = r + margin;
        const float win = rm >= rs ? rm : rs, win2 = win * win;
        for (int pass = 0; pass < passes; ++pass) {
            int y0 = (int)floorf(cr - win), y1 = (int)ceilf(cr + win), x0 = (int)floorf(cc - win), x1 = (int)ceilf(cc + win);
            y0 = y0 < 0 ? 0 : y0; x0 = x0 < 0 ? 0 : x0;
            y1 = y1 > qy - 1 ? qy - 1 : y1; x1 = x1 > qx - 1 ? qx - 1 : x1;
            float sI = 0.0f, sX = 0.0f, sY = 0.0f;
            for (int y = y0; y <= y1; ++y) {
                const float dy = (float)y - cr, dy2 = dy * dy;
                if (dy2 > win2) continue;
                for (int x = x0; x <= x1; ++x) {
                    const float dx = (float)x - cc;
                    if (fmaf(dx, dx, dy2) > win2) continue;
                    const float f = iw_filled(pat, y, x, qy, qx);
                    const float v = 0.0f >= f ? 0.0f : f;
                    sI += v;
                    sX = fmaf(v, (float)x, sX);
                    sY = fmaf(v, (float)y, sY);
                }
            }
            if (!(sI > 0.0f)) break;
            const float nc = sX / sI, nr = sY / sI;
            const int settled = fabsf(nc - cc) < 1e-4f && fabsf(nr - cr) < 1e-4f;
            cc = nc; cr = nr;
            if (settled) break;
        }
        out_r[k] = (double)cr;
        out_c[k] = (double)cc;
    }
}
