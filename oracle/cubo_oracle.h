/*
 * cubo_oracle.h — CPU restatement of CuboKit's first-processing hot path, used ONLY as a test checker
 * and as the CPU baseline of the benchmark. Nothing in the product library links or calls it.
 *
 * Every function restates one reference function (cited at its definition in cubo_oracle.c) in plain C
 * with Double arithmetic. Cubes are passed as Float32 values in storage order [scan position][detector
 * pixel], because the reference converts every element type to Float32 on read
 * (Sources/CuboCore/MappedCube.swift:56-83).
 */
#ifndef CUBO_ORACLE_H
#define CUBO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (mirroring the refusals of DetectorMask.make and CuboPhase). */
enum {
    OC_OK = 0,
    OC_INVALID_SHAPE = -1,
    OC_NON_FINITE = -2,
    OC_NEGATIVE_RADIUS = -3,
    OC_INNER_NOT_BELOW_OUTER = -4,
    OC_NEGATIVE_RECT_BOUND = -5,
    OC_POINT_OUTSIDE = -6,
    OC_PLANE_DEGENERATE = -7,
    OC_ZERO_FIELD = -8,
    OC_SCAN_TOO_SMALL = -9,
    OC_INVALID_PARAMETER = -10,
};

/* Detector masks (py4DSTEM make_detector, evaluated in Double). */
int oc_mask_point(int qy, int qx, int row, int col, float *out);
int oc_mask_circle(int qy, int qx, double cr, double cc, double r, float *out);
int oc_mask_annulus(int qy, int qx, double cr, double cc, double inner, double outer, float *out);
int oc_mask_rectangle(int qy, int qx, double r0, double r1, double c0, double c1, float *out);

/* Virtual image: out[p] = sum_i double(x[p,i]) * double(m[i]). */
void oc_virtual_image(const float *cube, int64_t npos, int64_t ndet, const float *mask, double *out);

/* Mean (TwoSum per pixel in scan order, seeded -0) and max (Float32, seeded -inf, NaN wins). */
void oc_mean_max(const float *cube, int64_t npos, int64_t ndet, double *mean, double *max);

/* Centre-of-mass sums with Neumaier compensation; weights NULL means all ones. */
void oc_com_sums(const float *cube, int64_t npos, int qy, int qx, const float *weights,
                 double *s, double *mr, double *mc);

/* Mask stack: Y[p,k] = sum_i double(x[p,i]) * W[k,i], W given in Double, K rows. */
void oc_mask_stack(const float *cube, int64_t npos, int64_t ndet, const double *w, int k, double *y);

/* Bilinear centre-shifted virtual image (MaskShift.bilinear). origin_* are per position. */
int oc_vi_bilinear(const float *cube, int64_t npos, int qy, int qx, const double *mask,
                   const double *origin_r, const double *origin_c, double *out);

/* Rounded centre-shifted virtual image (MaskShift.rounded). */
int oc_vi_rounded(const float *cube, int64_t npos, int qy, int qx, const double *mask,
                  const double *origin_r, const double *origin_c, double *out);

/* Adjoints of the reductions above, in Double. */
void oc_mask_stack_grad_w(const float *cube, int64_t npos, int64_t ndet, const double *dy, int k, double *dw);
void oc_vi_bilinear_grad_w(const float *cube, int64_t npos, int qy, int qx, const double *origin_r,
                           const double *origin_c, const double *dy, double *dw);
int oc_vi_bilinear_grad_origin(const float *cube, int64_t npos, int qy, int qx, const double *mask,
                               const double *origin_r, const double *origin_c, const double *dy,
                               double *d_origin_r, double *d_origin_c);

/* Selected-area and centre-shifted patterns: stat 0 = mean, 1 = max; take (selection flags) or w (weights) or
 * neither; orow/ocol an origin map for PatternShift.rounded (fourier = 0) or .fourier (1), or NULL; ties_later: Swift.max's tie rule
 * (meanAndMax(of:shift:)) instead of the first-kept one. OC_INVALID_PARAMETER for an empty selection or take with
 * w, OC_NON_FINITE for non-finite weights or origins, OC_ZERO_FIELD for a mean over weights summing to zero. */
int oc_pattern(const float *cube, int64_t npos, int qy, int qx, int stat, const uint8_t *take, const double *w,
               const double *orow, const double *ocol, int ties_later, int fourier, double *out);

/* Origins: py4DSTEM's get_origin (Double), ProbeSize.estimate (out: radius, centre row, centre column, threshold)
 * and the iterated-window estimator (Float32). Gaussian weights return R (the kernel is 2R + 1 long), or -1. */
int oc_gaussian_weights(double sigma, double *w, int max_len);
void oc_gaussian_filter(const double *p, int qy, int qx, const double *w, int R, double *out);
int oc_origin_py4dstem(const float *cube, int64_t npos, int qy, int qx, double radius, double rscale,
                       const uint8_t *take, int fast_center, double *out_r, double *out_c);
int oc_probe_size(const double *v, int rows, int cols, double lower, double upper, int count, double out[4]);
void oc_origin_iterated(const float *cube, int64_t npos, int qy, int qx, float r, float rscale, int passes, float margin,
                        int binned, double *out_r, double *out_c);

/* NumPy pairwise float64 sum. */
double oc_pairwise_sum(const double *a, int64_t n);

/* NumPy-convention 2-D complex FFT in Double, in place (inverse divides by rows*cols). */
void oc_fft2d(double *re, double *im, int rows, int cols, int inverse);

/* CuboPhase chain. */
int oc_plane_fit(const double *m, int R, int C, double tol, double *out);
/* reference: 0 none, 1 plane, 2 constant, 3 the given field (gx, gy; unused otherwise). */
int oc_normalize(const double *mx, const double *my, int R, int C, int reference, double sampling, double tol,
                 const double *gx, const double *gy, double *fx, double *fy, double *nx, double *ny);
void oc_rotate(const double *x, const double *y, int64_t n, double radians, int transpose, double *ox, double *oy);
double oc_curl(const double *x, const double *y, int R, int C);
double oc_divergence(const double *x, const double *y, int R, int C);
int oc_rotation_curve(const double *nx, const double *ny, int R, int C, const double *angles_deg, int nangles,
                      int transpose, int maximize, double *curve);
/* force_deg: NULL or the forced angle; force_t: -1 (search the orientation), 0 or 1. n_curve / n_curve_t receive
 * the length of each curve evaluated (0 when that orientation was not evaluated). */
int oc_rotation_solve(const double *nx, const double *ny, int R, int C, const double *angles_deg, int nangles,
                      int maximize, const double *force_deg, int force_t, double *degrees, int *transposed,
                      int *n_curve, int *n_curve_t, double *curve, double *curve_t, double *cx, double *cy);
int oc_dpc_reconstruct(const double *fx, const double *fy, int R, int C, double sr, double sc, double padding,
                       int max_iter, double step, double stop, int backtrack, double *phase, double *errors,
                       int *accepted, int *n_iter, double *final_step);

/* One DPC phase step on the padded grid (forward difference + adjoint update), and its adjoint. */
int oc_dpc_step(const double *phi, const double *fx, const double *fy, int R, int C, int Rp, int Cp,
                double sr, double sc, double step, double *phi_out, double *error_out);
int oc_dpc_step_vjp(const double *phi, const double *fx, const double *fy, int R, int C, int Rp, int Cp,
                    double sr, double sc, double step, const double *g_phi_out, double g_err,
                    double *g_phi, double *g_fx, double *g_fy);

/* The reference's .cpu first-processing pass on a uint16 cube, timed as the CPU baseline: tiles of at
 * most 64 MiB converted to Float32, then VI per position (serial), mean+max over 128-pixel blocks on
 * `threads` threads, CoM in parallel over positions. */
void oc_cpu_first_processing_u16(const uint16_t *cube, int64_t npos, int qy, int qx, int64_t rows_per_tile,
                                 int64_t positions_per_row, const float *bf, const float *adf, int threads,
                                 double *vi_bf, double *vi_adf, double *mean, double *max,
                                 double *s, double *mr, double *mc);

#ifdef __cplusplus
}
#endif
#endif
