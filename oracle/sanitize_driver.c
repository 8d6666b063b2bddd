/*
 * sanitize_driver.c — test infrastructure only: runs every function of the oracle (cubo_oracle.c) on small and
 * ragged inputs (odd detector sides, one position, partial tiles, non-finite values, refusals) under AddressSanitizer
 * and UndefinedBehaviorSanitizer (oracle/Makefile, target `sanitize`). It checks memory safety and undefined
 * behaviour of the checker's own indexing, not its numbers: those are pinned by the tests under tests/.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cubo_oracle.h"

static uint64_t state = 0x9e3779b97f4a7c15ull;
static double uniform(void) {
    state = state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(state >> 11) * (1.0 / 9007199254740992.0);
}
static float *cube_f32(int64_t n, int special) {
    float *c = (float *)malloc(sizeof(float) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) c[i] = (float)(uniform() * 100.0);
    if (special && n > 3) { c[0] = NAN; c[1] = INFINITY; c[2] = -INFINITY; c[3] = -0.0f; }
    return c;
}
static double *vec(int64_t n, double lo, double hi) {
    double *v = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) v[i] = lo + (hi - lo) * uniform();
    return v;
}

static void reductions(int ry, int rx, int qy, int qx, int special) {
    const int64_t npos = (int64_t)ry * rx, ndet = (int64_t)qy * qx;
    float *cube = cube_f32(npos * ndet, special);
    float *m = (float *)malloc(sizeof(float) * (size_t)ndet);
    double *out = vec(npos, 0, 0), *mean = vec(ndet, 0, 0), *mx = vec(ndet, 0, 0);
    double *s = vec(npos, 0, 0), *mr = vec(npos, 0, 0), *mc = vec(npos, 0, 0);
    oc_mask_point(qy, qx, qy - 1, qx - 1, m);
    oc_mask_point(qy, qx, qy, 0, m);                                 /* refused: outside */
    oc_mask_circle(qy, qx, qy / 2.0, qx / 2.0, qy / 3.0 + 0.5, m);
    oc_mask_annulus(qy, qx, qy / 2.0, qx / 2.0, 1.0, qy / 2.0 + 0.3, m);
    oc_mask_annulus(qy, qx, 1, 1, 3, 2, m);                          /* refused: inner >= outer */
    oc_mask_rectangle(qy, qx, 0.5, qy + 3.0, 1.5, qx - 0.5, m);
    oc_mask_rectangle(qy, qx, -1, 2, 0, 2, m);                       /* refused: negative bound */
    oc_mask_circle(qy, qx, qy / 2.0, qx / 2.0, qy / 3.0 + 0.5, m);
    oc_virtual_image(cube, npos, ndet, m, out);
    oc_mean_max(cube, npos, ndet, mean, mx);
    oc_com_sums(cube, npos, qy, qx, NULL, s, mr, mc);
    oc_com_sums(cube, npos, qy, qx, m, s, mr, mc);
    const int k = 3;
    double *w = vec((int64_t)k * ndet, -1, 1), *y = vec(npos * k, 0, 0), *dy = vec(npos * k, -1, 1);
    double *dw = vec((int64_t)k * ndet, 0, 0);
    oc_mask_stack(cube, npos, ndet, w, k, y);
    oc_mask_stack_grad_w(cube, npos, ndet, dy, k, dw);
    double *orr = vec(npos, qy / 2.0 - 3.7, qy / 2.0 + 3.7), *occ = vec(npos, qx / 2.0 - 2.2, qx / 2.0 + 5.1);
    double *dr = vec(npos, 0, 0), *dc = vec(npos, 0, 0), *d1 = vec(npos, -1, 1);
    oc_vi_bilinear(cube, npos, qy, qx, w, orr, occ, out);
    oc_vi_rounded(cube, npos, qy, qx, w, orr, occ, out);
    oc_vi_bilinear_grad_w(cube, npos, qy, qx, orr, occ, d1, dw);
    oc_vi_bilinear_grad_origin(cube, npos, qy, qx, w, orr, occ, d1, dr, dc);
    uint8_t *take = (uint8_t *)malloc((size_t)npos);
    for (int64_t i = 0; i < npos; ++i) take[i] = (uint8_t)(i % 3 != 1);
    double *pw = vec(npos, 0.1, 2);
    for (int stat = 0; stat < 2; ++stat)
        for (int ties = 0; ties < 2; ++ties) {
            oc_pattern(cube, npos, qy, qx, stat, take, NULL, NULL, NULL, ties, 0, mean);
            oc_pattern(cube, npos, qy, qx, stat, NULL, pw, NULL, NULL, ties, 0, mean);
            oc_pattern(cube, npos, qy, qx, stat, take, NULL, orr, occ, ties, 0, mean);
            oc_pattern(cube, npos, qy, qx, stat, take, NULL, orr, occ, ties, 1, mean);
            oc_pattern(cube, npos, qy, qx, stat, NULL, NULL, NULL, NULL, ties, 0, mean);
        }
    double gw[257];
    const int R = oc_gaussian_weights(1.7, gw, 257);
    double *pat = vec(ndet, 0, 10), *filt = vec(ndet, 0, 0);
    if (R >= 0) oc_gaussian_filter(pat, qy, qx, gw, R, filt);
    oc_origin_py4dstem(cube, npos, qy, qx, qy / 6.0 + 0.5, 1.2, take, 0, dr, dc);
    oc_origin_py4dstem(cube, npos, qy, qx, qy / 6.0 + 0.5, 1.2, NULL, 1, dr, dc);
    for (int binned = 0; binned < 2; ++binned)
        oc_origin_iterated(cube, npos, qy, qx, (float)(qy / 6.0 + 0.5), 1.2f, 4, 1.5f, binned, dr, dc);
    double est[4];
    oc_probe_size(pat, qy, qx, 0.01, 0.99, 100, est);
    uint16_t *u16 = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(npos * ndet));
    for (int64_t i = 0; i < npos * ndet; ++i) u16[i] = (uint16_t)(uniform() * 65535.0);
    float *bf = (float *)malloc(sizeof(float) * (size_t)ndet), *adf = (float *)malloc(sizeof(float) * (size_t)ndet);
    oc_mask_circle(qy, qx, qy / 2.0, qx / 2.0, qy / 4.0 + 0.5, bf);
    oc_mask_annulus(qy, qx, qy / 2.0, qx / 2.0, qy / 4.0 + 0.5, qy / 2.0 + 0.5, adf);
    for (int64_t tile = 1; tile <= ry; tile += 2)
        oc_cpu_first_processing_u16(u16, npos, qy, qx, tile, rx, bf, adf, 2, out, dr, mean, mx, s, mr, mc);
    free(cube); free(m); free(out); free(mean); free(mx); free(s); free(mr); free(mc); free(w); free(y); free(dy);
    free(dw); free(orr); free(occ); free(dr); free(dc); free(d1); free(take); free(pw); free(pat); free(filt);
    free(u16); free(bf); free(adf);
}

static void phase(int R, int C) {
    const int64_t n = (int64_t)R * C;
    double *mx = vec(n, -1, 1), *my = vec(n, -1, 1), *fx = vec(n, 0, 0), *fy = vec(n, 0, 0);
    double *nx = vec(n, 0, 0), *ny = vec(n, 0, 0), *cx = vec(n, 0, 0), *cy = vec(n, 0, 0);
    mx[0] = NAN;
    my[n - 1] = INFINITY;
    for (int ref = 0; ref < 4; ++ref) oc_normalize(mx, my, R, C, ref, 2.0, 1e-12, mx, my, fx, fy, nx, ny);
    oc_normalize(mx, my, R, C, 1, 1.0, -1.0, NULL, NULL, fx, fy, nx, ny);   /* refused: negative tolerance */
    oc_normalize(mx, my, R, C, 0, 1.0, 1e-12, NULL, NULL, fx, fy, nx, ny);
    double ang[179], curve[179], curve_t[179], deg, forced = 23.0;
    int t, n0, n1;
    for (int a = 0; a < 179; ++a) ang[a] = a - 89;
    for (int maximize = 0; maximize < 2; ++maximize)
        for (int ft = -1; ft < 2; ++ft) {
            oc_rotation_solve(nx, ny, R, C, ang, 179, maximize, NULL, ft, &deg, &t, &n0, &n1, curve, curve_t, cx, cy);
            oc_rotation_solve(nx, ny, R, C, ang, 179, maximize, &forced, ft, &deg, &t, &n0, &n1, curve, curve_t, cx,
                              cy);
        }
    oc_divergence(nx, ny, R, C);
    oc_curl(nx, ny, R, C);
    int Rp = (int)lround(R * 2.0), Cp = (int)lround(C * 2.0);
    double *phase = vec(n, 0, 0), *err = vec(8, 0, 0), *phi = vec((int64_t)Rp * Cp, -1, 1);
    double *phi2 = vec((int64_t)Rp * Cp, 0, 0), *gphi = vec((int64_t)Rp * Cp, 0, 0), e2, fs;
    int acc[8], it;
    oc_dpc_reconstruct(nx, ny, R, C, 1.0, 1.5, 2.0, 8, 0.5, 1e-6, 1, phase, err, acc, &it, &fs);
    oc_dpc_step(phi, nx, ny, R, C, Rp, Cp, 1.0, 1.0, 0.5, phi2, &e2);
    oc_dpc_step_vjp(phi, nx, ny, R, C, Rp, Cp, 1.0, 1.0, 0.5, phi2, 0.3, gphi, fx, fy);
    double *re = vec((int64_t)Rp * Cp, -1, 1), *im = vec((int64_t)Rp * Cp, -1, 1);
    oc_fft2d(re, im, Rp, Cp, 0);
    oc_fft2d(re, im, Rp, Cp, 1);
    oc_pairwise_sum(re, (int64_t)Rp * Cp);
    oc_pairwise_sum(re, 0);
    free(mx); free(my); free(fx); free(fy); free(nx); free(ny); free(cx); free(cy); free(phase); free(err); free(phi);
    free(phi2); free(gphi); free(re); free(im);
}

int main(void) {
    const int shapes[][5] = {{1, 1, 8, 8, 0}, {3, 5, 12, 12, 1}, {4, 6, 7, 9, 1}, {2, 3, 16, 16, 0}, {5, 1, 5, 13, 1}};
    for (size_t i = 0; i < sizeof shapes / sizeof shapes[0]; ++i)
        reductions(shapes[i][0], shapes[i][1], shapes[i][2], shapes[i][3], shapes[i][4]);
    const int fields[][2] = {{3, 3}, {7, 9}, {10, 13}, {16, 16}};
    for (size_t i = 0; i < sizeof fields / sizeof fields[0]; ++i) phase(fields[i][0], fields[i][1]);
    /* transforms above 2048 a side: Bluestein (2050) and radix-2 (4096) lengths, both axes */
    const int ffts[][2] = {{3, 2050}, {2050, 2}, {2, 4096}};
    for (size_t i = 0; i < sizeof ffts / sizeof ffts[0]; ++i) {
        const int64_t n = (int64_t)ffts[i][0] * ffts[i][1];
        double *re = vec(n, -1, 1), *im = vec(n, -1, 1);
        oc_fft2d(re, im, ffts[i][0], ffts[i][1], 0);
        oc_fft2d(re, im, ffts[i][0], ffts[i][1], 1);
        free(re);
        free(im);
    }
    puts("oracle sanitize driver: done");
    return 0;
}
