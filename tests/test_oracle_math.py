"""Pins for the oracle functions that CuboKit's fixtures do not cover (CPU only).

Each check is independent of the formula it pins: an independent library routine (numpy.fft, numpy's
own pairwise sum, float64 BLAS), a closed form (zero and integer shifts), an adjoint identity
<J v, u> = <v, J^T u>, or central finite differences, so that a plausible mistake (a dropped term, a
wrong sign, a transposed operand) fails at least one of them.
"""
import numpy as np

from oracle import oracle as O


def _rng(seed=1):
    return np.random.default_rng(seed)


def test_pairwise_sum_is_numpys():
    r = _rng(2)
    for n in [0, 1, 7, 8, 9, 127, 128, 129, 1000, 4097]:
        a = r.standard_normal(n) * 10 ** r.uniform(-3, 3, n)
        assert O.pairwise_sum(a) == float(np.sum(a))


def test_fft2_matches_numpy_both_directions():
    r = _rng(3)
    # the last three go through Bluestein's chirp convolution (lengths above 2048 that are not powers of two) and the
    # radix-2 transform above 2048
    for shape in [(8, 16), (6, 10), (48, 40), (1, 5), (3, 2050), (2, 3001), (2101, 3), (2, 4096)]:
        a = r.standard_normal(shape) + 1j * r.standard_normal(shape)
        re, im = O.fft2(a.real, a.imag)
        ref = np.fft.fft2(a)
        assert np.max(np.abs(re + 1j * im - ref)) <= 1e-11 * np.max(np.abs(ref))
        re, im = O.fft2(a.real, a.imag, inverse=True)
        ref = np.fft.ifft2(a)
        assert np.max(np.abs(re + 1j * im - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_mask_stack_equals_blas_and_single_mask_images():
    r = _rng(4)
    cube = r.poisson(50, (3, 4, 6, 8)).astype(np.uint16)
    w = r.uniform(-1, 1, (5, 48))
    y = O.mask_stack(cube, w)
    ref = cube.reshape(12, 48).astype(np.float64) @ w.T
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))
    m = (r.uniform(size=48) > 0.5).astype(np.float32)
    assert np.array_equal(O.mask_stack(cube, m[None].astype(np.float64))[:, 0], O.virtual_image(cube, m))


def test_mask_stack_grad_is_the_adjoint():
    r = _rng(5)
    cube = r.uniform(0, 100, (2, 3, 5, 7)).astype(np.float32)
    w = r.standard_normal((4, 35))
    dy = r.standard_normal((6, 4))
    dw = O.mask_stack_grad_w(cube, dy)
    lhs = np.sum(dy * O.mask_stack(cube, w))
    rhs = np.sum(dw * w)
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


def _origins(r, n, spread=2.0):
    return r.uniform(-spread, spread, n) + 10.0, r.uniform(-spread, spread, n) - 3.0


def test_bilinear_zero_and_integer_shifts():
    r = _rng(6)
    cube = r.uniform(0, 10, (2, 3, 6, 7)).astype(np.float32)
    x = cube.reshape(6, 42).astype(np.float64)
    m = r.uniform(0, 1, 42)
    same = np.full(6, 3.25)
    got = O.vi_bilinear(cube, m, same, same)
    assert np.max(np.abs(got - x @ m)) <= 1e-13 * np.max(np.abs(got))
    # integer offsets from the mean roll the mask exactly (np.roll semantics)
    dr = np.array([0, 1, -2, 3, 0, -2], np.float64)
    dc = np.array([1, -1, 0, 2, -3, 1], np.float64)
    got = O.vi_bilinear(cube, m, dr + 10.0, dc - 4.0)
    for p in range(6):
        rolled = np.roll(m.reshape(6, 7), (int(dr[p]), int(dc[p])), axis=(0, 1)).reshape(-1)
        assert abs(got[p] - x[p] @ rolled) <= 1e-12 * abs(got[p])


def test_bilinear_is_affine_inside_a_cell():
    r = _rng(7)
    cube = r.uniform(0, 10, (1, 3, 6, 7)).astype(np.float32)
    m = r.uniform(0, 1, 42)
    # position 0 sits at dr = 2t/3 in [0, 1) for these t: Y_0 is affine in t there
    vals = [O.vi_bilinear(cube, m, np.array([5.0 + t, 5.0, 5.0]), np.zeros(3))[0] for t in (0.1, 0.4, 0.7)]
    assert abs((vals[1] - vals[0]) - (vals[2] - vals[1])) <= 1e-11 * abs(vals[0])
    # and it is not constant (the shift reaches the image)
    assert abs(vals[1] - vals[0]) > 1e-6 * abs(vals[0])


def test_bilinear_grad_w_is_the_adjoint():
    r = _rng(8)
    cube = r.uniform(0, 10, (2, 3, 6, 7)).astype(np.float32)
    orr, occ = _origins(r, 6)
    m = r.uniform(0, 1, 42)
    dy = r.standard_normal(6)
    dw = O.vi_bilinear_grad_w(cube, orr, occ, dy)
    lhs = np.sum(dy * O.vi_bilinear(cube, m, orr, occ))
    assert abs(lhs - np.sum(dw * m)) <= 1e-11 * abs(lhs)


def test_bilinear_grad_origin_matches_finite_differences():
    r = _rng(9)
    cube = r.uniform(0, 10, (2, 2, 8, 8)).astype(np.float32)
    m = r.uniform(0, 1, 64)
    orr = np.array([0.3, 0.45, 0.62, 0.2]) + 4
    occ = np.array([0.55, 0.3, 0.25, 0.7]) - 1
    dy = r.standard_normal(4)
    gr, gc = O.vi_bilinear_grad_origin(cube, m, orr, occ, dy)
    h = 1e-6
    for p in range(4):
        e = np.zeros(4)
        e[p] = h
        fd_r = (dy @ O.vi_bilinear(cube, m, orr + e, occ) - dy @ O.vi_bilinear(cube, m, orr - e, occ)) / (2 * h)
        fd_c = (dy @ O.vi_bilinear(cube, m, orr, occ + e) - dy @ O.vi_bilinear(cube, m, orr, occ - e)) / (2 * h)
        assert abs(gr[p] - fd_r) <= 1e-6 * (abs(fd_r) + 1)
        assert abs(gc[p] - fd_c) <= 1e-6 * (abs(fd_c) + 1)


def _field(r, R, C):
    return r.standard_normal(R * C), r.standard_normal(R * C)


def test_dpc_step_is_the_first_reconstruction_iteration():
    r = _rng(10)
    R, C = 6, 5
    fx, fy = _field(r, R, C)
    phi0 = np.zeros(4 * R * C)
    phi1, err = O.dpc_step(phi0, fx, fy, R, C, 2 * R, 2 * C)
    rec = O.dpc_reconstruct(fx, fy, R, C, max_iter=1)
    assert rec["errors"][0] == err
    assert np.array_equal(phi1.reshape(2 * R, 2 * C)[:R, :C].reshape(-1), rec["phase"])


def test_dpc_step_vjp_is_the_adjoint_and_matches_finite_differences():
    r = _rng(11)
    R, C, Rp, Cp = 5, 6, 10, 12
    fx, fy = _field(r, R, C)
    phi = r.standard_normal(Rp * Cp)
    g = r.standard_normal(Rp * Cp)
    g_err = 0.7
    gp, gfx, gfy = O.dpc_step_vjp(phi, fx, fy, R, C, Rp, Cp, g, g_err)

    def loss(ph, a, b):
        out, e = O.dpc_step(ph, a, b, R, C, Rp, Cp)
        return g @ out + g_err * e

    h = 1e-6
    for _ in range(6):
        v, va, vb = r.standard_normal(Rp * Cp), r.standard_normal(R * C), r.standard_normal(R * C)
        fd = (loss(phi + h * v, fx + h * va, fy + h * vb) - loss(phi - h * v, fx - h * va, fy - h * vb)) / (2 * h)
        an = gp @ v + gfx @ va + gfy @ vb
        assert abs(fd - an) <= 1e-6 * (abs(an) + 1), (fd, an)
