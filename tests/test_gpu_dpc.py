"""GPU parity of iterative DPC (Sources/CuboPhase/DPC.swift:26-102) and of its adjoint.

- reconstruct against CuboKit's golden chains (com-dpc-pins.json: the rotation-corrected fields of 4 cubes and
  3 planted fields): the backtracking sequence and final step exact, errors within 1e-9 relative, the phase
  within 1e-8 of its maximum (CuboPhaseTests.swift:16-139's bars);
- one step and its vector-Jacobian product against the oracle (oc_dpc_step, oc_dpc_step_vjp), on power-of-two
  and other grid sizes: within 1e-11 of the largest value;
- a 5-step chain's adjoint against the oracle's composition of step adjoints (1e-10 of the largest value) and
  against central differences (h = 1e-3) of the oracle's forward chain (1e-5 relative: the error terms are
  rational in the field, so the central difference is exact only to second order).
"""
import numpy as np
import pytest

from cubokit_amd.phase import DPC
from cubokit_amd.virtual import CuboVirtualError
from oracle import oracle as O
from tests.support import com_pins, doubles, load_cube, near, of_max

pytestmark = pytest.mark.gpu


def _chains():
    pins = com_pins()
    for e in pins["com"]:
        if "chain" in e:
            R, C = load_cube(e["cube"]).shape[:2]
            yield e["cube"], R, C, e["chain"]
    for f in pins["fields"]:
        R, C = f["shape"]
        yield f"field {f['theta']}", R, C, f


def test_reconstruct_matches_the_golden_chains():
    n = 0
    for name, R, C, ch in _chains():
        got = DPC(R, C).reconstruct(doubles(ch["corrected_x"]), doubles(ch["corrected_y"]))
        assert got["accepted"] == ch["accepted"], f"{name}: backtracking sequence"
        assert got["final_step"] == float(ch["final_step"]), f"{name}: final step"
        near(got["errors"], doubles(ch["errors"]), 1e-9, f"{name} errors")
        of_max(got["phase"], doubles(ch["phase"]), 1e-8, f"{name} phase")
        n += 1
    assert n == 7


def _field(R, C, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=R * C), rng.normal(size=R * C)


@pytest.mark.parametrize("R,C", [(32, 16), (7, 9), (40, 24), (128, 64), (1024, 64)])
def test_step_and_its_adjoint_match_the_oracle(R, C):
    """One step and its adjoint against the oracle. Grids 64x32, 14x18 and 80x48 take the separate element-wise
    kernels; 256x128 and 2048x128 the passes with them fused in (128-point rows: a radix-2 stage first, so the
    per-stage twiddle tables start at span 2; 2048-point rows: the global twiddle table)."""
    d = DPC(R, C, sampling=(0.7, 1.3))
    Rp, Cp = d.padded_shape
    fx, fy = _field(R, C, R * C)
    rng = np.random.default_rng(R + C)
    phi = rng.normal(size=Rp * Cp)
    out, err = d.step(phi, fx, fy, step=0.4)
    ref, ref_err = O.dpc_step(phi, fx, fy, R, C, Rp, Cp, sampling=(0.7, 1.3), step=0.4)
    assert np.max(np.abs(out - ref)) <= 1e-11 * np.max(np.abs(ref)), f"{R}x{C} phase"
    assert abs(err - ref_err) <= 1e-12 * abs(ref_err)
    g_out, g_err = rng.normal(size=Rp * Cp), 0.8
    g = d.step_backward(phi, fx, fy, g_out, g_err, step=0.4)
    want = O.dpc_step_vjp(phi, fx, fy, R, C, Rp, Cp, g_out, g_err, sampling=(0.7, 1.3), step=0.4)
    for a, b, what in zip(g, want, ("phi", "fx", "fy")):
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b)), f"{R}x{C} dL/d{what}"


def _oracle_chain(fx, fy, R, C, Rp, Cp, n, step, sampling=(1.0, 1.0)):
    phis, errs = [np.zeros(Rp * Cp)], []
    for _ in range(n):
        p, e = O.dpc_step(phis[-1], fx, fy, R, C, Rp, Cp, sampling=sampling, step=step)
        phis.append(p)
        errs.append(e)
    return phis, np.array(errs)


def _crop(p, R, C, Cp):
    return p.reshape(-1, Cp)[:R, :C].reshape(-1)


def test_chain_adjoint_against_the_oracle_and_central_differences():
    R, C, n, step = 10, 12, 5, 0.5
    d = DPC(R, C)
    Rp, Cp = d.padded_shape
    fx, fy = _field(R, C, 3)
    rng = np.random.default_rng(4)
    g_phase, g_err = rng.normal(size=R * C), rng.normal(size=n)
    phase, errors, gfx, gfy = d.chain_backward(fx, fy, n, g_phase, g_err, step=step)

    phis, errs = _oracle_chain(fx, fy, R, C, Rp, Cp, n, step)
    assert np.max(np.abs(phase - _crop(phis[-1], R, C, Cp))) <= 1e-11 * np.max(np.abs(phis[-1]))
    assert np.max(np.abs(errors - errs)) <= 1e-12 * np.max(errs)
    g = np.zeros(Rp * Cp)
    g.reshape(Rp, Cp)[:R, :C] = g_phase.reshape(R, C)
    wx, wy = np.zeros(R * C), np.zeros(R * C)
    for i in reversed(range(n)):
        g, ax, ay = O.dpc_step_vjp(phis[i], fx, fy, R, C, Rp, Cp, g, g_err[i], step=step)
        wx += ax
        wy += ay
    assert np.max(np.abs(gfx - wx)) <= 1e-10 * np.max(np.abs(wx))
    assert np.max(np.abs(gfy - wy)) <= 1e-10 * np.max(np.abs(wy))

    def loss(x, y):
        p, e = _oracle_chain(x, y, R, C, Rp, Cp, n, step)
        return float(np.dot(g_phase, _crop(p[-1], R, C, Cp)) + np.dot(g_err, e))

    h = 1e-3
    for trial in range(2):
        vx, vy = rng.normal(size=R * C), rng.normal(size=R * C)
        fd = (loss(fx + h * vx, fy + h * vy) - loss(fx - h * vx, fy - h * vy)) / (2 * h)
        an = float(np.dot(gfx, vx) + np.dot(gfy, vy))
        assert abs(fd - an) <= 1e-5 * (abs(an) + np.abs(gfx * vx).sum() + np.abs(gfy * vy).sum()), f"{fd} vs {an}"


def test_a_larger_grid_with_other_lengths():
    """128 x 96 scan: a 256 x 192 grid (192 is not a power of two: the direct-DFT rows)."""
    R, C = 128, 96
    d = DPC(R, C)
    assert d.padded_shape == (256, 192)
    fx, fy = _field(R, C, 5)
    got = d.reconstruct(fx, fy, max_iter=6, step=0.5)
    want = O.dpc_reconstruct(fx, fy, R, C, max_iter=6, step=0.5)
    assert got["accepted"] == want["accepted"]
    near(got["errors"], want["errors"], 1e-10, "errors")
    of_max(got["phase"], want["phase"], 1e-10, "phase")


def test_refusals():
    with pytest.raises(CuboVirtualError):
        DPC(3, 3).reconstruct(np.zeros(9), np.zeros(9))
    with pytest.raises(CuboVirtualError):
        DPC(20000, 20000)                              # a padded grid of 1.6e9 points (above 2^28)
    with pytest.raises(CuboVirtualError):
        DPC(3, 3).reconstruct(np.array([1.0] * 8 + [np.inf]), np.zeros(9))
    with pytest.raises(CuboVirtualError):
        DPC(3, 3, padding=0.5)


@pytest.mark.parametrize("R,C,padding", [(1100, 1025, 2.0), (1024, 600, 4.0), (1500, 10, 2.0)])
def test_grids_above_2048_a_side(R, C, padding):
    """DPC.swift:42-44 sets no size limit. Padded sides above 2048 run four-step transforms (powers of two) and
    Bluestein's chirp convolution (other lengths): grids 2200 x 2050 (both Bluestein), 4096 x 2400 (four-step rows,
    Bluestein columns) and 3000 x 20 (Bluestein rows, the LDS transform for the columns). One step and its adjoint
    against the oracle (whose transforms above 2048 are pinned to numpy.fft) within 1e-10 of the largest value, and
    three iterations of reconstruct."""
    print(f"case: DPC grid of a {R} x {C} scan, padding {padding}", flush=True)
    d = DPC(R, C, padding=padding, sampling=(0.9, 1.1))
    Rp, Cp = d.padded_shape
    assert (Rp, Cp) == (int(round(R * padding)), int(round(C * padding)))
    fx, fy = _field(R, C, 11)
    rng = np.random.default_rng(12)
    phi = rng.normal(size=Rp * Cp)
    out, err = d.step(phi, fx, fy, step=0.4)
    ref, ref_err = O.dpc_step(phi, fx, fy, R, C, Rp, Cp, sampling=(0.9, 1.1), step=0.4)
    assert np.max(np.abs(out - ref)) <= 1e-10 * np.max(np.abs(ref)), "phase"
    assert abs(err - ref_err) <= 1e-10 * abs(ref_err)
    g = rng.normal(size=Rp * Cp)
    gp, gx, gy = d.step_backward(phi, fx, fy, g, g_error=0.7, step=0.4)
    wp, wx, wy = O.dpc_step_vjp(phi, fx, fy, R, C, Rp, Cp, g, 0.7, sampling=(0.9, 1.1), step=0.4)
    for a, b, name in ((gp, wp, "phi"), (gx, wx, "fx"), (gy, wy, "fy")):
        assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(b)), f"adjoint {name}"
    got = d.reconstruct(fx, fy, max_iter=3, step=0.5)
    want = O.dpc_reconstruct(fx, fy, R, C, sampling=(0.9, 1.1), padding=padding, max_iter=3, step=0.5)
    assert got["accepted"] == want["accepted"]
    near(got["errors"], want["errors"], 1e-10, "errors")
    of_max(got["phase"], want["phase"], 1e-10, "phase")
