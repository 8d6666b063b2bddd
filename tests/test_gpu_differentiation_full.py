"""The differentiation workload at its full size (512x512x128x128 uint16, 8 GiB), through size-independent
properties and sampled checks against the oracle.

- forward: a sample of positions against oc_vi_bilinear (a zero pattern with a balancing origin keeps the
  sample's origin mean equal to the whole map's), within 1e-6 of sum |x w|;
- mask adjoint: the adjoint identity <dY, Y(V)> = <dL/dmasks(dY), V> for random masks V (Y is linear in the
  masks), within 1e-6 of sum |dY| |Y(|V|)|;
- origin adjoint: central differences of the GPU forward (h = 1e-3) within 1e-3 of the scale, both for the pass
  re-reading the cube and for the adjoint taken from the forward's kept Jacobian rows (which agree within 1e-6);
- DPC: 20 fixed-step iterations on the 512x512 field (a 1024x1024 grid) and their adjoint against the oracle's
  composition, within 1e-10 of the largest value.
"""
import numpy as np
import pytest

import cubokit_amd as C
from cubokit_amd import synthetic
from cubokit_amd.phase import DPC
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def workload():
    shape, dtype, params = synthetic.differentiation_u16()
    cube = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    cube.fill_synthetic(params)
    fl = C.first_look(cube, [], com=True)
    yield cube, fl.com.row.values.copy(), fl.com.column.values.copy()
    cube.release()


def test_forward_on_sampled_positions(workload):
    cube, orow, ocol = workload
    masks = synthetic.differentiation_masks()
    y = C.ShiftedStack(cube, masks).forward(orow, ocol)
    n = orow.size
    rng = np.random.default_rng(1)
    idx = np.sort(rng.choice(n, 48, replace=False))
    pats = np.stack([cube.read_pattern(int(i) // 512, int(i) % 512) for i in idx])   # one scan row each
    sample = np.concatenate([pats, np.zeros((1,) + pats.shape[1:], pats.dtype)]).reshape(1, 49, 128, 128)
    mr, mc = O.pairwise_sum(orow) / n, O.pairwise_sum(ocol) / n
    srow = np.append(orow[idx], 49 * mr - O.pairwise_sum(orow[idx]))
    scol = np.append(ocol[idx], 49 * mc - O.pairwise_sum(ocol[idx]))
    for k in (0, 7, 15, 16, 17, 18):
        ref = O.vi_bilinear(sample, masks[k].astype(np.float64), srow, scol)[:48]
        cond = O.vi_bilinear(sample, np.abs(masks[k]).astype(np.float64), srow, scol)[:48]
        err = np.max(np.abs(y[idx, k] - ref) / np.maximum(cond, 1e-300))
        assert err <= 1e-6, f"row {k}: {err:.3e}"


def test_mask_adjoint_identity(workload):
    cube, orow, ocol = workload
    rng = np.random.default_rng(2)
    V = rng.uniform(-1, 1, (19, 128 * 128)).astype(np.float32)
    dy = rng.normal(size=(orow.size, 19))
    st = C.ShiftedStack(cube, V)
    yv = st.forward(orow, ocol)
    dm, _, _ = st.backward(orow, ocol, dy, origin=False)
    lhs = float(np.sum(dy * yv))
    rhs = float(np.sum(dm * V.astype(np.float64)))
    scale = float(np.sum(np.abs(dy) * C.ShiftedStack(cube, np.abs(V)).forward(orow, ocol)))
    assert abs(lhs - rhs) <= 1e-6 * scale, f"{lhs} vs {rhs} (scale {scale:.3e})"


def test_origin_adjoint_against_central_differences(workload):
    cube, orow, ocol = workload
    masks = synthetic.differentiation_masks()
    rng = np.random.default_rng(3)
    dy = rng.normal(size=(orow.size, 19))
    st = C.ShiftedStack(cube, masks)
    _, dr, dc = st.backward(orow, ocol, dy, masks=False)
    vr, vc = rng.normal(size=orow.size), rng.normal(size=orow.size)
    h = 1e-3
    lp = float(np.sum(dy * st.forward(orow + h * vr, ocol + h * vc)))
    lm = float(np.sum(dy * st.forward(orow - h * vr, ocol - h * vc)))
    fd = (lp - lm) / (2 * h)
    an = float(np.dot(dr, vr) + np.dot(dc, vc))
    y0 = st.forward(orow, ocol)
    scale = float(np.sum(np.abs(dy) * np.abs(y0)))
    assert abs(fd - an) <= 1e-3 * scale, f"{fd} vs {an} (scale {scale:.3e})"
    # a forward that keeps its Jacobian rows (Y unchanged): the backward takes dL/dorigin from them (no pass over
    # the cube); the rows are added across flushes in FP32 there, in FP64 in the pass
    st.keep_jacobian(True)
    assert np.array_equal(st.forward(orow, ocol), y0)
    st.jacobian_uses()
    _, drk, dck = st.backward(orow, ocol, dy, masks=False)
    assert st.jacobian_uses() == 1
    ank = float(np.dot(drk, vr) + np.dot(dck, vc))
    assert abs(fd - ank) <= 1e-3 * scale, f"{fd} vs {ank} (scale {scale:.3e})"
    assert abs(ank - an) <= 1e-6 * scale, f"kept {ank} vs pass {an} (scale {scale:.3e})"


def test_dpc_chain_adjoint_on_the_full_field():
    """The differentiation workload's 20 fixed-step DPC iterations on the 512x512 field (a 1024x1024 padded grid) and
    their adjoint against the oracle's chain of steps and step VJPs."""
    R = C_ = 512
    n, step = 20, 0.5
    rng = np.random.default_rng(4)
    yy, xx = np.mgrid[0:R, 0:C_] / 512.0
    fx = (np.sin(2 * np.pi * 3 * yy) * np.cos(2 * np.pi * 2 * xx)).reshape(-1) + 0.01 * rng.normal(size=R * C_)
    fy = (np.cos(2 * np.pi * 3 * yy) * np.sin(2 * np.pi * 2 * xx)).reshape(-1) + 0.01 * rng.normal(size=R * C_)
    g_phase = rng.normal(size=R * C_)
    d = DPC(R, C_)
    Rp, Cp = d.padded_shape
    assert (Rp, Cp) == (1024, 1024)
    phase, errors, gfx, gfy = d.chain_backward(fx, fy, n, g_phase, step=step)
    phis = [np.zeros(Rp * Cp)]
    for _ in range(n):
        phis.append(O.dpc_step(phis[-1], fx, fy, R, C_, Rp, Cp, step=step)[0])
    want_phase = phis[-1].reshape(Rp, Cp)[:R, :C_].reshape(-1)
    assert np.max(np.abs(phase - want_phase)) <= 1e-10 * np.max(np.abs(want_phase))
    g = np.zeros(Rp * Cp)
    g.reshape(Rp, Cp)[:R, :C_] = g_phase.reshape(R, C_)
    wx, wy = np.zeros(R * C_), np.zeros(R * C_)
    for i in reversed(range(n)):
        g, ax, ay = O.dpc_step_vjp(phis[i], fx, fy, R, C_, Rp, Cp, g, 0.0, step=step)
        wx += ax
        wy += ay
    assert np.max(np.abs(gfx - wx)) <= 1e-10 * np.max(np.abs(wx))
    assert np.max(np.abs(gfy - wy)) <= 1e-10 * np.max(np.abs(wy))
