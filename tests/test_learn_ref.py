"""CPU: the reference network used to check the device network is itself
pinned: its backward pass against central finite differences (float64 mode,
one random direction per layer) and the energy loss against its closed form."""
import importlib.util
import os

import numpy as np
import pytest

_spec = importlib.util.spec_from_file_location(
    "unet_ref", os.path.join(os.path.dirname(__file__), "..", "oracle", "unet_ref.py"))
U = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(U)


def _batch(rng, B, H, W):
    in5 = rng.normal(size=(B, H, W, 5))
    d = rng.uniform(1.0, 2.0, (B, H, W))
    cx = rng.uniform(0.0, 0.4, (B, H, W))
    cy = rng.uniform(0.0, 0.4, (B, H, W))
    cx[:, :, -1] = 0.0
    cy[:, -1, :] = 0.0
    op = np.stack([d, cx, cy], axis=1)
    return in5, rng.normal(size=(B, H, W)), op


@pytest.mark.parametrize("layer", list(range(15)))
def test_reference_backward_matches_finite_differences(layer):
    rng = np.random.default_rng(7 + layer)
    p = U.he_init(3)
    in5, rhat, op = _batch(rng, 2, 8, 8)
    m = U.UNetRef(p, use_bf16=False)
    e, acts = m.forward(in5)
    _, dout = m.loss(e, rhat, op)
    g = m.backward(dout, acts)
    off, n = U.offsets()
    hi = off[layer + 1] if layer < 14 else n
    v = np.zeros(n)
    v[off[layer]:hi] = rng.normal(size=hi - off[layer])
    eps = 1e-6
    mp, mm = U.UNetRef(p + eps * v, use_bf16=False), U.UNetRef(p - eps * v, use_bf16=False)
    fd = (mp.loss(mp.forward(in5)[0], rhat, op)[0] - mm.loss(mm.forward(in5)[0], rhat, op)[0]) / (2 * eps)
    an = float(g @ v)
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (layer, fd, an)


def test_energy_loss_closed_form():
    """L(e) = (e.A e - 2 e.r)/HW is minimal at e = A^-1 r with value -r.A^-1 r/HW,
    and its gradient vanishes there (A assembled densely from the same stencil)."""
    rng = np.random.default_rng(2)
    _, rhat, op = _batch(rng, 1, 6, 5)
    H, W = 6, 5
    n = H * W
    A = np.zeros((n, n))
    for k in range(n):
        e = np.zeros((1, H, W))
        e.flat[k] = 1.0
        A[:, k] = U.UNetRef.apply_op(e, op).ravel()
    assert np.allclose(A, A.T)
    x = np.linalg.solve(A, rhat.ravel()).reshape(1, H, W)
    m = U.UNetRef(U.he_init(0), use_bf16=False)
    L, g = m.loss(x, rhat, op)
    assert L == pytest.approx(-float(rhat.ravel() @ x.ravel()) / n, rel=1e-12)
    assert np.abs(g).max() < 1e-12
    L2, _ = m.loss(x + 0.01 * rng.normal(size=x.shape), rhat, op)
    assert L2 > L


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], dtype=np.float32)
    r = U.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 2 ** -8 * 3.0e-3
