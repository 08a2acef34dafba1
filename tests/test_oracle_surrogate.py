"""CPU checks of the surrogate network oracle (oracle/surrogate.py): its
hand-written backward pass against central differences of its loss (FP64),
its bf16 rounding against torch's, and invariances of the input features."""

import numpy as np
import pytest

from oracle import surrogate as osur


def _random_case(P=256, seed=0):
    rng = np.random.default_rng(seed)
    off, n = osur.layout()
    flat = np.zeros(n)
    for k, (o, s) in off.items():
        fan = s[1] if len(s) == 2 else 1
        scale = 0.1 if k in ("Wo", "bo") else (np.sqrt(2.0 / fan) if len(s) == 2 else 0.05)
        flat[o : o + int(np.prod(s))] = rng.normal(0.0, scale, int(np.prod(s)))
    X0 = np.zeros((P, osur.NIN))
    X0[:, :78] = rng.uniform(-1, 1, (P, 78))
    mag = np.exp(-rng.uniform(0, 3, P))
    ph = rng.uniform(-2, 2, P)
    g = np.stack([mag * np.cos(ph), mag * np.sin(ph)], axis=1)
    return flat, X0, g


def test_backward_matches_central_differences():
    flat, X0, g = _random_case()
    L, grad, _ = osur.loss_grad(flat, X0, g, None, block=128, bf16_mode=False)
    off, _ = osur.layout()
    rng = np.random.default_rng(1)
    idx = []
    for k, (o, s) in off.items():
        sz = int(np.prod(s))
        cols = np.arange(sz)
        if k == "W0":
            cols = cols[(cols % osur.NIN) < 78]  # padded inputs carry no gradient
        idx += list(o + rng.choice(cols, min(3, cols.size), replace=False))
    num = osur.numeric_gradient(flat, X0, g, idx, h=1e-5)
    ana = grad[idx]
    assert np.allclose(ana, num, rtol=2e-4, atol=1e-9 * np.abs(ana).max())


def test_backward_catches_a_transposed_weight():
    """A planted mistake (W3^T instead of W3 in the backward product) is
    caught by the same central-difference check."""
    flat, X0, g = _random_case(seed=3)
    p = osur.unpack(flat)
    phi, xs, zs = osur.forward(p, X0, False)
    _, dphi = osur.loss_and_dphi(phi, g, 128)
    bad = dict(p)
    bad["W3"] = p["W3"].T.copy()
    grads = {}
    dz = (dphi @ p["Wo"]) * osur.dsilu(zs[4])
    for l in range(4, -1, -1):
        grads[f"W{l}"] = dz.T @ xs[l]
        if l > 0:
            dz = (dz @ bad[f"W{l}"]) * osur.dsilu(zs[l - 1])
    off, _ = osur.layout()
    o = off["W1"][0]
    idx = [o + 5, o + 3000, o + 40000]
    num = osur.numeric_gradient(flat, X0, g, idx, h=1e-5)
    assert not np.allclose(grads["W1"].ravel()[[5, 3000, 40000]], num, rtol=2e-4)


def test_bf16_rounding_matches_torch():
    import torch

    rng = np.random.default_rng(4)
    x = np.concatenate([rng.normal(0, 1, 10000), rng.normal(0, 1e-3, 1000), [0.0, -0.0, 1.0, 3.0e38]])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(osur.bf16(x), ref)


def test_features_depend_on_the_wind_frame_only():
    """Rotating wavenumber and wind together, or mirroring the crosswind
    component, leaves the inputs unchanged; sin^2 + cos^2 = 1."""
    rng = np.random.default_rng(5)
    static = np.zeros(256)
    static[:128] = rng.normal(0, 2, 128)
    static[128:132] = [0.05, 0.05, 0.2, 0.2]
    static[140:148] = 1.0
    P = 500
    lx, ly = rng.uniform(-0.3, 0.3, P), rng.uniform(-0.3, 0.3, P)
    th = rng.uniform(0, 2 * np.pi)
    sm = np.zeros((P, 76))
    sm[:, 8], sm[:, 9], sm[:, 10], sm[:, 11] = np.cos(0.3), np.sin(0.3), 15.0, 5
    sm[:, 12:22] = rng.uniform(0.1, 5.0, 10)
    X = osur.features(static, sm, lx, ly)
    sm2 = sm.copy()
    sm2[:, 8], sm2[:, 9] = np.cos(0.3 + th), np.sin(0.3 + th)
    c, s = np.cos(th), np.sin(th)
    X2 = osur.features(static, sm2, c * lx - s * ly, s * lx + c * ly)
    assert np.allclose(X, X2, atol=1e-9)
    # mirror about the wind axis
    w = np.array([np.cos(0.3), np.sin(0.3)])
    k = np.stack([lx, ly], 1)
    km = 2 * (k @ w)[:, None] * w[None, :] - k
    X3 = osur.features(static, sm, km[:, 0], km[:, 1])
    assert np.allclose(X, X3, atol=1e-9)
    assert np.allclose(X[:, :32] ** 2 + X[:, 32:64] ** 2, 1.0)


def test_phi0_closed_forms():
    """The prior reduces to h |k| for no wind, to (1+i) h sqrt(kpar w / 2)
    for k along the wind with |k|^2 negligible, and is zero at k = 0."""
    sm = np.zeros((3, 76))
    sm[:, 8], sm[:, 9], sm[:, 11] = 1.0, 0.0, 2
    sm[:, 12:16] = [2.0, 0.0, 3.0, 0.0]  # h = 2, 3; w = 0
    lx, ly = np.array([0.1, 0.0, 0.0]), np.array([0.2, 0.3, 0.0])
    p = osur.phi0(sm, lx, ly)
    assert np.allclose(p[:, 0], 5.0 * np.hypot(lx, ly)) and np.allclose(p[:, 1], 0.0)
    sm[:, 12:16] = [2.0, 1e6, 3.0, 4e6]
    p = osur.phi0(sm[:1], np.array([1e-4]), np.array([0.0]))
    exp = 2.0 * np.sqrt(1j * 1e-4 * 1e6) + 3.0 * np.sqrt(1j * 1e-4 * 4e6)
    assert np.allclose(p[0, 0] + 1j * p[0, 1], exp, rtol=1e-6)
