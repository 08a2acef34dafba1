"""Oracle: the surrogate network in numpy FP64 (test infrastructure only).

Restates the network the HIP surrogate kernels evaluate and train
(input features, five SiLU layers of width 256, linear output Phi = -log G,
the training loss and its gradient) as plain matrix algebra, so the tests
can check the kernels' forward pass and hand-written backward pass.  There
is no reference counterpart: the reference has no learned component.  The
network maps the flux transfer of the reference's footprint solve
(src/bldfm/solver.py:220-235, 307-374), which ``oracle.solver`` restates.

With ``bf16=True`` the kernels' storage roundings are applied at the same
places (bf16 weights, inputs, hidden activations, stored pre-activations
and gradients); with ``bf16=False`` everything is FP64, and
``numeric_gradient`` checks that backward pass by central differences.
"""

import numpy as np

W, NIN, NL = 256, 96, 5
N_RFF, N_COORD, N_MET = 32, 4, 8


def layout():
    """Offsets of the flat parameter vector: W_l (row-major [out][in]), b_l, Wo, bo."""
    off, o = {}, 0
    for l in range(NL):
        kin = NIN if l == 0 else W
        off[f"W{l}"] = (o, (W, kin))
        o += W * kin
        off[f"b{l}"] = (o, (W,))
        o += W
    off["Wo"] = (o, (2, W))
    o += 2 * W
    off["bo"] = (o, (2,))
    o += 2
    return off, o


def unpack(flat):
    off, n = layout()
    assert flat.size >= n
    return {k: np.asarray(flat[o : o + int(np.prod(s))], dtype=np.float64).reshape(s) for k, (o, s) in off.items()}


def pack(parts):
    off, n = layout()
    flat = np.zeros(n)
    for k, (o, s) in off.items():
        flat[o : o + int(np.prod(s))] = np.asarray(parts[k]).ravel()
    return flat


def bf16(x):
    """Round to bfloat16 (round to nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s + z * s * (1.0 - s)


def phi0(surmet, lx, ly):
    """WKB prior of -log G: sum_j h_j sqrt(|k|^2 + i kpar w_j) over the
    met sample's intervals below zm (principal root), (P, 2)."""
    sm = np.asarray(surmet, dtype=np.float64)
    wx, wy = sm[:, 8], sm[:, 9]
    nw = sm[:, 11].astype(int)
    tab = sm[:, 12:].reshape(-1, 32, 2)
    k2 = lx * lx + ly * ly
    kpar = lx * wx + ly * wy
    out = np.zeros(lx.size, dtype=np.complex128)
    for j in range(32):
        lam = np.sqrt(k2 + 1j * kpar * tab[:, j, 1])
        out += np.where(j < nw, tab[:, j, 0] * lam, 0.0)
    return np.stack([out.real, out.imag], axis=1)


def features(static, surmet, lx, ly):
    """Network inputs (P, 96): RFF of the wind-frame, zm-scaled wavenumber
    coordinates, the coordinates, the normalised met values, the prior
    Phi0 / 8, zero padding.  static: the 256 floats of the network's fixed
    header; surmet: the 76 network-input floats of each point's met sample."""
    static = np.asarray(static, dtype=np.float64)
    B = static[: N_RFF * N_COORD].reshape(N_RFF, N_COORD)
    cs, mu, sd = static[128:132], static[132:140], static[140:148]
    sm = np.asarray(surmet, dtype=np.float64)
    wx, wy, zm = sm[:, 8], sm[:, 9], sm[:, 10]
    kpar = (lx * wx + ly * wy) * zm
    kperp = np.abs(ly * wx - lx * wy) * zm
    c = np.stack([kpar * cs[0], kperp * cs[1], np.sign(kpar) * np.sqrt(np.abs(kpar)) * cs[2],
                  np.sqrt(kperp) * cs[3]], axis=1)
    th = 2.0 * np.pi * c @ B.T
    X = np.zeros((lx.size, NIN))
    X[:, :N_RFF] = np.sin(th)
    X[:, N_RFF : 2 * N_RFF] = np.cos(th)
    X[:, 2 * N_RFF : 2 * N_RFF + N_COORD] = c
    X[:, 2 * N_RFF + N_COORD : 2 * N_RFF + N_COORD + N_MET] = (sm[:, :N_MET] - mu) / sd
    X[:, 2 * N_RFF + N_COORD + N_MET : 2 * N_RFF + N_COORD + N_MET + 2] = phi0(sm, lx, ly) / 8.0
    return X


def forward(p, X0, bf16_mode=True):
    """Network output (P, 2; Phi = Phi0 + output) and the stored
    activations of the forward pass."""
    r = bf16 if bf16_mode else (lambda v: v)
    xs, zs = [r(X0)], []
    x = xs[0]
    for l in range(NL):
        z = x @ r(p[f"W{l}"]).T + p[f"b{l}"]
        zs.append(z)
        x_full = silu(z)
        xs.append(r(x_full))
        x = xs[-1]
    phi = x_full @ p["Wo"].T + p["bo"]
    return phi, xs, zs


def loss_and_dphi(phi, g, block=128, p0=None):
    """The training loss of the kernels and dLoss/dPhi.  Per block of
    ``block`` points (one met sample) with S = sum |G|^2:
    sum_i |G_i - exp(-Phi_i)|^2 + |G_i|^2 |Phi_i + log G_i|^2 (imaginary
    part wrapped to [-pi, pi]), divided by S and the number of blocks;
    points with |G| > 1.5, a non-finite G, or (given the prior p0)
    |G| > 100 exp(-Re p0) + 1e-8 are left out."""
    P = phi.shape[0]
    nb = P // block
    G = g[:, 0] + 1j * g[:, 1]
    valid = np.abs(G) ** 2 <= 2.25  # False for NaN
    if p0 is not None:  # roundoff, not physics: far above the WKB prior
        valid &= np.abs(G) <= 100.0 * np.exp(-p0[:, 0]) + 1e-8
    G = np.where(valid, G, 0.0)
    s2 = np.abs(G) ** 2
    mag = np.exp(-np.maximum(phi[:, 0], -20.0))
    Gp = mag * np.exp(-1j * phi[:, 1])
    e = Gp - G
    dre = -2.0 * np.real(np.conj(e) * Gp)
    dim = 2.0 * np.imag(np.conj(e) * Gp)
    li = np.abs(e) ** 2
    pos = s2 > 1e-30
    tr = np.where(pos, -0.5 * np.log(np.where(pos, s2, 1.0)), 0.0)
    ti = np.where(pos, -np.angle(G), 0.0)
    dr = phi[:, 0] - tr
    di = phi[:, 1] - ti
    di = di - 2.0 * np.pi * np.rint(di / (2.0 * np.pi))
    li = li + np.where(pos, s2 * (dr**2 + di**2), 0.0)
    dre = dre + np.where(pos, 2.0 * s2 * dr, 0.0)
    dim = dim + np.where(pos, 2.0 * s2 * di, 0.0)
    S = s2.reshape(nb, block).sum(axis=1)
    sc = np.repeat(1.0 / ((S + 1e-30) * nb), block) * valid
    return float(np.sum(li * sc)), np.stack([dre * sc, dim * sc], axis=1)


def backward(p, xs, zs, dphi, bf16_mode=True):
    """Gradient of the loss (dict like ``unpack``) from dLoss/dPhi."""
    r = bf16 if bf16_mode else (lambda v: v)
    g = {"Wo": dphi.T @ xs[NL], "bo": dphi.sum(axis=0)}
    dz = r((dphi @ p["Wo"]) * dsilu(r(zs[NL - 1])))
    for l in range(NL - 1, -1, -1):
        g[f"W{l}"] = dz.T @ xs[l]
        g[f"b{l}"] = dz.sum(axis=0)
        if l > 0:
            dz = r((dz @ r(p[f"W{l}"])) * dsilu(r(zs[l - 1])))
    return g


def loss_grad(flat, X0, g, p0=None, block=128, bf16_mode=True):
    """Loss, gradient (flat) and Phi = p0 + network output."""
    p = unpack(flat)
    out, xs, zs = forward(p, X0, bf16_mode)
    phi = out + (0.0 if p0 is None else p0)
    L, dphi = loss_and_dphi(phi, g, block, p0)
    return L, pack(backward(p, xs, zs, dphi, bf16_mode)), phi


def numeric_gradient(flat, X0, g, idx, p0=None, h=1e-6, block=128):
    """Central differences of the FP64 loss at parameter indices idx."""
    out = []
    for i in idx:
        a = flat.copy()
        b = flat.copy()
        a[i] += h
        b[i] -= h
        La, _, _ = loss_grad(a, X0, g, p0, block, bf16_mode=False)
        Lb, _, _ = loss_grad(b, X0, g, p0, block, bf16_mode=False)
        out.append((La - Lb) / (2 * h))
    return np.asarray(out)
