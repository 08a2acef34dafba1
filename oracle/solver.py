"""Oracle: steady-state advection-diffusion solver (numpy, FP64).

Test infrastructure only (see ``oracle/__init__.py``).  Restates
src/bldfm/solver.py:16-374 with explicit index maps instead of the
reference's fftshift/pad/ifftshift sequence:

* truncated index j along an axis with nl modes holds the frequency
  f_j = fftfreq order (j for j < ceil(nl/2), else j - nl), and that
  frequency lives at padded index f_j mod ne (solver.py:139-145, 265-278);
* the vertical step, the two initial-value problems, the shooting
  coefficient and the (0,0) trapezoid follow solver.py:190-251 and
  solver.py:334-374 operation by operation;
* the padded 2-D transforms use numpy's pocketfft with the reference's
  ``norm`` arguments (solver.py:136, 282-287).

``_mutation`` plants a known mistake (tests only) so the tests can show that
the reference fixtures detect it.
"""

import numpy as np

VALID_MUTATIONS = (
    None,
    "b_cubic_sign",  # + instead of - on the cubic term of b (solver.py:362)
    "upper_level",  # profile coefficients of level i+1 instead of i (solver.py:357)
    "shift_sign",  # e^{-i k.x} instead of e^{+i k.x} (solver.py:255)
    "top_bc_sign",  # q = -Kz lambda p at the top (solver.py:228-230)
    "halo_in_shift",  # drop the halo from the footprint shift (solver.py:255)
)


def fft_frequencies(nl):
    """Integer frequencies in numpy ``fftfreq`` order (j, then j - nl)."""
    j = np.arange(nl)
    return np.where(j < (nl + 1) // 2, j, j - nl)


def padded_index(nl, ne):
    """Padded-grid index of each truncated index: f_j mod ne."""
    return np.mod(fft_frequencies(nl), ne)


def wavenumbers(nl, d, ne):
    """2 pi / d / ne * fftfreq(nl, d=1/nl)  (solver.py:148-153)."""
    ilx = np.fft.fftfreq(nl, d=1.0 / nl)
    return 2.0 * np.pi / d / ne * ilx


def _levels_array(levels):
    return np.array([levels]) if np.ndim(levels) == 0 else levels


def _cmul(ar, ai, br, bi):
    """Complex product with every real product and sum rounded separately:
    (ar br - ai bi, ar bi + ai br), the arithmetic numba compiles (no
    fast-math, hence no fused multiply-add).  numpy's own complex multiply
    is not used here: its AVX-512 loop fuses, which numba does not."""
    return ar * br - ai * bi, ar * bi + ai * br


def ivp(p_init, q_init, profiles, z, levels, lx, ly, mutation=None):
    """March dp/dz = -q/Kz, dq/dz = T p from z[0] to z[-1] for every mode.

    Follows solver.py:334-374 as numba compiles it (the reference jit-compiles
    ivp_solver, src/bldfm/utils.py:95-106): the 2x2 step uses the profile
    values of the LOWER level i; literal integer powers are repeated
    multiplications (h**2 = h*h, h**3 = h*(h*h), Kzinv**2, Ti**2 = Ti*Ti);
    complex arithmetic is plain (each product and sum rounded).  Snapshots are
    taken before the step when ``i in levels`` and after the last step when
    ``nz-1 in levels``.  Returns the top state and the snapshots
    (nlvls, nmodes), unfilled slots zero.
    """
    u, v, Kx, Ky, Kz = profiles
    nz = len(z)
    dz = np.diff(z)
    nlvls = len(levels)
    pr, pi = np.array(p_init.real, dtype=np.float64), np.array(p_init.imag, dtype=np.float64)
    qr, qi = np.array(q_init.real, dtype=np.float64), np.array(q_init.imag, dtype=np.float64)
    snap_p = np.zeros((nlvls, pr.shape[0]), dtype=np.complex128)
    snap_q = np.zeros((nlvls, pr.shape[0]), dtype=np.complex128)
    lx2, ly2 = lx * lx, ly * ly
    lvl = 0
    for i in range(nz - 1):
        if i in levels:
            snap_p[lvl], snap_q[lvl] = pr + 1j * pi, qr + 1j * qi
            lvl += 1
        k = i + 1 if mutation == "upper_level" else i
        # Ti = -(Kx lx^2 + Ky ly^2) - 1j u lx - 1j v ly
        tr = -(Kx[k] * lx2 + Ky[k] * ly2)
        ti = -(u[k] * lx) - v[k] * ly
        kinv = 1.0 / Kz[k]
        h = dz[i]
        h2, h3 = h * h, h * (h * h)
        sa, sb, sc = 0.5 * kinv, 1.0 / 6.0 * (kinv * kinv), 1.0 / 6.0 * kinv
        # a = d = 1 - 0.5 Kzinv Ti dz^2
        ar, ai = 1.0 - sa * tr * h2, -(sa * ti * h2)
        # b = -Kzinv dz -/+ 1/6 Kzinv^2 Ti dz^3
        cr, ci = sb * tr * h3, sb * ti * h3
        if mutation == "b_cubic_sign":
            br, bi = -kinv * h + cr, ci
        else:
            br, bi = -kinv * h - cr, -ci
        # c = Ti dz - 1/6 Kzinv Ti^2 dz^3
        t2r, t2i = _cmul(tr, ti, tr, ti)
        cr, ci = tr * h - sc * t2r * h3, ti * h - sc * t2i * h3
        apr, api = _cmul(ar, ai, pr, pi)
        bqr, bqi = _cmul(br, bi, qr, qi)
        cpr, cpi = _cmul(cr, ci, pr, pi)
        dqr, dqi = _cmul(ar, ai, qr, qi)
        pr, pi, qr, qi = apr + bqr, api + bqi, cpr + dqr, cpi + dqi
    if nz - 1 in levels:
        snap_p[lvl], snap_q[lvl] = pr + 1j * pi, qr + 1j * qi
    return pr + 1j * pi, qr + 1j * qi, snap_p, snap_q


def steady_state_transport_solver(
    srf_flx,
    z,
    profiles,
    domain,
    levels,
    modes=(512, 512),
    meas_pt=(0.0, 0.0),
    srf_bg_conc=0.0,
    footprint=False,
    analytic=False,
    halo=None,
    precision="single",
    cache=None,
    *,
    _mutation=None,
    _return_spectra=False,
    _alpha_ulps=0,
):
    """CPU restatement of src/bldfm/solver.py:16-304.

    Returns ``((X, Y, Z), conc, flx)`` squeezed, like the reference.  With
    ``_return_spectra`` the truncated (nlvls, nly, nlx) spectra p^, q^ after
    the shift are returned as well.  ``_alpha_ulps`` (tests only) scales the
    shooting coefficient by 1 + k*2^-52 before the combination: the change of
    the output measures how strongly a level depends on the last bit of
    alpha (near the top of a tall 3-D field the growing and decaying
    solutions cancel, and that dependence exceeds any fixed tolerance).
    """
    assert _mutation in VALID_MUTATIONS
    q0 = np.asarray(srf_flx)
    u, v, Kx, Ky, Kz = profiles
    xmx, ymx = domain
    nlx, nly = modes
    xm, ym = meas_pt
    if (nlx % 2 > 0) or (nly % 2 > 0):
        raise ValueError("modes must consist of even numbers.")
    ny, nx = q0.shape
    nz = len(z)
    dx, dy = xmx / nx, ymx / ny
    dz = np.diff(z)
    levels = _levels_array(levels)
    nlvls = len(levels)
    if halo is None:
        halo = max(xmx, ymx)
    px, py = int(halo / dx), int(halo / dy)
    nxe, nye = nx + 2 * px, ny + 2 * py
    if (nlx > nxe) or (nly > nye):
        nlx, nly = nxe, nye
    # The untruncation pads the nl retained modes by (ne - nl)//2 on each side
    # (solver.py:265-278), so the inverse transforms have lx_t = nl + 2 dl
    # points: ne, or ne - 1 when ne - nl is odd.
    dlx, dly = (nxe - nlx) // 2, (nye - nly) // 2
    ltx, lty = nlx + 2 * dlx, nly + 2 * dly

    if footprint:
        tq0 = np.ones((nly, nlx), dtype=np.complex128) / nxe / nye
    else:
        qpad = np.zeros((nye, nxe))
        qpad[py : py + ny, px : px + nx] = q0
        fq = np.fft.fft2(qpad, norm="forward")
        if ltx == nxe and lty == nye:
            tq0 = fq[np.ix_(padded_index(nly, nye), padded_index(nlx, nxe))]
        else:
            # odd ne - nl: the reference's fftshift/slice/ifftshift keeps
            # nl + 1 modes (solver.py:138-145); its mode mask below fails on it
            tq0 = np.fft.ifftshift(np.fft.fftshift(fq)[dly : nye - dly, dlx : nxe - dlx])
    iy = padded_index(nly, lty)
    ix = padded_index(nlx, ltx)

    lx = wavenumbers(nlx, dx, nxe)
    ly = wavenumbers(nly, dy, nye)
    Lx, Ly = np.meshgrid(lx, ly)
    msk = np.ones((nly, nlx), dtype=bool)
    msk[0, 0] = False

    kinv_t = 1.0 / Kz[nz - 1]
    eig = np.sqrt(
        Kx[nz - 1] * kinv_t * Lx[msk] ** 2
        + Ky[nz - 1] * kinv_t * Ly[msk] ** 2
        + 1j * u[nz - 1] * kinv_t * Lx[msk]
        + 1j * v[nz - 1] * kinv_t * Ly[msk]
    )

    if precision == "single":
        cdt = np.complex64
    elif precision == "double":
        cdt = np.complex128
    else:
        raise ValueError("precision must be single (default) or double.")
    tp = np.zeros((nlvls, nly, nlx), dtype=cdt)
    tq = np.zeros((nlvls, nly, nlx), dtype=cdt)
    tp[0, 0, 0] = srf_bg_conc
    tq[:, 0, 0] = tq0[0, 0]

    if analytic:
        h = z[levels] - z[0]
        tp[:, 0, 0] = srf_bg_conc - tq0[0, 0] * kinv_t * h
        tq[:, msk] = tq0[msk] * np.exp(-eig * h)  # IndexError for odd ne - nl (solver.py:197)
        tp[:, msk] = tq[:, msk] * kinv_t / eig
    else:
        one = np.ones(msk.sum(), dtype=np.complex128)
        zero = np.zeros(msk.sum(), dtype=np.complex128)
        p1, q1, sp1, sq1 = ivp(one, zero, profiles, z, levels, Lx[msk], Ly[msk], _mutation)
        p2, q2, sp2, sq2 = ivp(zero, tq0[msk], profiles, z, levels, Lx[msk], Ly[msk], _mutation)
        kzt = Kz[nz - 1]
        if _mutation == "top_bc_sign":
            alpha = -(q2 + kzt * eig * p2) / (q1 + kzt * eig * p1)
        else:
            alpha = -(q2 - kzt * eig * p2) / (q1 - kzt * eig * p1)
        if _alpha_ulps:
            alpha = alpha * (1.0 + _alpha_ulps * 2.0**-52)
        tp[:, msk] = alpha * sp1 + sp2
        tq[:, msk] = alpha * sq1 + sq2
        # degenerate mode (0,0): trapezoid in z, flux conserved (solver.py:239-251)
        lvl = 0
        p00 = srf_bg_conc
        for i in range(nz - 1):
            if i in levels:
                tp[lvl, 0, 0] = p00
                lvl += 1
            p00 = p00 - tq0[0, 0] * dz[i] * (0.5 / Kz[i] + 0.5 / Kz[i + 1])
        if nz - 1 in levels:
            tp[lvl, 0, 0] = p00

    if footprint:
        sx = xm if _mutation == "halo_in_shift" else xm + halo
        sy = ym if _mutation == "halo_in_shift" else ym + halo
        sgn = -1.0 if _mutation == "shift_sign" else 1.0
        shift = np.exp(sgn * 1j * (Lx * sx + Ly * sy))
        tp, tq = tp * shift, tq * shift
    elif xm**2 + ym**2 > 0.0:
        shift = np.exp(1j * (Lx * (xm - xmx / 2) + Ly * (ym - ymx / 2)))
        tp, tq = tp * shift, tq * shift

    P = np.zeros((nlvls, lty, ltx), dtype=tp.dtype)
    Q = np.zeros((nlvls, lty, ltx), dtype=tq.dtype)
    P[:, iy[:, None], ix[None, :]] = tp
    Q[:, iy[:, None], ix[None, :]] = tq
    if footprint:
        # unnormalised forward transform mirrors the Green's function
        p = np.fft.fft2(P, norm="backward").real
        q = np.fft.fft2(Q, norm="backward").real
    else:
        p = np.fft.ifft2(P, norm="forward").real
        q = np.fft.ifft2(Q, norm="forward").real
    conc = p[:, py : nye - py, px : nxe - px]
    flx = q[:, py : nye - py, px : nxe - px]

    x = np.linspace(0, xmx, nx, endpoint=False)
    y = np.linspace(0, ymx, ny, endpoint=False)
    Zg, Yg, Xg = np.meshgrid(z[levels], y, x, indexing="ij")
    result = ((np.squeeze(Xg), np.squeeze(Yg), np.squeeze(Zg)), np.squeeze(conc), np.squeeze(flx))
    if _return_spectra:
        return result, (tp, tq)
    return result
