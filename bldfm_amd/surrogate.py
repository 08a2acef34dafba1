"""Learned spectral surrogate of the footprint transfer function.

No counterpart in the reference.  In footprint mode the reference solves,
per retained Fourier mode k, the vertical ODE from a unit surface flux and
keeps the flux at the measurement height (src/bldfm/solver.py:132-134,
220-235, 307-374): the transfer G(k) = q^(zm; k) / q^(0; k).  The footprint is
the inverse transform of (1/N) G(k) e^{i k.x_tower}.  This module trains a
coordinate MLP on G, with the FP64 solver as data source, spot-check judge
and fallback:

* ``MetSet``: met samples (u*, 1/L, z0, zm, wind) on the training grid
  convention (``n`` points below zm, top at ``top_ratio * zm``, MOST), their
  level tables on the device, and the FP64 point solver
  (``bldfm_transfer_solve``) for any wavenumbers.
* ``SurrogateNet``: fp32 master parameters, Adam moments and the packed
  bf16 network buffer; ``train`` runs the hand-written forward, backward and
  Adam kernels on freshly solved points every step; ``evaluate`` reports the
  spectral relative L2 error (equal to the footprint's, by Parseval) against
  the solver on whole mode lattices.
* ``FootprintBatch(..., surrogate=net)`` (``bldfm_amd.batch``) serves whole
  footprints: the network on every mode, 1 % of the modes re-solved in FP64
  as an error gate, full FP64 fallback above the tolerance or outside the
  training hull.

Wavenumber and met inputs are rotated into the wind frame and scaled by zm:
with u and K proportional to u* under MOST, G depends on the met only
through zm/L and z0/zm (and the grid convention), which the network sees
among its inputs.
"""

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib

KARMAN = 0.4
N_RFF = 32
N_COORD = 4
N_MET = 8
STATIC_FLOATS = 256  # SurStatic region of the network buffer (1024 bytes)
SURMET_FLOATS = 76  # network inputs per met sample (SurMet)
TRANSFER_BLOCK = 256  # points per met sample block of the point solver


def sizes():
    """(fp32 parameter count, network buffer bytes, training bytes per point, points per block)."""
    lib = _lib.load()
    a, b, c = _lib._I64(), _lib._I64(), _lib._I64()
    d = _lib._I32()
    _lib.check(lib.bldfm_sur_sizes(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)))
    return a.value, b.value, c.value, d.value


@dataclass
class Hull:
    """Met box and grid convention of a network's training data."""

    ustar: tuple = (0.1, 1.0)
    invl: tuple = (-0.1, 0.05)
    z0: tuple = (0.01, 1.0)
    z0_zm_max: float = 0.1
    zm: tuple = (2.0, 30.0)
    kmax: float = 0.28
    top_ratio: float = 100.0 / 15.0
    n: int = 26
    closure: str = "MOST"

    def c_struct(self):
        h = _lib.SurHull()
        h.ustar_lo, h.ustar_hi = self.ustar
        h.invl_lo, h.invl_hi = self.invl
        h.z0_lo, h.z0_hi = self.z0
        h.z0_zm_max = self.z0_zm_max
        h.zm_lo, h.zm_hi = self.zm
        h.kmax = self.kmax
        h.top_ratio = self.top_ratio
        h.n = self.n
        h.closure = _lib.CLOSURES[self.closure]
        return h

    def as_dict(self):
        return dict(ustar=list(self.ustar), invl=list(self.invl), z0=list(self.z0), z0_zm_max=self.z0_zm_max,
                    zm=list(self.zm), kmax=self.kmax, top_ratio=self.top_ratio, n=self.n, closure=self.closure)


def sample_met(count, seed, hull: Hull = Hull()):
    """Seeded met samples inside ``hull``: u* ~ U, 1/L ~ U, z0 log-uniform
    (z0 <= z0_zm_max zm), zm ~ U, direction ~ U(0, 360).  The wind speed at
    zm follows from the log law so that vertical_profiles(..., ustar=u*)
    derives this z0 (src/bldfm/pbl_model.py:65-77)."""
    from .pbl_model import psi

    rng = np.random.default_rng(seed)
    out = {k: [] for k in ("ustar", "invl", "z0", "zm", "speed", "dir")}
    while len(out["ustar"]) < count:
        m = 2 * (count - len(out["ustar"])) + 16
        us = rng.uniform(*hull.ustar, m)
        il = rng.uniform(*hull.invl, m)
        z0 = np.exp(rng.uniform(np.log(hull.z0[0]), np.log(hull.z0[1]), m))
        zm = rng.uniform(*hull.zm, m)
        dr = rng.uniform(0.0, 360.0, m)
        keep = z0 <= hull.z0_zm_max * zm
        us, il, z0, zm, dr = us[keep], il[keep], z0[keep], zm[keep], dr[keep]
        sp = us / KARMAN * (np.log(zm / z0) + psi(zm * il))
        keep = sp > 0.1
        for k, v in zip(("ustar", "invl", "z0", "zm", "speed", "dir"), (us, il, z0, zm, sp, dr)):
            out[k].extend(v[keep].tolist())
    return {k: np.asarray(v[:count]) for k, v in out.items()}


def met_structs(samples, hull: Hull = Hull()):
    """``bldfm_met`` of each sample on the hull's grid convention (u* given, z0 derived)."""
    mets = []
    for us, il, zm, sp, dr in zip(samples["ustar"], samples["invl"], samples["zm"], samples["speed"],
                                  samples["dir"]):
        d = math.radians(dr)
        wind = (-sp * math.sin(d), -sp * math.cos(d))  # compute_wind_fields (src/bldfm/utils.py:7-27)
        mets.append(_lib.make_met(hull.n, zm, wind, us, None, 1.0 / il, 1.0, hull.closure,
                                  hull.top_ratio * zm, None, None))
    return mets


class MetSet:
    """Met samples with their device level tables and network inputs."""

    def __init__(self, samples, hull: Hull = Hull(), dev=None):
        lib = _lib.load()
        self.device = dev if dev is not None else _lib.device()
        self.hull = hull
        self.samples = samples
        mets = met_structs(samples, hull)
        n = len(mets)
        self.n = n
        arr = (_lib.Met * n)(*mets)
        nlev = []
        for m in mets:
            c = _lib._I32()
            _lib.check(lib.bldfm_profile_query(ctypes.byref(m), ctypes.byref(c), None, None), "profile_query")
            nlev.append(c.value)
        self.rows = max(nlev)
        self.tab = torch.empty(n * self.rows * 16, dtype=torch.float64, device=self.device)
        nz = (ctypes.c_int32 * n)()
        _lib.check(lib.bldfm_transfer_tables(self.device.index, arr, n, self.rows, _lib.ptr(self.tab), nz,
                                             _lib.stream_handle(self.device)), "transfer_tables")
        self.nz = torch.tensor(list(nz), dtype=torch.int32, device=self.device)
        self.lvl = torch.full((n,), hull.n, dtype=torch.int32, device=self.device)
        sm = np.zeros((n, SURMET_FLOATS), dtype=np.float32)
        _lib.check(lib.bldfm_sur_met(arr, n, sm.ctypes.data_as(ctypes.c_void_p)), "sur_met")
        self.surmet_host = sm
        self.surmet = torch.from_numpy(sm).to(self.device)

    def conditioned(self, kmax, probe=32):
        """Samples whose FP64 transfer is physics rather than roundoff on a
        probe x probe lattice of [-kmax, kmax]^2: finite, |G| <= 1 + 1e-6
        and |G| <= 100 |exp(-Phi0)| + 1e-8 with the WKB prior Phi0.  In
        deep, stable columns the shooting's cancellation (the reference's
        method, reproduced here) leaves noise where G is exponentially
        small; such samples are no training data, and at inference the
        spot check sends them to the solver."""
        k = np.linspace(-kmax, kmax, probe)
        kx, ky = np.meshgrid(k, k)
        pts = np.stack([kx.ravel(), ky.ravel()], 1)
        per = (pts.shape[0] + TRANSFER_BLOCK - 1) // TRANSFER_BLOCK
        pad = np.zeros((per * TRANSFER_BLOCK, 2))
        pad[: pts.shape[0]] = pts
        kxy = torch.from_numpy(np.tile(pad, (self.n, 1))).to(self.device)
        blk = torch.arange(self.n, dtype=torch.int32, device=self.device).repeat_interleave(per)
        g = self.solve(blk, kxy).view(self.n, per * TRANSFER_BLOCK, 2)
        prior = SurrogateNet(np.zeros(N_MET), np.ones(N_MET), hull=self.hull, dev=self.device)  # untrained: Phi0
        p0 = prior.phi(self, blk.repeat_interleave(2), kxy)[:, 0].double().view(self.n, -1)
        a = (g * g).sum(dim=2).sqrt()
        bad = ~torch.isfinite(a) | (a > 1.0 + 1e-6) | (a > 100.0 * torch.exp(-p0) + 1e-8)
        return (~bad.any(dim=1)).cpu().numpy()

    def subset(self, mask):
        """A MetSet of the samples where mask is true."""
        return MetSet({k: np.asarray(v)[mask] for k, v in self.samples.items()}, self.hull, self.device)

    def solve(self, blk_met, kxy, exact=True):
        """FP64 transfer of points kxy ((P, 2) float64, P a multiple of 256);
        block b of 256 points uses sample blk_met[b]."""
        lib = _lib.load()
        g = torch.empty_like(kxy)
        nblk = blk_met.numel()
        assert kxy.shape[0] == nblk * TRANSFER_BLOCK and blk_met.dtype == torch.int32
        _lib.check(lib.bldfm_transfer_solve(_lib.ptr(self.tab), self.rows, _lib.ptr(self.nz), _lib.ptr(self.lvl),
                                            self.n, _lib.ptr(blk_met), nblk, _lib.ptr(kxy), _lib.ptr(g),
                                            1 if exact else 0, _lib.stream_handle(self.device)), "transfer_solve")
        return g


def lattice(domain, modes, nxy=None):
    """Wavenumbers (lx, ly) of every retained mode of a footprint solve on
    ``domain`` with ``modes`` (grid = modes, default halo), as the solver
    forms them (src/bldfm/solver.py:148-153)."""
    nlx, nly = modes
    nx, ny = nxy if nxy is not None else modes
    dx, dy = domain[0] / nx, domain[1] / ny
    halo = max(domain)
    nxe, nye = nx + 2 * int(halo / dx), ny + 2 * int(halo / dy)
    lx = 2.0 * np.pi / dx / nxe * np.fft.fftfreq(nlx, d=1.0 / nlx)
    ly = 2.0 * np.pi / dy / nye * np.fft.fftfreq(nly, d=1.0 / nly)
    Lx, Ly = np.meshgrid(lx, ly)
    return np.stack([Lx.ravel(), Ly.ravel()], axis=1)


@dataclass
class TrainLog:
    steps: int = 0
    seconds: float = 0.0
    loss: list = field(default_factory=list)


MODE_SPEC = 1  # the network as specified: RFF(kx, ky) + 6 met scalars -> (Re G, Im G), no prior
STATIC_MODE = 148  # index of the mode in the feature header


class SurrogateNet:
    """Parameters, Adam state and packed buffer of one network.

    ``mode`` 0 (default): wind-frame, zm-scaled coordinates, 8 met inputs and
    the WKB prior; the outputs correct Phi = -log G.  ``mode`` MODE_SPEC: 64
    random Fourier features of the raw wavenumber (kx, ky), 6 normalised met
    scalars (u*, 1/L, ln z0, zm, |U|, direction), outputs Re G and Im G."""

    def __init__(self, met_mu, met_sd, hull: Hull = Hull(), seed=0, rff_sigma=2.0, dev=None, mode=0):
        self.device = dev if dev is not None else _lib.device()
        self.hull = hull
        nparam, net_bytes, bpp, blk = sizes()
        self.nparam, self.net_bytes, self.bytes_per_point, self.block = nparam, net_bytes, bpp, blk
        rng = np.random.default_rng(seed)
        static = np.zeros(STATIC_FLOATS, dtype=np.float32)
        static[: N_RFF * N_COORD] = rng.normal(0.0, rff_sigma, N_RFF * N_COORD)
        if mode == MODE_SPEC:  # B acts on (kx, ky) scaled by 1/kmax
            static[: N_RFF * N_COORD].reshape(N_RFF, N_COORD)[:, 2:] = 0.0
            static[128:132] = [1.0 / hull.kmax, 1.0 / hull.kmax, 0.0, 0.0]
        else:
            kz = hull.kmax * hull.zm[1] * math.sqrt(2.0)  # largest |k| zm
            static[128:132] = [1.0 / kz, 1.0 / kz, 1.0 / math.sqrt(kz), 1.0 / math.sqrt(kz)]
        static[132:140] = np.asarray(met_mu, dtype=np.float32)
        static[140:148] = np.asarray(met_sd, dtype=np.float32)
        static[STATIC_MODE] = float(mode)
        self.mode = int(mode)
        self.static = torch.from_numpy(static)
        self.params = torch.from_numpy(self._init_params(rng)).to(self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.step = 0
        self.net = torch.zeros(net_bytes, dtype=torch.uint8, device=self.device)
        self.net[: 4 * STATIC_FLOATS].copy_(self.static.view(torch.uint8).to(self.device))
        self.pack()

    @staticmethod
    def _init_params(rng):
        """He-normal hidden weights (fan-in of the real inputs), zero biases,
        zero output layer: the untrained network predicts G = 1."""
        W, NIN, NL = 256, 96, 5
        real_in = 2 * N_RFF + N_COORD + N_MET + 2
        p = []
        w0 = np.zeros((W, NIN), dtype=np.float32)
        w0[:, :real_in] = rng.normal(0.0, math.sqrt(2.0 / real_in), (W, real_in))
        p += [w0.ravel(), np.zeros(W, np.float32)]
        for _ in range(NL - 1):
            p += [rng.normal(0.0, math.sqrt(2.0 / W), W * W).astype(np.float32), np.zeros(W, np.float32)]
        p += [np.zeros(2 * W, np.float32), np.zeros(2, np.float32)]
        flat = np.concatenate(p)
        nparam = sizes()[0]
        assert flat.size == nparam, (flat.size, nparam)
        return flat

    def pack(self):
        _lib.check(_lib.load().bldfm_sur_pack(_lib.ptr(self.params), _lib.ptr(self.net),
                                               _lib.stream_handle(self.device)), "sur_pack")

    # ------------------------------------------------------------------ use
    def phi(self, metset: MetSet, blk_met, kxy):
        """Network output Phi = -log G, (P, 2) float32; block b of
        ``self.block`` points uses sample blk_met[b]."""
        P = kxy.shape[0]
        out = torch.empty((P, 2), dtype=torch.float32, device=self.device)
        _lib.check(_lib.load().bldfm_sur_forward(_lib.ptr(self.net), _lib.ptr(metset.surmet), metset.n,
                                                 _lib.ptr(blk_met), _lib.ptr(kxy), P, _lib.ptr(out),
                                                 _lib.stream_handle(self.device)), "sur_forward")
        return out

    def transfer(self, phi):
        """G from the network outputs, complex128: exp(-Phi) with |G| <= 1, or
        the outputs themselves in MODE_SPEC."""
        p = phi.double()
        if self.mode == MODE_SPEC:
            return torch.complex(p[:, 0], p[:, 1])
        return torch.exp(-torch.complex(p[:, 0].clamp(min=0.0), p[:, 1]))

    def grad_step(self, metset: MetSet, blk_met, kxy, g, ws=None, loss=None, phi=None):
        """Loss and gradient (accumulated into self.grad) of one batch."""
        P = kxy.shape[0]
        if ws is None:
            ws = torch.empty(P * self.bytes_per_point, dtype=torch.uint8, device=self.device)
        if loss is None:
            loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        _lib.check(_lib.load().bldfm_sur_train_grad(
            _lib.ptr(self.net), _lib.ptr(metset.surmet), metset.n, _lib.ptr(blk_met), _lib.ptr(kxy), _lib.ptr(g),
            P, _lib.ptr(ws), _lib.ptr(self.grad), _lib.ptr(loss), _lib.ptr(phi), _lib.stream_handle(self.device)),
            "sur_train_grad")
        return loss

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _lib.check(_lib.load().bldfm_sur_adam(_lib.ptr(self.params), _lib.ptr(self.grad), _lib.ptr(self.m),
                                              _lib.ptr(self.v), _lib.ptr(self.net), float(lr), float(beta1),
                                              float(beta2), float(eps), self.step,
                                              _lib.stream_handle(self.device)), "sur_adam")

    def sample_points(self, metset: MetSet, nblk, gen):
        """nblk blocks of 256 points: a random met sample per block and
        wavenumbers in [-kmax, kmax]^2; in every training block of
        ``self.block`` points half are uniform and half concentrated at
        small |k| (k = kmax u^3), so that each block's energy sum, which
        normalises its loss, includes the energetic low modes."""
        mids = torch.randint(0, metset.n, (nblk,), generator=gen, device=self.device, dtype=torch.int64)
        u = torch.rand((nblk * TRANSFER_BLOCK, 2), generator=gen, device=self.device, dtype=torch.float64) * 2 - 1
        u = u.view(-1, self.block, 2)
        u[:, self.block // 2 :] = u[:, self.block // 2 :] ** 3
        return mids.to(torch.int32), u.reshape(-1, 2) * self.hull.kmax

    def train(self, metset: MetSet, steps, batch=65536, lr=1e-3, warmup=200, seed=1, log_every=100,
              callback=None):
        """Adam with cosine learning-rate decay on freshly solved points."""
        import time

        assert batch % (2 * self.block) == 0 and self.block * 2 == TRANSFER_BLOCK
        nblk = batch // TRANSFER_BLOCK
        gen = torch.Generator(device=self.device)
        gen.manual_seed(seed)
        ws = torch.empty(batch * self.bytes_per_point, dtype=torch.uint8, device=self.device)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        log = TrainLog()
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        for it in range(steps):
            mids, kxy = self.sample_points(metset, nblk, gen)
            g = metset.solve(mids, kxy)
            loss.zero_()
            self.grad_step(metset, mids.repeat_interleave(2), kxy, g, ws=ws, loss=loss)
            scale = min(1.0, (it + 1) / warmup) * 0.5 * (1.0 + math.cos(math.pi * it / steps))
            self.adam(lr * scale)
            if log_every and (it % log_every == 0 or it == steps - 1):
                log.loss.append((it, float(loss.item())))
                if callback:
                    callback(it, log.loss[-1][1])
        torch.cuda.synchronize(self.device)
        log.steps = steps
        log.seconds = time.perf_counter() - t0
        return log

    def evaluate(self, metset: MetSet, domain=(2000.0, 2000.0), modes=(512, 512), exact=True):
        """Per-sample spectral relative L2 error of the network against the
        FP64 solver on the full mode lattice of a footprint solve.  By
        Parseval this is the relative L2 error of the whole (uncropped)
        footprint, the inverse DFT of G times a unit-modulus shift.  An
        untrained network (zero output layer) gives the error of the WKB
        prior alone."""
        lat = lattice(domain, modes)
        P = lat.shape[0]
        Pp = (P + TRANSFER_BLOCK - 1) // TRANSFER_BLOCK * TRANSFER_BLOCK
        kxy = torch.zeros((Pp, 2), dtype=torch.float64, device=self.device)
        kxy[:P] = torch.from_numpy(lat).to(self.device)
        errs = []
        for m in range(metset.n):
            blk = torch.full((Pp // TRANSFER_BLOCK,), m, dtype=torch.int32, device=self.device)
            g = metset.solve(blk, kxy, exact=exact)[:P]
            gs = torch.complex(g[:, 0], g[:, 1])
            gn = self.transfer(self.phi(metset, blk.repeat_interleave(2), kxy)[:P])
            errs.append(float(torch.linalg.vector_norm(gn - gs) / torch.linalg.vector_norm(gs)))
        return np.asarray(errs)

    # ------------------------------------------------------------ checkpoint
    def state(self, with_optimizer=True):
        st = {"params": self.params.cpu(), "step": torch.tensor(self.step), "static": self.static.clone(),
              "hull": torch.tensor([*self.hull.ustar, *self.hull.invl, *self.hull.z0, self.hull.z0_zm_max,
                                    *self.hull.zm, self.hull.kmax, self.hull.top_ratio, self.hull.n],
                                   dtype=torch.float64)}
        if with_optimizer:
            st["m"] = self.m.cpu()
            st["v"] = self.v.cpu()
        return st

    def save(self, path, with_optimizer=True):
        """Checkpoint: fp32 master parameters, the fixed feature header, the
        hull and (to resume training) the Adam moments and step."""
        torch.save(self.state(with_optimizer), path)

    @classmethod
    def load(cls, path, dev=None):
        st = torch.load(path, map_location="cpu", weights_only=True)
        h = st["hull"].tolist()
        hull = Hull(ustar=(h[0], h[1]), invl=(h[2], h[3]), z0=(h[4], h[5]), z0_zm_max=h[6], zm=(h[7], h[8]),
                    kmax=h[9], top_ratio=h[10], n=int(h[11]))
        static = st["static"].numpy()
        net = cls(static[132:140], static[140:148], hull=hull, dev=dev, mode=int(static[STATIC_MODE]))
        net.static = st["static"].clone()
        net.net[: 4 * STATIC_FLOATS].copy_(net.static.view(torch.uint8).to(net.device))
        net.params.copy_(st["params"].to(net.device))
        if "m" in st:
            net.m.copy_(st["m"].to(net.device))
            net.v.copy_(st["v"].to(net.device))
        net.step = int(st["step"])
        net.pack()
        return net


def met_normalisation(metset: MetSet):
    """Mean and standard deviation of the 8 met inputs over a sample set."""
    sm = metset.surmet_host[:, :N_MET].astype(np.float64)
    mu = sm.mean(axis=0)
    sd = sm.std(axis=0)
    sd[sd < 1e-6] = 1.0
    return mu.astype(np.float32), sd.astype(np.float32)


DEFAULT_CHECKPOINT = __import__("pathlib").Path(__file__).resolve().parent / "data" / "surrogate_most_n26.pt"


def train_default(steps=400000, pool=22000, seed=1, heldout=48, out=None, log=print):
    """The shipped network: ``pool`` met samples of the default hull (those
    the FP64 solver resolves), ``steps`` Adam steps of 65,536 fresh points
    each, lr 1e-3 with cosine decay; reports the held-out footprint error
    of the prior alone and of the trained network."""
    ms = MetSet(sample_met(pool, seed=seed))
    ms = ms.subset(ms.conditioned(ms.hull.kmax))
    ho = MetSet(sample_met(heldout, seed=seed + 1))
    ho = ho.subset(ho.conditioned(ho.hull.kmax))
    mu, sd = met_normalisation(ms)
    net = SurrogateNet(mu, sd, seed=0)
    prior = net.evaluate(ho)
    log(f"training pool {ms.n} samples, held out {ho.n}; prior alone: median {np.median(prior):.3e}")
    tl = net.train(ms, steps, log_every=max(steps // 20, 1),
                   callback=lambda it, l: log(f"  step {it} loss {l:.4e}"))
    err = net.evaluate(ho)
    log(f"{steps} steps in {tl.seconds:.1f} s; held-out footprint rel-L2 median {np.median(err):.3e} "
        f"p90 {np.percentile(err, 90):.3e} max {err.max():.3e}")
    if out is not None:
        net.save(out, with_optimizer=False)
    return net, prior, err


SPEC_CHECKPOINT = DEFAULT_CHECKPOINT.parent / "surrogate_spec_c4.pt"


def spec_dataset(metset: MetSet, domain=(1000.0, 1000.0), modes=(256, 256), chunk=64):
    """G of every sample on one fixed mode lattice (the 256^2 modes of a 256^2
    footprint on ``domain``: with 1000 m the production wavenumber range at
    half the density), solved in FP64, kept as fp32 pairs (n, P, 2)."""
    lat = torch.from_numpy(lattice(domain, modes)).to(metset.device)
    P = lat.shape[0]
    assert P % TRANSFER_BLOCK == 0
    data = torch.empty((metset.n, P, 2), dtype=torch.float32, device=metset.device)
    for c0 in range(0, metset.n, chunk):
        c1 = min(c0 + chunk, metset.n)
        blk = torch.arange(c0, c1, dtype=torch.int32, device=metset.device).repeat_interleave(P // TRANSFER_BLOCK)
        data[c0:c1] = metset.solve(blk, lat.repeat(c1 - c0, 1)).view(c1 - c0, P, 2).float()
    return lat, data


def train_spec(samples=20000, steps=200000, batch=65536, lr=1e-3, warmup=200, seed=5, heldout=32, out=None,
               log=print):
    """The network exactly as specified: ``samples`` seeded met draws (those
    the FP64 solver resolves), their G on a fixed 256^2 lattice, 64 RFF of
    (kx, ky) + 6 met scalars -> (Re G, Im G), batches of ``batch`` lattice
    points (half uniform, half from the quarter of the lattice nearest k = 0),
    Adam lr 1e-3 with cosine decay.  Returns the network, its held-out
    footprint rel-L2 on the production 512^2 lattice and the logged losses."""
    import time

    ms = MetSet(sample_met(samples, seed=seed))
    ms = ms.subset(ms.conditioned(ms.hull.kmax))
    ho = MetSet(sample_met(heldout, seed=seed + 1))
    ho = ho.subset(ho.conditioned(ho.hull.kmax))
    lat, data = spec_dataset(ms)
    P = lat.shape[0]
    near = torch.argsort(lat.norm(dim=1))[: P // 4]  # lattice points nearest k = 0
    mu, sd = met_normalisation(ms)
    net = SurrogateNet(mu, sd, seed=0, mode=MODE_SPEC)
    log(f"spec network: {ms.n} samples x {P} lattice points ({data.numel() * 4 / 2**30:.1f} GiB), held out {ho.n}")
    nblk = batch // net.block
    gen = torch.Generator(device=net.device)
    gen.manual_seed(seed)
    ws = torch.empty(batch * net.bytes_per_point, dtype=torch.uint8, device=net.device)
    loss = torch.zeros(1, dtype=torch.float32, device=net.device)
    torch.cuda.synchronize(net.device)
    t0 = time.perf_counter()
    losses = []
    for it in range(steps):
        sid = torch.randint(0, ms.n, (nblk,), generator=gen, device=net.device)
        pid = torch.randint(0, P, (nblk, net.block), generator=gen, device=net.device)
        pid[:, net.block // 2 :] = near[torch.randint(0, near.numel(), (nblk, net.block // 2), generator=gen,
                                                      device=net.device)]
        g = data[sid[:, None], pid].reshape(-1, 2).double()
        kxy = lat[pid.reshape(-1)]
        loss.zero_()
        net.grad_step(ms, sid.to(torch.int32), kxy, g, ws=ws, loss=loss)
        scale = min(1.0, (it + 1) / warmup) * 0.5 * (1.0 + math.cos(math.pi * it / steps))
        net.adam(lr * scale)
        if it % max(steps // 20, 1) == 0 or it == steps - 1:
            losses.append((it, float(loss.item())))
            log(f"  step {it} loss {losses[-1][1]:.4e}")
    torch.cuda.synchronize(net.device)
    err = net.evaluate(ho)
    log(f"{steps} steps in {time.perf_counter() - t0:.1f} s; held-out footprint rel-L2 median "
        f"{np.median(err):.3e} p90 {np.percentile(err, 90):.3e} max {err.max():.3e}")
    if out is not None:
        net.save(out, with_optimizer=False)
    return net, err, losses


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="Train the footprint-transfer surrogate on the GPU")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--spec", action="store_true",
                    help="the network exactly as specified (RFF of (kx, ky) + 6 met -> Re, Im; 20,000 samples)")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.spec:
        train_spec(steps=a.steps or 200000, out=a.out or str(SPEC_CHECKPOINT), log=lambda m: print(m, flush=True))
    else:
        train_default(a.steps or 400000, out=a.out or str(DEFAULT_CHECKPOINT), log=lambda m: print(m, flush=True))
