"""Learned surrogate on the GPU: the FP64 point solver against the oracle's
footprint spectra, the bf16-MFMA forward and hand-written backward against
the numpy network with the same roundings, Adam against numpy, the error
gate and its FP64 fallback inside a plan.

Bars: point solver vs oracle transfer rel-L2 <= 1e-8 (FP64, unit instead of
1/N surface flux); network output vs numpy (bf16 storage mirrored) rel-L2
<= 1e-3, loss rel 1e-3, every gradient block rel-L2 <= 1e-2; Adam rel 1e-5;
fallback footprints bitwise equal to the plain batch.
"""

import dataclasses

import numpy as np
import pytest

from conftest import rel_l2
from oracle import solver as osolver
from oracle import surrogate as osur

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def metset():
    import torch

    assert torch.cuda.is_available()
    from bldfm_amd import surrogate as sur

    return sur.MetSet(sur.sample_met(8, seed=7))


@pytest.fixture(scope="module")
def net(metset):
    from bldfm_amd import surrogate as sur

    mu, sd = sur.met_normalisation(metset)
    return sur.SurrogateNet(mu, sd, seed=3)


def test_point_solver_matches_oracle_spectra(metset):
    """G from bldfm_transfer_solve equals the oracle footprint spectrum
    q^(zm)/(q00 e^{i k.x}) on a whole mode lattice."""
    import torch

    from bldfm_amd import surrogate as sur
    from oracle import pbl as opbl

    s = metset.samples
    for m in (0, 3):
        d = np.deg2rad(s["dir"][m])
        wind = (-s["speed"][m] * np.sin(d), -s["speed"][m] * np.cos(d))
        zm = s["zm"][m]
        z, prof = opbl.vertical_profiles(26, zm, wind, ustar=s["ustar"][m], mol=1.0 / s["invl"][m],
                                         domain_height=zm * 100.0 / 15.0)
        dom, modes = (500.0, 500.0), (64, 64)
        _, (tp, tq) = osolver.steady_state_transport_solver(
            np.zeros((64, 64)), z, prof, dom, 26, modes=modes, meas_pt=(250.0, 250.0), footprint=True,
            precision="double", _return_spectra=True)
        lat = sur.lattice(dom, modes)
        nxe = 64 * 3
        shift = np.exp(1j * (lat[:, 0] * 750.0 + lat[:, 1] * 750.0))
        g_or = tq[0].ravel() / shift * nxe * nxe
        P = lat.shape[0]
        kxy = torch.from_numpy(lat).cuda()
        blk = torch.full((P // 256,), m, dtype=torch.int32, device="cuda")
        g = metset.solve(blk, kxy).cpu().numpy()
        g = g[:, 0] + 1j * g[:, 1]
        assert rel_l2(g, g_or) < 1e-8


def _ws_views(ws, P):
    import torch

    NIN, W, NL = 96, 256, 5
    b = ws.view(torch.bfloat16)
    x0 = b[: NIN * P].view(NIN, P).float().cpu().numpy().T.astype(np.float64)
    return x0


def test_forward_and_gradient_match_numpy(metset, net):
    import torch

    from bldfm_amd import surrogate as sur

    # a non-zero output layer so every gradient block is exercised
    rng = np.random.default_rng(11)
    p = net.params.cpu().numpy().astype(np.float64)
    off, n = osur.layout()
    o, s = off["Wo"]
    p[o : o + 512] = rng.normal(0.0, 0.05, 512)
    p[off["bo"][0] : off["bo"][0] + 2] = [0.3, -0.2]
    net.params.copy_(torch.from_numpy(p.astype(np.float32)).cuda())
    net.pack()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    mids, kxy = net.sample_points(metset, 4, gen)  # 1024 points, 4 met samples
    g = metset.solve(mids, kxy)
    P = kxy.shape[0]
    ws = torch.zeros(P * net.bytes_per_point, dtype=torch.uint8, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    phi = torch.empty((P, 2), dtype=torch.float32, device="cuda")
    net.grad.zero_()
    blk = mids.repeat_interleave(2)
    net.grad_step(metset, blk, kxy, g, ws=ws, loss=loss, phi=phi)
    torch.cuda.synchronize()
    X0 = _ws_views(ws, P)
    # inputs: the kernel's features against the numpy features (1 bf16 ulp,
    # plus the fp32 rounding of RFF arguments up to ~60 rad)
    sm = metset.surmet_host[np.repeat(mids.cpu().numpy(), 256)]
    k = kxy.cpu().numpy()
    Xr = osur.features(net.static.numpy(), sm, k[:, 0], k[:, 1])
    assert np.all(np.abs(X0 - osur.bf16(Xr)) <= 2.0 ** -7 * np.abs(Xr) + 1e-4)
    # forward, loss, gradient against numpy with the kernels' roundings
    p0 = osur.phi0(sm, k[:, 0], k[:, 1])
    L, grad, phi_r = osur.loss_grad(p, X0, g.cpu().numpy(), p0, block=128, bf16_mode=True)
    assert rel_l2(phi.cpu().numpy(), phi_r) < 1e-3
    assert abs(float(loss.item()) - L) <= 1e-3 * abs(L)
    gk = net.grad.cpu().numpy().astype(np.float64)
    for name, (o, s) in off.items():
        sz = int(np.prod(s))
        assert rel_l2(gk[o : o + sz], grad[o : o + sz]) < 1e-2, name
    # the inference entry point computes the same network (fp32 summation
    # order may differ between instantiations)
    phi2 = net.phi(metset, blk, kxy)
    assert rel_l2(phi2.cpu().numpy(), phi.cpu().numpy()) < 1e-5


def test_adam_matches_numpy(net):
    import torch

    rng = np.random.default_rng(2)
    p0 = net.params.cpu().numpy().astype(np.float64)
    g = rng.normal(0.0, 1e-3, p0.size)
    net.grad.copy_(torch.from_numpy(g.astype(np.float32)).cuda())
    m0 = net.m.cpu().numpy().astype(np.float64)
    v0 = net.v.cpu().numpy().astype(np.float64)
    t = net.step + 1
    net.adam(1e-3)
    torch.cuda.synchronize()
    m = 0.9 * m0 + 0.1 * g
    v = 0.999 * v0 + 0.001 * g * g
    p = p0 - 1e-3 * (m / (1 - 0.9**t)) / (np.sqrt(v / (1 - 0.999**t)) + 1e-8)
    assert np.allclose(net.params.cpu().numpy(), p, rtol=1e-5, atol=1e-8)
    assert float(net.grad.abs().max()) == 0.0


def _specs(zms, L=(-20.0, 100.0)):
    from bldfm_amd.batch import FootprintSpec
    from bldfm_amd.pbl_model import psi

    specs = []
    for i, zm in enumerate(zms):
        mol = L[i % len(L)]
        speed = 0.3 / 0.4 * (np.log(zm / 0.05) + float(psi(zm / mol)))
        wd = np.deg2rad(30.0 + 70.0 * i)
        specs.append(FootprintSpec(nx=128, ny=128, domain=(500.0, 500.0), modes=(128, 128), meas_pt=(250.0, 250.0),
                                   n=26, meas_height=zm, wind=(-speed * np.sin(wd), -speed * np.cos(wd)), ustar=0.3,
                                   mol=mol, domain_height=zm * 100.0 / 15.0))
    return specs


def test_gate_falls_back_to_the_solver(net):
    """An untrained network (the WKB prior alone) fails every spot check, and an
    instance outside the hull is never served: all outputs equal the plain
    FP64 batch bit for bit."""
    import torch

    from bldfm_amd import surrogate as sur
    from bldfm_amd.batch import FootprintBatch

    untrained = sur.SurrogateNet(net.static.numpy()[132:140], net.static.numpy()[140:148], seed=1)
    specs = _specs([10.0, 15.0, 60.0])  # the last is above the hull's zm range
    plain = FootprintBatch(specs)
    plain.run()
    fb = FootprintBatch(specs, surrogate=untrained, tol=1e-3)
    fb.run()
    torch.cuda.synchronize()
    err, mode = fb.surrogate_stats()
    assert mode == [1, 1, 1]
    assert err[0] > 1e-3 and err[1] > 1e-3 and np.isnan(err[2])
    for i in range(len(specs)):
        for a, b in zip(fb.output(i), plain.output(i)):
            assert torch.equal(a, b)


def test_small_instances_are_never_served_unchecked(net):
    """Instances with few modes get fewer than BLDFM_SUR_MIN_SPOT spot-checked
    modes (with 8x8 modes, possibly none at all): whatever the tolerance and
    wherever they sit in the batch, the solver serves them, bit for bit; a
    large instance beside them is still served by the network."""
    import torch

    from bldfm_amd import surrogate as sur
    from bldfm_amd.batch import FootprintBatch

    untrained = sur.SurrogateNet(net.static.numpy()[132:140], net.static.numpy()[140:148], seed=1)
    base = _specs([10.0, 12.0, 14.0, 15.0, 18.0])
    modes = [(8, 8), (16, 16), (128, 128), (32, 32), (8, 8)]
    specs = [dataclasses.replace(s, modes=m) for s, m in zip(base, modes)]
    plain = FootprintBatch(specs)
    plain.run()
    fb = FootprintBatch(specs, surrogate=untrained, tol=1e9)
    fb.run()
    torch.cuda.synchronize()
    err, mode = fb.surrogate_stats()
    assert mode == [1, 1, 0, 1, 1], (err, mode)
    for i in (0, 1, 3, 4):
        for a, b in zip(fb.output(i), plain.output(i)):
            assert torch.equal(a, b)


def test_gate_accepts_within_tolerance(net):
    """With a tolerance above the untrained network's error every in-hull
    instance is served by the network (conc 0), and the spot-check estimate
    is of the size of the whole footprint's relative L2 error."""
    import torch

    from bldfm_amd import surrogate as sur
    from bldfm_amd.batch import FootprintBatch

    untrained = sur.SurrogateNet(net.static.numpy()[132:140], net.static.numpy()[140:148], seed=1)
    specs = _specs([10.0, 15.0])
    plain = FootprintBatch(specs)
    plain.run()
    fb = FootprintBatch(specs, surrogate=untrained, tol=1e9)
    fb.run()
    torch.cuda.synchronize()
    err, mode = fb.surrogate_stats()
    assert mode == [0, 0]
    for i in range(2):
        conc, flx = fb.output(i)
        assert float(conc.abs().max()) == 0.0
        f = flx[0].cpu().numpy()
        assert np.all(np.isfinite(f))
        e = rel_l2(f, plain.output(i)[1][0].cpu().numpy())
        assert err[i] > 0.1 and e > 0.1
        assert e / 3.0 < err[i] < 3.0 * e


@pytest.fixture(scope="module")
def trained():
    from bldfm_amd import surrogate as sur

    if not sur.DEFAULT_CHECKPOINT.exists():
        pytest.skip("no trained network shipped")
    return sur.SurrogateNet.load(sur.DEFAULT_CHECKPOINT)


def test_trained_network_gain_over_the_prior(trained):
    """On held-out met samples (a seed not used in training) the trained
    network's footprint error is below 1e-3 where the WKB prior alone (the
    method without the network) is off by tens of percent."""
    import torch

    from bldfm_amd import surrogate as sur

    ho = sur.MetSet(sur.sample_met(12, seed=99))
    ho = ho.subset(ho.conditioned(trained.hull.kmax))
    assert ho.n >= 6
    e = trained.evaluate(ho)
    static = trained.static.numpy()
    prior = sur.SurrogateNet(static[132:140], static[140:148], hull=trained.hull)
    prior.static = trained.static.clone()
    prior.net[: 4 * sur.STATIC_FLOATS].copy_(prior.static.view(torch.uint8).cuda())
    e0 = prior.evaluate(ho)
    assert np.median(e) < 1e-3 and e.max() < 3e-3
    assert np.median(e0) > 0.1
    assert np.median(e0) / np.median(e) > 100.0


def test_trained_surrogate_batch_within_tolerance(trained):
    """Served instances' flux footprints are within 1.5 tol of the FP64
    batch (the gate's 1 % estimate decides at tol); fallbacks are bitwise
    the FP64 footprint."""
    import torch

    from bldfm_amd.batch import FootprintBatch

    specs = _specs([8.0, 12.0, 20.0, 25.0], L=(-20.0, -100.0, 100.0, -50.0))
    plain = FootprintBatch(specs)
    plain.run()
    tol = 1e-3
    fb = FootprintBatch(specs, surrogate=trained, tol=tol)
    fb.run()
    torch.cuda.synchronize()
    err, mode = fb.surrogate_stats()
    assert sum(1 for m in mode if m == 0) >= 2, (err, mode)
    for i in range(len(specs)):
        a, b = fb.output(i)[1], plain.output(i)[1]
        if mode[i] == 0:
            assert err[i] <= tol
            assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1.5 * tol
        else:
            assert torch.equal(a, b)


def test_training_from_the_prior_on_the_gpu():
    """SurrogateNet.train with the batch of the training configuration
    (65,536 freshly FP64-solved points per step) from the untrained network,
    i.e. the WKB prior: 300 Adam steps of the hand-written kernels lower the
    loss and the held-out footprint error."""
    from bldfm_amd import surrogate as sur

    ms = sur.MetSet(sur.sample_met(600, seed=11))
    ms = ms.subset(ms.conditioned(ms.hull.kmax))
    ho = sur.MetSet(sur.sample_met(12, seed=12))
    ho = ho.subset(ho.conditioned(ho.hull.kmax))
    mu, sd = sur.met_normalisation(ms)
    net = sur.SurrogateNet(mu, sd, seed=0)
    e0 = net.evaluate(ho)
    log = net.train(ms, 300, batch=65536, log_every=30)
    e1 = net.evaluate(ho)
    losses = [v for _, v in log.loss]
    print(f"  {ms.n} samples: loss {losses[0]:.3e} -> {losses[-1]:.3e} in {log.seconds:.1f} s; held-out median "
          f"{np.median(e0):.3e} (prior) -> {np.median(e1):.3e}")
    assert losses[-1] < 0.5 * losses[0]
    assert np.median(e1) < 0.5 * np.median(e0)


def test_specified_network_trains():
    """The network exactly as specified (64 RFF of (kx, ky) + 6 met scalars ->
    Re G, Im G; MODE_SPEC) runs on the same kernels from a fixed-lattice
    dataset: 200 steps lower its loss (it starts from G = 0, with no prior)."""
    from bldfm_amd import surrogate as sur

    net, err, losses = sur.train_spec(samples=64, steps=200, heldout=6, log=lambda m: None)
    assert net.mode == sur.MODE_SPEC
    print(f"  loss {losses[0][1]:.3e} -> {losses[-1][1]:.3e}; held-out median {np.median(err):.3e}")
    assert losses[-1][1] < 0.8 * losses[0][1]


def test_specified_network_against_the_shipped_design():
    """Held-out footprint rel-L2 on the production 512^2 lattice (12 met
    samples of a seed used by neither training run) of the network trained
    exactly as specified and of the shipped design (wind-frame coordinates,
    8 met inputs, WKB prior, Phi = -log G)."""
    from bldfm_amd import surrogate as sur

    if not (sur.SPEC_CHECKPOINT.exists() and sur.DEFAULT_CHECKPOINT.exists()):
        pytest.skip("both checkpoints are needed")
    ho = sur.MetSet(sur.sample_met(12, seed=99))
    ho = ho.subset(ho.conditioned(ho.hull.kmax))
    spec = sur.SurrogateNet.load(sur.SPEC_CHECKPOINT)
    shipped = sur.SurrogateNet.load(sur.DEFAULT_CHECKPOINT)
    es, ed = spec.evaluate(ho), shipped.evaluate(ho)
    print(f"  held-out median / max: as specified {np.median(es):.3e} / {es.max():.3e}, "
          f"shipped {np.median(ed):.3e} / {ed.max():.3e}")
    assert spec.mode == sur.MODE_SPEC and shipped.mode == 0
    assert np.median(ed) < 1e-3 and np.median(ed) < np.median(es)
