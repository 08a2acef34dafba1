"""GPU: the drop-in chain on random geometries against the oracle chain.

100 seeded random cases: grids of 12-96 cells per axis, even mode counts up
to the grid, domains of 50-3000 m, default or random halos, footprint or
dispersion (a random source, one to three random levels), MOST meteorology
(u*, L of either sign, wind speed and direction, measurement height, with
or without a domain height).  Each case runs device compute_wind_fields ->
device vertical_profiles -> device solver, and the oracle chain from the
same scalars.

Bars: where the oracle raises, the device raises the same exception type
(the reference's errors: odd padded-length parity in dispersion mode, empty
vertical grids).  Otherwise rel-L2 of conc and flx <= max(1e-10, 8 s), with
s the change of the oracle's own output when alpha moves by one ulp --
dispersion levels above the measurement height can depend on the last bit
of alpha (the growing and decaying solutions cancel), and no rounding-exact
restatement can be held closer there than the reference is to itself on
another CPU.  Every case's error and bar are printed.
"""

import numpy as np
import pytest

from conftest import rel_l2
from oracle import pbl as opbl
from oracle import solver as osolver

pytestmark = pytest.mark.gpu


def _cases(seed, count):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        nx, ny = int(rng.integers(12, 97)), int(rng.integers(12, 97))
        c = dict(nx=nx, ny=ny, modes=(2 * int(rng.integers(2, nx // 2 + 1)), 2 * int(rng.integers(2, ny // 2 + 1))),
                 domain=(float(rng.uniform(50, 3000)), float(rng.uniform(50, 3000))))
        c["halo"] = [None, float(rng.uniform(0, 1.5) * max(c["domain"]))][int(rng.integers(0, 2))]
        c["footprint"] = bool(rng.integers(0, 2))
        c["zm"], c["n"] = float(rng.uniform(2, 40)), int(rng.integers(6, 30))
        c["ustar"], c["mol"] = float(rng.uniform(0.1, 1.0)), float(rng.choice([-1, 1]) * rng.uniform(5, 500))
        c["speed"], c["dir"] = float(rng.uniform(1, 10)), float(rng.uniform(0, 360))
        c["domain_height"] = [None, float(c["zm"] * rng.uniform(1.2, 6))][int(rng.integers(0, 2))]
        c["level_draw"] = rng.integers(0, 10**6, size=int(rng.integers(1, 4))).tolist()
        c["meas_pt"] = (float(rng.uniform(0, c["domain"][0])), float(rng.uniform(0, c["domain"][1])))
        c["src_seed"] = int(rng.integers(0, 2**31))
        yield c


def _err(g, o):
    """rel-L2 over the entries the oracle has finite; the device must have
    NaN exactly where the oracle has (the reference's own grid can end in a
    NaN level: a short stretch under a tall domain height)."""
    g, o = np.asarray(g), np.asarray(o)
    gn, on = np.isnan(g), np.isnan(o)
    assert np.array_equal(gn, on), "NaN entries differ"
    return 0.0 if on.all() else rel_l2(g[~on], o[~on])


def _run(fn, *a, **k):
    try:
        return fn(*a, **k), None
    except Exception as e:  # noqa: BLE001 -- the exception type is compared
        return None, e


@pytest.fixture(scope="module")
def dropins():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from bldfm_amd import _lib, pbl_model, solver, utils

    _lib.load()
    return utils, pbl_model, solver


def test_random_geometries_match_the_oracle_chain(dropins):
    utils, pbl_model, solver = dropins
    compared = raised = 0
    worst = (0.0, None)
    for i, c in enumerate(_cases(seed=20261203, count=100)):
        u, v = utils.compute_wind_fields(c["speed"], c["dir"])
        d = np.deg2rad(c["dir"])
        ou, ov = c["speed"] * -np.sin(d), c["speed"] * -np.cos(d)  # src/bldfm/utils.py:23-27
        kwp = dict(ustar=c["ustar"], mol=c["mol"], domain_height=c["domain_height"])
        g, ge = _run(pbl_model.vertical_profiles, c["n"], c["zm"], (u, v), **kwp)
        o, oe = _run(opbl.vertical_profiles, c["n"], c["zm"], (ou, ov), **kwp)
        if oe is not None or ge is not None:
            assert type(ge) is type(oe), (i, c, ge, oe)
            raised += 1
            continue
        (z, prof), (oz, oprof) = g, o
        nz = len(oz)
        levels = c["n"] if c["footprint"] else sorted({k % nz for k in c["level_draw"]})
        q0 = (np.zeros((c["ny"], c["nx"])) if c["footprint"]
              else np.random.default_rng(c["src_seed"]).random((c["ny"], c["nx"])))
        kw = dict(modes=c["modes"], meas_pt=c["meas_pt"], footprint=c["footprint"], halo=c["halo"],
                  precision="double")
        g, ge = _run(solver.steady_state_transport_solver, q0, z, prof, c["domain"], levels, **kw)
        o, oe = _run(osolver.steady_state_transport_solver, q0, oz, oprof, c["domain"], levels, **kw)
        if oe is not None or ge is not None:
            assert type(ge) is type(oe), (i, c, ge, oe)
            raised += 1
            continue
        (_, gc, gf), (_, oc, of) = g, o
        _, ac, af = osolver.steady_state_transport_solver(q0, oz, oprof, c["domain"], levels, _alpha_ulps=1, **kw)
        err = max(rel_l2(gc, oc), rel_l2(gf, of))
        bar = max(1e-10, 8.0 * max(rel_l2(ac, oc), rel_l2(af, of)))
        print(f"  case {i}: {c['nx']}x{c['ny']} modes {c['modes']} {'footprint' if c['footprint'] else 'dispersion'}"
              f" levels {levels} of {nz}: rel-L2 {err:.2e} (bar {bar:.1e})", flush=True)
        assert err <= bar, (i, c, err, bar)
        compared += 1
        if err > worst[0]:
            worst = (err, i)
    print(f"  {compared} compared, {raised} raised the reference's errors on both sides; worst {worst[0]:.2e} (case {worst[1]})")
    assert compared >= 50


def _option_cases(seed, count):
    """Random closures and options: MOST/MOSTM/CONSTANT/OAAHOC, z0 instead of
    u*, stretch, single precision, the analytic branch, a background
    concentration, no halo, mode counts above the padded grid."""
    rng = np.random.default_rng(seed)
    for c in _cases(seed + 1, count):
        c["closure"] = ["MOST", "MOSTM", "CONSTANT", "OAAHOC"][int(rng.integers(0, 4))]
        c["prsc"] = float(rng.uniform(0.5, 1.5))
        c["tke"] = float(rng.uniform(0.1, 3.0)) if c["closure"] == "OAAHOC" else None
        c["z0_given"] = c["closure"] != "OAAHOC" and bool(rng.integers(0, 3) == 0)
        c["z0"] = float(rng.uniform(0.01, 0.1) * c["zm"])
        c["stretch"] = [None, float(c["zm"] * rng.uniform(0.5, 4))][int(rng.integers(0, 2))]
        c["precision"] = ["double", "single"][int(rng.integers(0, 2))]
        c["analytic"] = c["closure"] == "CONSTANT" and bool(rng.integers(0, 2))
        c["srf_bg_conc"] = [0.0, float(rng.uniform(-2, 2))][int(rng.integers(0, 2))]
        edge = int(rng.integers(0, 4))
        if edge == 1:  # no halo: the padded grid is the grid
            c["halo"] = 0.0
        elif edge == 2:  # mode counts up to three times the grid: clamped to the padded size when above it
            c["modes"] = (2 * int(rng.integers(2, 3 * c["nx"] // 2 + 1)), 2 * int(rng.integers(2, 3 * c["ny"] // 2 + 1)))
        yield c


def test_random_options_match_the_oracle_chain(dropins):
    """Bars as above; single precision: rel-L2 <= max(1e-6, 8 s) (float32
    outputs; the unshifted single-precision transform differs from the
    reference's float32 pyFFTW at that level)."""
    utils, pbl_model, solver = dropins
    compared = raised = 0
    for i, c in enumerate(_option_cases(seed=20261204, count=100)):
        u, v = utils.compute_wind_fields(c["speed"], c["dir"])
        d = np.deg2rad(c["dir"])
        ou, ov = c["speed"] * -np.sin(d), c["speed"] * -np.cos(d)
        kwp = dict(mol=c["mol"], prsc=c["prsc"], closure=c["closure"], domain_height=c["domain_height"],
                   stretch=c["stretch"], tke=c["tke"])
        if c["z0_given"]:
            kwp["z0"] = c["z0"]
        else:
            kwp["ustar"] = c["ustar"]
        g, ge = _run(pbl_model.vertical_profiles, c["n"], c["zm"], (u, v), **kwp)
        o, oe = _run(opbl.vertical_profiles, c["n"], c["zm"], (ou, ov), **kwp)
        if oe is not None or ge is not None:
            assert type(ge) is type(oe), (i, c, ge, oe)
            raised += 1
            continue
        (z, prof), (oz, oprof) = g, o
        nz = len(oz)
        analytic = c["analytic"]
        levels = c["n"] if (c["footprint"] or analytic) else sorted({k % nz for k in c["level_draw"]})
        q0 = (np.zeros((c["ny"], c["nx"])) if c["footprint"]
              else np.random.default_rng(c["src_seed"]).random((c["ny"], c["nx"])))
        kw = dict(modes=c["modes"], meas_pt=c["meas_pt"], footprint=c["footprint"], halo=c["halo"],
                  precision=c["precision"], analytic=analytic, srf_bg_conc=c["srf_bg_conc"])
        g, ge = _run(solver.steady_state_transport_solver, q0, z, prof, c["domain"], levels, **kw)
        o, oe = _run(osolver.steady_state_transport_solver, q0, oz, oprof, c["domain"], levels, **kw)
        if oe is not None or ge is not None:
            assert type(ge) is type(oe), (i, c, ge, oe)
            raised += 1
            continue
        (_, gc, gf), (_, oc, of) = g, o
        assert gc.dtype == oc.dtype and gc.shape == oc.shape, (i, c, gc.dtype, oc.dtype, gc.shape, oc.shape)
        for a, b in zip((z,) + tuple(prof), (oz,) + tuple(oprof)):
            assert np.array_equal(np.isnan(a), np.isnan(b)), (i, c)
        s = 0.0
        if not analytic:
            _, ac, af = osolver.steady_state_transport_solver(q0, oz, oprof, c["domain"], levels, _alpha_ulps=1, **kw)
            s = max(_err(ac, oc), _err(af, of))
        err = max(_err(gc, oc), _err(gf, of))
        bar = max(1e-6 if c["precision"] == "single" else 1e-10, 8.0 * s)
        print(f"  case {i}: {c['closure']}{' z0' if c['z0_given'] else ''}{' stretch' if c['stretch'] else ''} "
              f"{c['precision']}{' analytic' if analytic else ''} {'footprint' if c['footprint'] else 'dispersion'} "
              f"levels {levels} of {nz}: rel-L2 {err:.2e} (bar {bar:.1e})", flush=True)
        assert err <= bar, (i, c, err, bar)
        compared += 1
    print(f"  {compared} compared, {raised} raised the reference's errors on both sides")
    assert compared >= 40


def test_launch_shapes_are_bit_identical_on_random_geometries(dropins):
    """Every launch shape of a plan (modes per workgroup of the march 64 /
    128 / 256, x-pass input layout 0 / 1 / 2, the widest FFT workgroup) gives
    the same bits on random geometries: odd-parity footprints, dispersion
    with several levels, halos of every size."""
    import torch

    from bldfm_amd import _lib

    _, _, solver = dropins
    dev = _lib.device()
    shapes = [dict(shoot_block=b, xpass_layout=x, fft_threads=0) for b in (64, 128, 256) for x in (0, 1, 2)]
    checked = 0
    for i, c in enumerate(_cases(seed=20261207, count=40)):
        z, prof = opbl.vertical_profiles(c["n"], c["zm"], (3.0, -1.0), ustar=c["ustar"], mol=c["mol"])
        nz = len(z)
        levels = c["n"] if c["footprint"] else sorted({k % nz for k in c["level_draw"]})
        try:
            g = solver._geometry((c["ny"], c["nx"]), z, c["domain"], levels, c["modes"], c["meas_pt"],
                                 c["footprint"], c["halo"], "double")
        except Exception:  # noqa: BLE001 -- geometries the reference refuses
            continue
        slot, idx = solver._level_maps(g["levels"], nz)
        ints = np.concatenate([slot, idx]).astype(np.int32)
        inst = solver._instance(g, len(g["levels"]), c["footprint"], False, "double", 0.0)
        pr = torch.cat([torch.as_tensor(np.asarray(a, dtype=np.float64)) for a in (z,) + tuple(prof)]).to(dev)
        src = None
        if not c["footprint"]:
            src = torch.as_tensor(np.random.default_rng(c["src_seed"]).random((c["ny"], c["nx"]))).to(dev)
        outs = []
        for t in shapes + [dict(shoot_block=256, xpass_layout=0, fft_threads=1024)]:
            try:
                plan = _lib.Plan([inst], ints, dev, tuning=t)
            except ValueError:  # the 256-thread FFT family: its widest workgroup
                plan = _lib.Plan([inst], ints, dev, tuning=dict(t, fft_threads=256))
            conc = torch.empty((len(g["levels"]), g["nyo"], g["nxo"]), dtype=torch.float64, device=dev)
            flx = torch.empty_like(conc)
            ws = torch.empty(max(plan.workspace_bytes, 1), dtype=torch.uint8, device=dev)
            plan.execute(pr, src, conc, flx, ws)
            outs.append((conc.cpu().numpy(), flx.cpu().numpy()))
        for k, (a, b) in enumerate(outs[1:], 1):
            assert np.array_equal(a, outs[0][0]) and np.array_equal(b, outs[0][1]), (i, c, k)
        checked += 1
    print(f"  {checked} geometries x {len(shapes) + 1} launch shapes bit-identical")
    assert checked >= 20


def test_tiny_geometries(dropins):
    """Grids of 1-8 cells per axis with 2-8 modes (odd padded lengths,
    modes clamped to the padded grid, single-mode axes): the same results or
    exceptions as the oracle; bars as in the first sweep."""
    _, pbl_model, solver = dropins
    rng = np.random.default_rng(20261208)
    z, prof = opbl.vertical_profiles(10, 8.0, (2.0, 3.0), ustar=0.3, mol=-80.0)
    compared = raised = 0
    for i in range(80):
        nx, ny = int(rng.integers(1, 9)), int(rng.integers(1, 9))
        modes = (2 * int(rng.integers(1, 5)), 2 * int(rng.integers(1, 5)))
        footprint = bool(rng.integers(0, 2))
        dom = (float(rng.uniform(20, 400)), float(rng.uniform(20, 400)))
        halo = [None, 0.0, float(rng.uniform(0, 2) * max(dom))][int(rng.integers(0, 3))]
        levels = 10 if footprint else sorted({int(k) for k in rng.integers(0, len(z), 2)})
        q0 = np.zeros((ny, nx)) if footprint else rng.random((ny, nx))
        kw = dict(modes=modes, meas_pt=(0.3 * dom[0], 0.6 * dom[1]), footprint=footprint, halo=halo,
                  precision="double")
        g, ge = _run(solver.steady_state_transport_solver, q0, z, prof, dom, levels, **kw)
        o, oe = _run(osolver.steady_state_transport_solver, q0, z, prof, dom, levels, **kw)
        if oe is not None or ge is not None:
            assert type(ge) is type(oe), (i, nx, ny, modes, footprint, halo, ge, oe)
            raised += 1
            continue
        (_, gc, gf), (_, oc, of) = g, o
        assert gc.shape == oc.shape, (i, gc.shape, oc.shape)
        _, ac, af = osolver.steady_state_transport_solver(q0, z, prof, dom, levels, _alpha_ulps=1, **kw)
        err = max(_err(gc, oc), _err(gf, of))
        bar = max(1e-10, 8.0 * max(_err(ac, oc), _err(af, of)))
        assert err <= bar, (i, nx, ny, modes, footprint, halo, err, bar)
        compared += 1
    print(f"  {compared} tiny geometries compared, {raised} raised the reference's errors on both sides")
    assert compared >= 30


def test_heterogeneous_plan_equals_single_instance_plans(dropins):
    """One plan holding random dispersion and footprint instances of
    different grids, padded lengths (several FFT plans in one launch chain)
    and level counts: every instance's output equals its own single-instance
    plan bit for bit."""
    import torch

    from bldfm_amd import _lib

    _, _, solver = dropins
    dev = _lib.device()
    items = []
    for c in _cases(seed=20261209, count=60):
        if len(items) == 8:
            break
        z, prof = opbl.vertical_profiles(c["n"], c["zm"], (2.0, -1.5), ustar=c["ustar"], mol=c["mol"],
                                         domain_height=c["domain_height"])
        nz = len(z)
        levels = c["n"] if c["footprint"] else sorted({k % nz for k in c["level_draw"]})
        try:
            g = solver._geometry((c["ny"], c["nx"]), z, c["domain"], levels, c["modes"], c["meas_pt"],
                                 c["footprint"], c["halo"], "double")
        except Exception:  # noqa: BLE001 -- geometries the reference refuses
            continue
        slot, idx = solver._level_maps(g["levels"], nz)
        inst = solver._instance(g, len(g["levels"]), c["footprint"], False, "double", 0.0)
        src = None if c["footprint"] else np.random.default_rng(c["src_seed"]).random((c["ny"], c["nx"]))
        items.append((g, inst, np.concatenate([slot, idx]).astype(np.int32),
                      np.concatenate([z] + [np.asarray(a, dtype=np.float64) for a in prof]), src))
    assert len(items) == 8 and {it[1].footprint for it in items} == {0, 1}

    def run(plan, parts):
        prof = torch.zeros(max(plan.prof_doubles, 1), dtype=torch.float64)
        src = torch.zeros(max(plan.src_doubles, 1), dtype=torch.float64)
        for k, (_, _, _, p, s) in enumerate(parts):
            op, osrc, _, _ = plan.offsets[k]
            prof[op: op + p.size] = torch.from_numpy(p)
            if s is not None:
                src[osrc: osrc + s.size] = torch.from_numpy(s.ravel())
        conc = torch.empty(max(plan.out_elems, 1), dtype=torch.float64, device=dev)
        flx = torch.empty_like(conc)
        ws = torch.empty(max(plan.workspace_bytes, 1), dtype=torch.uint8, device=dev)
        plan.execute(prof.to(dev), src.to(dev) if plan.src_doubles > 0 else None, conc, flx, ws)
        out = []
        for k, (g, _, _, _, _) in enumerate(parts):
            _, _, oo, _ = plan.offsets[k]
            n = len(g["levels"]) * g["nyo"] * g["nxo"]
            out.append((conc[oo: oo + n].cpu().numpy(), flx[oo: oo + n].cpu().numpy()))
        return out

    multi = run(_lib.Plan([it[1] for it in items], np.concatenate([it[2] for it in items]), dev), items)
    for k, it in enumerate(items):
        (c1, f1), = run(_lib.Plan([it[1]], it[2], dev), [it])
        assert np.array_equal(multi[k][0], c1) and np.array_equal(multi[k][1], f1), k
    print(f"  {len(items)} instances ({sum(it[1].footprint for it in items)} footprints) in one plan")
