"""Refused calls through the C ABI: each returns its GDPD_ERR_* code with a
message (the reference prints the message and exits, force/pair_dpd.cu:78-81,
140-143, pair_mdpd.cu:81-89, 154-162, system/read_data.cu), and the context
stays usable afterwards."""
import numpy as np
import pytest

import gpumdpd_amd as g
from gpumdpd_amd import _lib, configs

pytestmark = pytest.mark.gpu

ARG, STATE, NAN = 1, 4, 6


@pytest.fixture(scope="module")
def ctx():
    c = g.Context(0)
    yield c
    c.close()


def refused(code, fragment, fn, *a):
    with pytest.raises(_lib.GdpdError) as e:
        fn(*a)
    assert e.value.code == code, str(e.value)
    assert fragment in str(e.value), str(e.value)


def fresh(ctx, cfg, prec=g.PREC_F32):
    d = g.make_desc(cfg["N"], cfg["ntypes"], cfg["lo"], cfg["hi"], cfg["dt"], cfg["global_cut"], cfg["skin"],
                    cfg["max_rho"], 1, 10, prec)
    return g.System(ctx, d)


def test_system_descriptor_checks(ctx):
    cfg = configs.scaled_dpd(1000, seed=1)
    print("case: descriptor checks", flush=True)
    for change, frag in ((dict(N=0), "n_beads must be positive"), (dict(ntypes=0), "n_types"),
                         (dict(hi=np.asarray(cfg["lo"]) - 1.0), "box hi must exceed lo"),
                         (dict(dt=0.0), "timestep must be positive")):
        c = dict(cfg, **change)
        refused(ARG, frag, fresh, ctx, c)
    refused(ARG, "unknown precision", fresh, ctx, cfg, 7)


def test_pair_style_checks(ctx):
    cfg = configs.scaled_dpd(1000, seed=2)
    s = fresh(ctx, cfg)
    print("case: pair style checks", flush=True)
    c3 = configs.dpd_coeff(1, {(1, 1): (25.0, 4.5, 1.0)})
    refused(ARG, "global rc should not be bigger than the global_cut", s.set_pair_dpd, 1.0, 1.5, 42, c3)
    refused(ARG, "global rc should not be bigger than the global_cut", s.set_pair_dpd, 1.0, 1.0, 42,
            configs.dpd_coeff(1, {(1, 1): (25.0, 4.5, 1.2)}))
    refused(ARG, "must be positive", s.set_pair_dpd, 1.0, 1.0, 42, configs.dpd_coeff(1, {(1, 1): (25.0, 4.5, 0.0)}))
    refused(ARG, "gamma must be non-negative", s.set_pair_dpd, 1.0, 1.0, 42,
            configs.dpd_coeff(1, {(1, 1): (25.0, -1.0, 1.0)}))
    refused(ARG, "temperature must be non-negative", s.set_pair_dpd, -1.0, 1.0, 42, c3)
    c5 = configs.mdpd_coeff(1, {(1, 1): (-40.0, 25.0, 4.5, 1.0, 0.75)})
    refused(ARG, "rd should not be bigger than rc", s.set_pair_mdpd, 1.0, 1.0, 1.1, 42, c5)
    refused(ARG, "rd should not be bigger than rc", s.set_pair_mdpd, 1.0, 1.0, 0.75, 42,
            configs.mdpd_coeff(1, {(1, 1): (-40.0, 25.0, 4.5, 0.7, 0.75)}))
    # a valid style still goes through afterwards
    s.set_pair_dpd(1.0, 1.0, 42, c3)
    s.destroy()


def test_upload_checks(ctx):
    cfg = configs.scaled_dpd(1000, seed=3)
    s = fresh(ctx, cfg)
    s.set_pair_dpd(1.0, 1.0, 42, cfg["coeff"])
    print("case: upload checks", flush=True)
    ids = np.array(cfg["id"], copy=True)
    ids[3] = ids[4]
    refused(ARG, "duplicate atom id", s.upload, ids, cfg["type"], cfg["pos"], cfg["vel"], cfg["masses"])
    ids = np.array(cfg["id"], copy=True)
    ids[0] = 0
    refused(ARG, "outside 1..", s.upload, ids, cfg["type"], cfg["pos"], cfg["vel"], cfg["masses"])
    typ = np.array(cfg["type"], copy=True)
    typ[7] = 2
    refused(ARG, "atom type 2 outside", s.upload, cfg["id"], typ, cfg["pos"], cfg["vel"], cfg["masses"])
    pos = np.array(cfg["pos"], copy=True)
    pos[9, 2] = np.nan
    refused(NAN, "non-finite position", s.upload, cfg["id"], cfg["type"], pos, cfg["vel"], cfg["masses"])
    refused(ARG, "mass of type 1 must be positive", s.upload, cfg["id"], cfg["type"], cfg["pos"], cfg["vel"],
            np.array([0.0]))
    s.upload(cfg["id"], cfg["type"], cfg["pos"], cfg["vel"], cfg["masses"])
    s.run(3)
    s.destroy()


def test_bond_checks(ctx):
    cfg = configs.config_c2(n_chains=50, box=8.0)
    s = g.system_from_config(ctx, {k: v for k, v in cfg.items() if k != "bonds"}, precision=g.PREC_F32)
    print("case: bond checks", flush=True)
    b = np.array(cfg["bonds"], copy=True).reshape(-1, 3)
    b[2, 0] = cfg["n_bond_types"] + 1
    refused(ARG, "type exceeds n_bond_types", s.set_bonds, b, cfg["n_bond_types"], cfg["bond_coeff"])
    b = np.array(cfg["bonds"], copy=True).reshape(-1, 3)
    b[5, 2] = cfg["N"] + 1
    refused(ARG, "atom id out of range", s.set_bonds, b, cfg["n_bond_types"], cfg["bond_coeff"])
    s.set_bonds(cfg["bonds"], cfg["n_bond_types"], cfg["bond_coeff"])
    s.run(3)
    s.destroy()


def test_call_order_and_tuning_checks(ctx):
    cfg = configs.scaled_dpd(1000, seed=4)
    s = fresh(ctx, cfg)
    s.set_pair_dpd(1.0, 1.0, 42, cfg["coeff"])
    print("case: call order", flush=True)
    refused(STATE, "upload the system before computing", s.zero_step)
    s.upload(cfg["id"], cfg["type"], cfg["pos"], cfg["vel"], cfg["masses"])
    refused(STATE, "call gdpd_zero_step first", s.advance, 5)
    s.zero_step()
    s.advance(2)
    refused(STATE, "brick shape is fixed", s.set_tuning, (3, 3, 3))
    refused(ARG, "capacities must be in", s.set_capacity, 70000, 10)
    refused(STATE, "pair style is fixed", s.set_pair_dpd, 1.0, 1.0, 42, cfg["coeff"])
    s.advance(2)
    s.finish()
    s.destroy()


def test_box_too_small_for_the_stencil(ctx):
    cfg = configs.dpd_fluid(20, (2.5, 4.0, 4.0), seed=5)
    cfg.update(style="dpd", temp=1.0, rc=1.0, saru_seed=42, dt=0.04, skin=0.0, global_cut=1.0, max_rho=6.0,
               coeff=configs.dpd_coeff(1, {(1, 1): (25.0, 4.5, 1.0)}))
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F32)
    print("case: box under 3 cells", flush=True)
    refused(ARG, "fewer than 3 cells", s.run, 2)
    s.destroy()


@pytest.mark.parametrize("bad", ["fast", "nan"])
def test_runaway_bead_is_reported(ctx, bad):
    """A bead that would move more than 16 box lengths in one step, or whose
    velocity is not finite, stops the run with GDPD_ERR_NAN instead of being
    binned out of range (the kernels put it back at the box corner and flag
    the system)."""
    cfg = configs.scaled_dpd(2000, seed=6)
    vel = np.array(cfg["vel"], copy=True)
    vel[11] = [1.0e5, 0.0, 0.0] if bad == "fast" else [np.nan, 0.0, 0.0]
    print(f"case: runaway bead {bad}", flush=True)
    s = g.system_from_config(ctx, dict(cfg, vel=vel), precision=g.PREC_F32)
    refused(NAN, "", s.run, 5)
    s.destroy()
    # the context goes on working
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F32)
    s.run(5)
    s.destroy()
