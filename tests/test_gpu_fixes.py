"""Group fixes and computes on the GPU (fix/fix_*.cu, compute/compute_*.cu)
against the oracle's reference-order run, on the wetting geometry: a DPD fluid
between two frozen wall slabs with a probe cluster.  FP64 mode compares the
trajectory, the thermo records and every compute/spring record within 1e-9
(Saru streams are bit-identical, so only FP rounding separates the runs);
FP32 mode within the FP32 force tolerance over a few steps.  The hook path
(one ABI call per reference hook) must give the fused path's records.
"""
import numpy as np
import pytest

import gpumdpd_amd as g
from gpumdpd_amd import configs
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

CASES = {
    "two_wall": (
        [{"kind": "nve", "group": {1, 4}},
         {"kind": "two_wall", "group": {1}, "lo_group": {2}, "hi_group": {3}, "dir": 2}],
        [{"kind": "temp", "group": {1, 2, 3, 4}, "every": 5},
         {"kind": "rho", "group": {1}, "every": 5, "region": (0.0, 10.0, 0.0, 10.0, 3.0, 11.0)},
         {"kind": "wall_contact", "group": {1}, "every": 5, "wall_group": {2}, "side": 1, "dir": 2, "cutoff": 1.0},
         {"kind": "wall_contact", "group": {1}, "every": 5, "wall_group": {3}, "side": -1, "dir": 2, "cutoff": 1.0},
         {"kind": "tf", "group": {2}, "every": 4}]),
    "spring_rvcm": (
        [{"kind": "nve", "group": {1, 4}},
         {"kind": "spring", "group": {4}, "k": 50.0, "xyz": (6.0, 5.5, 7.5), "r0": 0.0, "every": 2},
         {"kind": "rvcm", "group": {1}, "every": 4, "flags": (1, 1, 0)}],
        [{"kind": "msd", "group": {4}, "every": 3}, {"kind": "temp", "group": {1}, "every": 3},
         {"kind": "tf", "group": {4}, "every": 2}]),
    "move_reflect_recenter": (
        [{"kind": "nve", "group": {1}},
         {"kind": "move", "group": {4}, "v": (0.2, -0.1, 0.05)},
         {"kind": "wall_reflect", "group": {1}, "lo": 1.3, "hi": 12.7, "dir": 2},
         {"kind": "recenter", "group": {1}, "every": 3, "xyz": (5.0, 5.0, 7.0)}],
        [{"kind": "msd", "group": {1, 4}, "every": 4}, {"kind": "rho", "group": {4}, "every": 2,
                                                          "region": (0.0, 10.0, 0.0, 10.0, 0.0, 14.0)}]),
}

NSTEPS, LOG = 20, 5


@pytest.fixture(scope="module")
def ctx():
    c = g.Context(0)
    yield c
    c.close()


def make(ctx, cfg, fixes, computes, prec):
    s = g.system_from_config(ctx, cfg, precision=prec, log_every=LOG)
    s.set_fixes(fixes)
    s.set_computes(computes)
    return s


@pytest.fixture(scope="module")
def cfg():
    return configs.config_wetting()


@pytest.fixture(scope="module")
def oracle_runs(cfg):
    return {name: O.run_fixes(cfg, NSTEPS, f, c, log_every=LOG) for name, (f, c) in CASES.items()}


def rec_keys(fixes, computes):
    keys = [("compute", k) for k in range(len(computes))]
    keys += [("fix", k) for k, f in enumerate(fixes) if f["kind"] == "spring"]
    return keys


def gpu_records(s, key):
    return s.records(key[1], fix=key[0] == "fix")


@pytest.mark.parametrize("name", list(CASES))
def test_fused_fix_path_matches_oracle_fp64(ctx, cfg, oracle_runs, name):
    fixes, comps = CASES[name]
    print(f"case: fixes {name} FP64 fused", flush=True)
    rows_o, rec_o, st_o = oracle_runs[name]
    s = make(ctx, cfg, fixes, comps, g.PREC_F64)
    s.run(NSTEPS)
    d = s.download()
    rows = s.thermo_rows()
    np.testing.assert_allclose(d["pos"], st_o["pos"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(d["vel"], st_o["vel"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(d["force"], st_o["force"], rtol=0, atol=1e-8)
    assert rows.shape == rows_o.shape
    np.testing.assert_allclose(rows, rows_o, rtol=1e-9, atol=1e-9)
    for key in rec_keys(fixes, comps):
        got, want = gpu_records(s, key), np.array(rec_o[key])
        assert got.shape == want.shape, key
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9, err_msg=str(key))
    s.destroy()


def test_fused_fix_path_fp32(ctx, cfg):
    fixes, comps = CASES["two_wall"]
    n = 5
    rows_o, rec_o, st_o = O.run_fixes(cfg, n, fixes, comps, log_every=LOG)
    print("case: fixes two_wall FP32 fused", flush=True)
    s = make(ctx, cfg, fixes, comps, g.PREC_F32)
    s.run(n)
    d = s.download()
    frms = np.sqrt(np.mean(np.sum(st_o["force"] ** 2, axis=1)))
    assert np.linalg.norm(d["force"] - st_o["force"]) / np.linalg.norm(st_o["force"]) < 1e-4
    assert np.max(np.abs(d["force"] - st_o["force"])) < 1e-2 * frms
    np.testing.assert_allclose(d["pos"], st_o["pos"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(s.thermo_rows(), rows_o, rtol=1e-5, atol=1e-5)
    for k in range(len(comps)):
        got, want = s.records(k), np.array(rec_o[("compute", k)])
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-3)
    s.destroy()


def hook_run(s, fixes, comps, nsteps):
    """The reference's loop (main/utils.cuh:166-314) as one ABI call per hook."""
    rows, recs = [], {}
    springs = [k for k, f in enumerate(fixes) if f["kind"] == "spring"]

    def post_force(step):
        for k in springs:
            out = s.fix_hook(k, "post_force", step)
            ev = fixes[k].get("every", 0)
            if ev and step % ev == 0:
                recs.setdefault(("fix", k), []).append([step] + list(out))

    def computes(step):
        for k, c in enumerate(comps):
            if step % c["every"] == 0:
                recs.setdefault(("compute", k), []).append([step] + list(s.compute_eval(k)) + [0.0, 0.0])

    s.pair_compute(0, log=True)
    post_force(0)
    rows.append(s.thermo())
    computes(0)
    for step in range(1, nsteps + 1):
        for hook in ("initial_integrate", "post_integrate"):
            for k in range(len(fixes)):
                s.fix_hook(k, hook, step)
        s.pair_compute(step, log=(step % LOG == 0))
        post_force(step)
        for hook in ("final_integrate", "end_of_step"):
            for k in range(len(fixes)):
                s.fix_hook(k, hook, step)
        if step % LOG == 0:
            rows.append(s.thermo())
        computes(step)
    return np.array(rows), recs


@pytest.mark.parametrize("name", ["spring_rvcm", "move_reflect_recenter"])
def test_hook_path_matches_oracle(ctx, cfg, oracle_runs, name):
    fixes, comps = CASES[name]
    print(f"case: fixes {name} FP64 hooks", flush=True)
    rows_o, rec_o, st_o = oracle_runs[name]
    s = make(ctx, cfg, fixes, comps, g.PREC_F64)
    rows, recs = hook_run(s, fixes, comps, NSTEPS)
    d = s.download()
    np.testing.assert_allclose(d["pos"], st_o["pos"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(d["vel"], st_o["vel"], rtol=0, atol=1e-9)
    # the hook thermo has no step column (-1)
    np.testing.assert_allclose(rows[:, :11], rows_o[:, :11], rtol=1e-9, atol=1e-9)
    for key in rec_keys(fixes, comps):
        np.testing.assert_allclose(np.array(recs[key]), np.array(rec_o[key]), rtol=1e-9, atol=1e-9,
                                   err_msg=str(key))
    s.destroy()


def test_fix_validation(ctx, cfg):
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F64)
    bad = [
        [{"kind": "nve", "group": {1}}, {"kind": "move", "group": {1}, "v": (0, 0, 0)}],
        [{"kind": "wall_reflect", "group": {1}, "lo": 5.0, "hi": 2.0, "dir": 2}],
        [{"kind": "wall_reflect", "group": {1}, "lo": -1.0, "hi": 2.0, "dir": 2}],
        [{"kind": "recenter", "group": {1}, "every": 0, "xyz": (1, 1, 1)}],
        [{"kind": "spring", "group": {1}, "k": 1.0, "xyz": (50.0, 1.0, 1.0), "r0": 0.0}],
        [{"kind": "nve", "group": {7}}],
    ]
    for fx in bad:
        with pytest.raises(g.GdpdError):
            s.set_fixes(fx)
    with pytest.raises(g.GdpdError):
        s.set_computes([{"kind": "rho", "group": {1}, "every": 1, "region": (0, 20, 0, 10, 0, 14)}])
    s.destroy()


def bonded_small():
    return configs.config_c2(n_chains=250, box=11.0)


def angles_small():
    return configs.config_angles(n_chains=250, box=11.0)


def mdpd_slab():
    return configs.config_c1()


OTHER = {
    # diblock melt: A = 1, B = 2; bonds enter the spring/recenter centres and tf (third law incl. bonds)
    "bonded": (bonded_small,
               [{"kind": "nve", "group": {1, 2}},
                {"kind": "rvcm", "group": {1, 2}, "every": 3, "flags": (1, 1, 1)},
                {"kind": "recenter", "group": {2}, "every": 4, "xyz": (5.5, 5.5, 5.5)},
                {"kind": "spring", "group": {1}, "k": 10.0, "xyz": (5.0, 6.0, 5.0), "r0": 0.5, "every": 1}],
               [{"kind": "msd", "group": {1}, "every": 5}, {"kind": "temp", "group": {2}, "every": 5},
                {"kind": "tf", "group": {1, 2}, "every": 5}]),
    "angles": (angles_small,
               [{"kind": "nve", "group": {1, 2}}, {"kind": "rvcm", "group": {1, 2}, "every": 2, "flags": (0, 1, 1)}],
               [{"kind": "tf", "group": {1}, "every": 5}]),
    # many-body slab kept centred, with the liquid density of its middle
    "mdpd": (mdpd_slab,
             [{"kind": "nve", "group": {1}},
              {"kind": "recenter", "group": {1}, "every": 5, "xyz": (20.0, 5.0, 5.0)},
              {"kind": "rvcm", "group": {1}, "every": 5, "flags": (1, 1, 1)}],
             [{"kind": "rho", "group": {1}, "every": 5, "region": (17.0, 23.0, 0.0, 10.0, 0.0, 10.0)},
              {"kind": "tf", "group": {1}, "every": 5}]),
}


@pytest.mark.parametrize("name", list(OTHER))
def test_fix_path_bonded_angles_mdpd_fp64(ctx, name):
    make_cfg, fixes, comps = OTHER[name]
    cfg = make_cfg()
    n = 15
    print(f"case: fixes {name} FP64 fused, N={cfg['N']}", flush=True)
    rows_o, rec_o, st_o = O.run_fixes(cfg, n, fixes, comps, log_every=LOG)
    s = make(ctx, cfg, fixes, comps, g.PREC_F64)
    s.run(n)
    d = s.download()
    np.testing.assert_allclose(d["pos"], st_o["pos"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(d["vel"], st_o["vel"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(s.thermo_rows(), rows_o, rtol=1e-9, atol=1e-9)
    for key in rec_keys(fixes, comps):
        np.testing.assert_allclose(gpu_records(s, key), np.array(rec_o[key]), rtol=1e-9, atol=1e-8,
                                   err_msg=str(key))
    s.destroy()
