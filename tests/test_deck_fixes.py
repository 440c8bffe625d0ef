"""Reference decks with groups, the hot-path fixes and computes, run through the
C++ shells (tests/cpp/deck_runner): the fused run and the one-call-per-hook run
write the same thermo.txt, compute files and spring file, and their values are
the oracle's (compute/compute_*.cu file rows, fix/fix_spring.cu file rows).
"""
import os
import subprocess

import numpy as np
import pytest

from gpumdpd_amd import configs, datafile
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "tests", "cpp", "deck_runner")

GROUPS = ["group fluid type 1", "group lower type 2", "group upper type 3", "group probe type 4",
          "group all type 1 2 3 4"]
FIXES = ["fix 1 fluid nve", "fix 2 probe nve", "fix 3 fluid two_wall lower upper z",
         "fix 4 probe spring 50 6 5.5 7.5 0 spring.txt 2", "fix 5 fluid rvcm 4 1 1 0"]
COMPUTES = ["compute t1 all temp 5 temp.txt", "compute r1 fluid rho 5 0 10 0 10 3 11 rho.txt",
            "compute w1 fluid wall_contact 5 lower lo z 1.0 contact.txt", "compute f1 probe tf 4 tf.txt",
            "compute m1 probe msd 3 msd.txt"]
ORACLE_FIXES = [{"kind": "nve", "group": {1}}, {"kind": "nve", "group": {4}},
                {"kind": "two_wall", "group": {1}, "lo_group": {2}, "hi_group": {3}, "dir": 2},
                {"kind": "spring", "group": {4}, "k": 50.0, "xyz": (6.0, 5.5, 7.5), "r0": 0.0, "every": 2},
                {"kind": "rvcm", "group": {1}, "every": 4, "flags": (1, 1, 0)}]
ORACLE_COMPUTES = [{"kind": "temp", "group": {1, 2, 3, 4}, "every": 5},
                   {"kind": "rho", "group": {1}, "every": 5, "region": (0.0, 10.0, 0.0, 10.0, 3.0, 11.0)},
                   {"kind": "wall_contact", "group": {1}, "every": 5, "wall_group": {2}, "side": 1, "dir": 2,
                    "cutoff": 1.0},
                   {"kind": "tf", "group": {4}, "every": 4}, {"kind": "msd", "group": {4}, "every": 3}]
FILES = ["temp.txt", "rho.txt", "contact.txt", "tf.txt", "msd.txt"]
NSTEPS, LOG = 20, 5


def read_rows(path):
    with open(path) as f:
        lines = f.read().splitlines()
    return lines[0].split(), np.array([list(map(float, l.split())) for l in lines[1:]])


def deck_dir(tmp_path, name, lines, nsteps=NSTEPS):
    d = tmp_path / name
    d.mkdir()
    cfg = configs.config_wetting()
    datafile.write_data(cfg, str(d / "in.data"))
    datafile.write_deck(str(d / "run.in"), cfg, "in.data", nsteps, LOG, groups_fixes=lines)
    return cfg, d


def run(d, mode):
    assert os.path.exists(RUNNER), "build() compiles tests/cpp/deck_runner"
    return subprocess.run([RUNNER, str(d / "run.in"), mode, "f64"], capture_output=True, text=True, timeout=120)


@pytest.mark.gpu
def test_deck_fixes_and_computes_fused_hook_oracle(tmp_path):
    print("case: deck with fixes and computes", flush=True)
    cfg, fused = deck_dir(tmp_path, "fused", GROUPS + FIXES + COMPUTES)
    _, hook = deck_dir(tmp_path, "hook", GROUPS + FIXES + COMPUTES)
    for d, mode in ((fused, "fused"), (hook, "hook")):
        r = run(d, mode)
        assert r.returncode == 0, r.stderr
    rows_o, rec_o, _ = O.run_fixes(cfg, NSTEPS, ORACLE_FIXES, ORACLE_COMPUTES, log_every=LOG)
    for k, name in enumerate(FILES):
        hf, a = read_rows(str(fused / name))
        hh, b = read_rows(str(hook / name))
        assert hf == hh and hf[0] == "step"
        np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-9, err_msg=name)
        want = np.array(rec_o[("compute", k)])[:, : a.shape[1]]
        np.testing.assert_allclose(a, want, rtol=2e-5, atol=1e-5, err_msg=name)
    _, sa = read_rows(str(fused / "spring.txt"))
    _, sb = read_rows(str(hook / "spring.txt"))
    np.testing.assert_allclose(sa, sb, rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(sa, np.array(rec_o[("fix", 3)]), rtol=2e-5, atol=1e-5)
    _, ta = datafile.read_thermo(str(fused / "thermo.txt"))
    _, tb = datafile.read_thermo(str(hook / "thermo.txt"))
    np.testing.assert_allclose(ta, tb, rtol=1e-6, atol=2e-6)
    np.testing.assert_allclose(ta[:, 1], rows_o[:, 0], rtol=1e-6, atol=2e-6)


@pytest.mark.gpu
def test_deck_compute_dump_frames(tmp_path):
    print("case: deck with compute dump", flush=True)
    _, d = deck_dir(tmp_path, "dump", GROUPS + ["fix 1 all nve", "compute d1 probe dump 4 probe.trj"], nsteps=8)
    r = run(d, "fused")
    assert r.returncode == 0, r.stderr
    frames = datafile.read_dump(str(d / "probe.trj"))
    assert [f[0] for f in frames] == [0, 4, 8]
    for f in frames:
        assert f[1] == ["id", "type", "x", "y", "z"]
        assert len(f[2]) == 20 and all(int(row[1]) == 4 for row in f[2])


@pytest.mark.gpu
@pytest.mark.parametrize("line,message", [
    ("fix 9 fluid wall_reflect 5 2 z", "lo_wall_pos should lower than hi_wall_pos"),
    ("fix 9 fluid wall_reflect 1 20 z", "Wall position is out of the box!"),
    ("fix 9 fluid spring 1 50 1 1 0", "spring position is out of the box in x direction!"),
    ("fix 9 nobody nve", "Group nobody not found"),
    ("compute c9 fluid rho 5 0 20 0 10 0 14 r.txt", "xhi > box.xhi"),
    ("compute c9 fluid wall_contact 5 lower side z 1.0 c.txt", "illegal compute_wall_contact command"),
])
def test_deck_fix_and_compute_errors(tmp_path, line, message):
    print(f"case: deck error {line}", flush=True)
    _, d = deck_dir(tmp_path, "bad", GROUPS + ["fix 1 fluid nve", line], nsteps=2)
    r = run(d, "fused")
    assert r.returncode == 1
    assert message in r.stderr
