"""Data-file front end: the shells' read_data/write_data (C++) against the
Python writer and reader, the reference's error cases, and a whole deck
(run.in + data file) run through the C++ shells on the GPU, compared with the
same system run through the C ABI directly."""
import os
import re
import subprocess

import numpy as np
import pytest

import gpumdpd_amd as g
from gpumdpd_amd import configs, datafile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROUNDTRIP = os.path.join(ROOT, "tests", "cpp", "data_roundtrip")
RUNNER = os.path.join(ROOT, "tests", "cpp", "deck_runner")


def roundtrip(style, src, dst):
    assert os.path.exists(ROUNDTRIP), "build() compiles tests/cpp/data_roundtrip"
    return subprocess.run([ROUNDTRIP, style, src, dst], capture_output=True, text=True, timeout=60)


def bonded_cfg():
    cfg = configs.config_c2(n_chains=60, box=8.0)
    cfg["mol"] = 1 + (np.arange(cfg["N"]) // 16)
    return cfg


def angle_cfg():
    cfg = configs.config_angles(n_chains=60, box=8.0)
    cfg["mol"] = 1 + (np.arange(cfg["N"]) // 16)
    return cfg


def compare(d, cfg, tol):
    order = np.argsort(cfg["id"])
    assert d["n"] == cfg["N"]
    assert np.array_equal(d["id"], np.arange(1, cfg["N"] + 1))
    assert np.array_equal(d["type"], np.asarray(cfg["type"])[order])
    assert np.allclose(d["pos"], np.asarray(cfg["pos"])[order], atol=tol, rtol=0)
    assert np.allclose(d["vel"], np.asarray(cfg["vel"])[order], atol=tol, rtol=0)
    assert np.allclose(d["masses"], cfg["masses"], atol=tol)
    assert np.allclose(d["lo"], cfg["lo"]) and np.allclose(d["hi"], cfg["hi"])


def test_roundtrip_atomic(tmp_path):
    cfg = configs.scaled_dpd(2000, seed=3)
    src, dst = str(tmp_path / "in.data"), str(tmp_path / "out.data")
    datafile.write_data(cfg, src)
    r = roundtrip("atomic", src, dst)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["ok", str(cfg["N"]), "0", "1"]
    compare(datafile.read_data(dst), cfg, 6e-10)


def test_roundtrip_bonded_keeps_molecules_and_bonds(tmp_path):
    cfg = bonded_cfg()
    src, dst = str(tmp_path / "in.data"), str(tmp_path / "out.data")
    datafile.write_data(cfg, src)
    r = roundtrip("bond", src, dst)
    assert r.returncode == 0, r.stderr
    d = datafile.read_data(dst)
    compare(d, cfg, 6e-10)
    assert np.array_equal(d["mol"], cfg["mol"])
    assert np.array_equal(d["bonds"], np.asarray(cfg["bonds"]))


def test_positions_outside_the_box_are_wrapped(tmp_path):
    cfg = configs.scaled_dpd(500, seed=4)
    L = np.asarray(cfg["hi"]) - np.asarray(cfg["lo"])
    shifted = dict(cfg, pos=cfg["pos"] + np.array([2.0, -1.0, 3.0]) * L)
    src, dst = str(tmp_path / "in.data"), str(tmp_path / "out.data")
    datafile.write_data(shifted, src)
    assert roundtrip("atomic", src, dst).returncode == 0
    compare(datafile.read_data(dst), cfg, 1e-8)


def test_roundtrip_angles(tmp_path):
    cfg = angle_cfg()
    src, dst = str(tmp_path / "in.data"), str(tmp_path / "out.data")
    datafile.write_data(cfg, src)
    r = roundtrip("angle", src, dst)
    assert r.returncode == 0, r.stderr
    d = datafile.read_data(dst)
    compare(d, cfg, 6e-10)
    assert np.array_equal(d["bonds"], np.asarray(cfg["bonds"]))
    assert np.array_equal(d["angles"], np.asarray(cfg["angles"]))
    # an angle file is refused under atom_style bond
    assert roundtrip("bond", src, dst).returncode == 1


@pytest.mark.parametrize("edit,message", [
    (lambda t: t.replace("Masses", "Massless"), "Massless"),
    (lambda t: t.replace("Atoms  # atomic", "Atoms  # bond"), "inconsistent"),
    (lambda t: t.replace("\n2 1 ", "\n1 1 ", 1), "duplicate atom id"),
    (lambda t: t.replace("\n1 1 ", "\n0 1 ", 1), "Atom ID cannot be 0"),
    (lambda t: t.replace(" atom types", " atom types\n3 angles"), "angles but atom_style"),
    # a coordinate far outside the box is refused, not walked back box by box
    (lambda t: re.sub(r"\n(1 1 )\S+", r"\n\g<1>99999999999", t, count=1), "more than 16 box lengths"),
])
def test_reference_error_cases(tmp_path, edit, message):
    cfg = configs.scaled_dpd(300, seed=5)
    src = str(tmp_path / "in.data")
    datafile.write_data(cfg, src)
    text = open(src).read()
    bad = edit(text)
    assert bad != text
    open(src, "w").write(bad)
    r = roundtrip("atomic", src, str(tmp_path / "out.data"))
    assert r.returncode == 1
    assert message in r.stderr


def run_deck(tmp_path, cfg, nsteps, log_every, mode, prec):
    data, deck, out = tmp_path / "in.data", tmp_path / "run.in", tmp_path / f"out_{mode}.data"
    datafile.write_data(cfg, str(data))
    datafile.write_deck(str(deck), cfg, "in.data", nsteps, log_every, write_name=out.name)
    assert os.path.exists(RUNNER), "build() compiles tests/cpp/deck_runner"
    r = subprocess.run([RUNNER, str(deck), mode, prec], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = np.array([list(map(float, l.split()[1:])) for l in r.stdout.splitlines() if l.startswith("thermo")])
    return datafile.read_data(str(out)), rows


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["dpd", "bonded", "angles"])
def test_deck_through_shells_matches_direct_run(tmp_path, name):
    cfg = {"dpd": lambda: configs.scaled_dpd(3000, seed=6), "bonded": bonded_cfg, "angles": angle_cfg}[name]()
    nsteps, every = 20, 5
    print(f"case: deck {name}", flush=True)
    d_fused, rows_fused = run_deck(tmp_path, cfg, nsteps, every, "fused", "f64")
    d_hook, rows_hook = run_deck(tmp_path, cfg, nsteps, every, "hook", "f64")
    ctx = g.Context(0)
    try:
        s = g.system_from_config(ctx, cfg, precision=g.PREC_F64, log_every=every)
        s.run(nsteps)
        ref = s.download()
        ref_rows = s.thermo_rows()
        s.destroy()
    finally:
        ctx.close()
    for d in (d_fused, d_hook):
        assert np.allclose(d["pos"], ref["pos"], atol=6e-10, rtol=0)
        assert np.allclose(d["vel"], ref["vel"], atol=6e-10, rtol=0)
    # step, T, pair PE, bond PE, P, angle PE of every thermo row
    assert rows_fused.shape == (nsteps // every + 1, 6)
    assert np.allclose(rows_fused[:, 0], ref_rows[:, 11])
    assert np.allclose(rows_fused[:, 1:], ref_rows[:, [0, 1, 2, 3, 12]], rtol=1e-12, atol=1e-12)
    assert np.allclose(rows_hook, rows_fused, rtol=1e-11, atol=1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("wrapped", [True, False])
def test_deck_dump_and_thermo_files(tmp_path, wrapped):
    """dump frames (main/utils.cuh:200-204, 296-300 timing; system/dump.cu
    columns) and thermo.txt (compute/thermo.cu columns) against a direct run."""
    cfg = configs.scaled_dpd(2500, seed=8)
    nsteps, every = 12, 4
    data, deck = tmp_path / "in.data", tmp_path / "run.in"
    datafile.write_data(cfg, str(data))
    mode = "wrapped" if wrapped else "unwrapped"
    datafile.write_deck(str(deck), cfg, "in.data", nsteps, every, dump=f"2 5 {mode} force_yes vel_yes")
    print(f"case: deck dump {mode}", flush=True)
    r = subprocess.run([RUNNER, str(deck), "fused", "f64"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    frames = datafile.read_dump(str(tmp_path / "all.trj"))
    assert [f[0] for f in frames] == [2, 7, 12]
    assert frames[0][1] == ["id", "type", "x", "y", "z", "fx", "fy", "fz", "vx", "vy", "vz"]
    head, rows = datafile.read_thermo(str(tmp_path / "thermo.txt"))
    assert head == ["step", "temp", "pair_pe", "pressure", "pxx", "pxy", "pxz", "pyy", "pyz", "pzz"]
    ctx = g.Context(0)
    try:
        s = g.system_from_config(ctx, cfg, precision=g.PREC_F64, log_every=every)
        s.run(nsteps)
        ref = s.download()
        ref_rows = s.thermo_rows()
        s.destroy()
    finally:
        ctx.close()
    last = frames[-1][2]
    assert np.array_equal(last[:, 0], np.arange(1, cfg["N"] + 1))
    L = np.asarray(cfg["hi"]) - np.asarray(cfg["lo"])
    want = ref["pos"] if wrapped else ref["pos"] + ref["image"] * L
    assert np.allclose(last[:, 2:5], want, atol=2e-9, rtol=0)
    assert np.allclose(last[:, 5:8], ref["force"], atol=2e-9, rtol=0)
    assert np.allclose(last[:, 8:11], ref["vel"], atol=2e-9, rtol=0)
    assert np.allclose(rows[:, 0], ref_rows[:, 11])
    assert np.allclose(rows[:, 1], ref_rows[:, 0], atol=2e-6)
    assert np.allclose(rows[:, 2], ref_rows[:, 1], atol=2e-6)
    assert np.allclose(rows[:, 3:], ref_rows[:, 3:10], atol=2e-6)


def test_mutated_files_are_read_or_refused(tmp_path):
    """Randomly damaged data files (fields replaced by junk, huge or non-finite
    numbers, lines dropped, added or cut off) are either read or refused with
    an error; the reader never hangs or crashes."""
    import random
    rng = random.Random(19)
    cfg = configs.scaled_dpd(200, seed=9)
    src = str(tmp_path / "base.data")
    datafile.write_data(cfg, src)
    base = open(src).read().splitlines()
    junk = ["", "-1", "0", "1e400", "nan", "abc", "99999999999", "3.5", "2147483648", "#", "Atoms", "Masses"]
    for it in range(150):
        lines = list(base)
        for _ in range(rng.randint(1, 3)):
            op = rng.random()
            if op < 0.4:
                i = rng.randrange(len(lines))
                f = lines[i].split()
                if f:
                    f[rng.randrange(len(f))] = rng.choice(junk)
                    lines[i] = " ".join(f)
            elif op < 0.6:
                del lines[rng.randrange(len(lines))]
            elif op < 0.8:
                lines.insert(rng.randrange(len(lines) + 1), rng.choice(junk + ["1 1 1 0 0 0", "5 atoms"]))
            else:
                lines = lines[:rng.randrange(1, len(lines))]
        path = str(tmp_path / "m.data")
        open(path, "w").write("\n".join(lines) + "\n")
        r = roundtrip("atomic", path, str(tmp_path / "out.data"))
        assert r.returncode in (0, 1), (it, r.returncode, r.stderr[-300:])
