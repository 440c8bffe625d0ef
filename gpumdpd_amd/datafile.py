"""LAMMPS-style data files and run.in decks for the configs in configs.py.

The layout is the one the reference's read_data/write_data use
(system/read_data.cu:201-381, 469-637): header counts, box, Masses, Atoms
(atomic: id type x y z; bond: id mol type x y z), Velocities, Bonds.  These
helpers let a test or a user turn a generated system into the files a
reference deck reads, and read back a file written by write_data.
"""
import numpy as np


def write_data(cfg, path, title="gdpd data file"):
    """Write cfg (a configs.* dictionary) as a data file; atom_style bond when
    cfg has bonds.  Rows are written in id order."""
    n = cfg["N"]
    order = np.argsort(cfg["id"])
    ids = np.asarray(cfg["id"])[order]
    types = np.asarray(cfg["type"])[order]
    pos = np.asarray(cfg["pos"])[order]
    vel = np.asarray(cfg["vel"])[order]
    bonded = "bonds" in cfg
    angled = "angles" in cfg
    lines = [f"# {title}", "", f"{n} atoms"]
    if bonded:
        lines.append(f"{len(cfg['bonds'])} bonds")
    if angled:
        lines.append(f"{len(cfg['angles'])} angles")
    lines.append(f"{cfg['ntypes']} atom types")
    if bonded:
        lines.append(f"{cfg['n_bond_types']} bond types")
    if angled:
        lines.append(f"{cfg['n_angle_types']} angle types")
    lo, hi = cfg["lo"], cfg["hi"]
    lines += ["", f"{lo[0]:.17g} {hi[0]:.17g} xlo xhi", f"{lo[1]:.17g} {hi[1]:.17g} ylo yhi",
              f"{lo[2]:.17g} {hi[2]:.17g} zlo zhi", "", "Masses", ""]
    lines += [f"{t + 1} {m:.17g}" for t, m in enumerate(cfg["masses"])]
    lines += ["", "Atoms  # " + ("angle" if angled else "bond" if bonded else "atomic"), ""]
    mol = cfg.get("mol")
    for k in range(n):
        x = pos[k]
        if bonded:
            m = int(mol[ids[k] - 1]) if mol is not None else 1 + (ids[k] - 1) // 16
            lines.append(f"{ids[k]} {m} {types[k]} {x[0]:.17g} {x[1]:.17g} {x[2]:.17g}")
        else:
            lines.append(f"{ids[k]} {types[k]} {x[0]:.17g} {x[1]:.17g} {x[2]:.17g}")
    lines += ["", "Velocities", ""]
    lines += [f"{ids[k]} {vel[k][0]:.17g} {vel[k][1]:.17g} {vel[k][2]:.17g}" for k in range(n)]
    if bonded:
        lines += ["", "Bonds", ""]
        b = np.asarray(cfg["bonds"])
        lines += [f"{i + 1} {int(b[i, 0])} {int(b[i, 1])} {int(b[i, 2])}" for i in range(len(b))]
    if angled:
        lines += ["", "Angles", ""]
        a = np.asarray(cfg["angles"])
        lines += [f"{i + 1} {int(a[i, 0])} {int(a[i, 1])} {int(a[i, 2])} {int(a[i, 3])}" for i in range(len(a))]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_data(path):
    """Parse a data file into id-ordered arrays: dict with n, types, box,
    masses, id, type, mol (or None), pos, vel (or None), bonds (or None)."""
    with open(path) as f:
        raw = f.read().split("\n")
    strip = [l.split("#")[0].strip() for l in raw]
    out = {"mol": None, "vel": None, "bonds": None, "angles": None}
    i = 1
    counts = {}
    sections = {"Masses", "Atoms", "Velocities", "Bonds", "Angles"}
    lo, hi = [0.0] * 3, [0.0] * 3
    while i < len(strip) and strip[i] not in sections:
        t = strip[i].split()
        if len(t) >= 2 and t[-1] in ("atoms", "bonds", "angles"):
            counts[t[-1]] = int(t[0])
        elif len(t) >= 3 and t[-1] == "types":
            counts[t[-2] + " types"] = int(t[0])
        elif len(t) == 4 and t[2].endswith("lo"):
            d = "xyz".index(t[2][0])
            lo[d], hi[d] = float(t[0]), float(t[1])
        i += 1
    n = counts["atoms"]
    out.update(n=n, lo=lo, hi=hi, ntypes=counts["atom types"])
    style_bond = False

    def rows(start, count):
        r, k = [], start
        while len(r) < count:
            if strip[k]:
                r.append(strip[k].split())
            k += 1
        return r, k

    while i < len(strip):
        name = strip[i]
        if name not in sections:
            i += 1
            continue
        if name == "Atoms":
            style_bond = "bond" in raw[i] or "angle" in raw[i]
        count = {"Masses": out["ntypes"], "Atoms": n, "Velocities": n, "Bonds": counts.get("bonds", 0),
                 "Angles": counts.get("angles", 0)}[name]
        r, i = rows(i + 1, count)
        a = np.array(r, dtype=np.float64) if r else np.zeros((0, 4))
        if name == "Masses":
            out["masses"] = a[np.argsort(a[:, 0]), 1]
        elif name == "Atoms":
            a = a[np.argsort(a[:, 0])]
            out["id"] = a[:, 0].astype(np.int64)
            if style_bond:
                out["mol"], out["type"], out["pos"] = a[:, 1].astype(np.int64), a[:, 2].astype(np.int64), a[:, 3:6]
            else:
                out["type"], out["pos"] = a[:, 1].astype(np.int64), a[:, 2:5]
        elif name == "Velocities":
            a = a[np.argsort(a[:, 0])]
            out["vel"] = a[:, 1:4]
        elif name == "Bonds":
            a = a[np.argsort(a[:, 0])]
            out["bonds"] = a[:, 1:4].astype(np.int64)
        elif name == "Angles":
            a = a[np.argsort(a[:, 0])]
            out["angles"] = a[:, 1:5].astype(np.int64)
    return out


def write_deck(path, cfg, data_name, nsteps, log_every, write_name=None, dump=None, groups_fixes=None):
    """A reference-style run.in for cfg (pair style, coefficients, bonds,
    neighbor, timestep, thermo, run, read_data, group and fix nve); dump is
    the argument string of a dump command, e.g. "0 5 wrapped force_yes vel_yes".
    groups_fixes (a list of group / fix / compute lines) replaces the default
    "group all" + "fix md all nve"."""
    bonded = "bonds" in cfg
    angled = "angles" in cfg
    nt = cfg["ntypes"]
    style = "angle" if angled else "bond" if bonded else "atomic"
    lines = ["# generated deck", f"atom_style {style}",
             f"neighbor {cfg['global_cut']} {cfg['skin']} {cfg['max_rho']} 1", f"timestep {cfg['dt']}",
             f"thermo {log_every}", f"run {nsteps}", f"read_data {data_name}"]
    if write_name:
        lines.append(f"write_data {write_name}")
    if dump:
        lines.append(f"dump {dump}")
    c = np.asarray(cfg["coeff"], dtype=np.float64)
    if cfg["style"] == "mdpd":
        lines.append(f"pair_style mdpd {cfg['temp']} {cfg['rc']} {cfg['rd']} {cfg['saru_seed']}")
        w = 5
    else:
        lines.append(f"pair_style dpd {cfg['temp']} {cfg['rc']} {cfg['saru_seed']}")
        w = 3
    c = c.reshape(-1, w)
    k = 0
    for i in range(1, nt + 1):
        for j in range(i, nt + 1):
            v = c[k]
            if w == 3:
                lines.append(f"pair_coeff {i} {j} {v[0]:.17g} {v[1]:.17g} {v[2]:.17g}")
            else:
                lines.append(f"pair_coeff {i} {j} {v[0]:.17g} {v[1]:.17g} {v[2]:.17g} {v[3]:.17g} {v[4]:.17g}")
            k += 1
    if bonded:
        lines.append("bond_style harmonic")
        bc = np.asarray(cfg["bond_coeff"], dtype=np.float64).reshape(-1, 2)
        lines += [f"bond_coeff {t + 1} {bc[t, 0]:.17g} {bc[t, 1]:.17g}" for t in range(len(bc))]
    if angled:
        lines.append("angle_style harmonic")
        ac = np.asarray(cfg["angle_coeff"], dtype=np.float64).reshape(-1, 2)
        lines += [f"angle_coeff {t + 1} {ac[t, 0]:.17g} {ac[t, 1]:.17g}" for t in range(len(ac))]
    if groups_fixes is None:
        groups_fixes = ["group all type " + " ".join(str(t) for t in range(1, nt + 1)), "fix md all nve"]
    lines += list(groups_fixes)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_dump(path):
    """Frames of a LAMMPS text dump: list of (step, columns, array rows x cols)."""
    frames = []
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines):
        if lines[i].startswith("ITEM: TIMESTEP"):
            step = int(lines[i + 1])
            n = int(lines[i + 3])
            cols = lines[i + 8].split()[2:]
            data = np.array([list(map(float, lines[i + 9 + k].split())) for k in range(n)])
            frames.append((step, cols, data))
            i += 9 + n
        else:
            i += 1
    return frames


def read_thermo(path):
    """thermo.txt written by the deck runner: (header columns, rows)."""
    with open(path) as f:
        head = f.readline().split()
        rows = np.array([list(map(float, l.split())) for l in f if l.strip()])
    return head, rows
