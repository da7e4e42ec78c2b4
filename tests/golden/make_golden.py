"""Write the golden fixtures of this directory from the reference CPU path.

Runs oracle/_ref/half_ref_driver (the reference's own Fortran CPU modules,
compiled by oracle/Makefile against MKL) on the seeded synthetic inputs and
stores its outputs as JSON.  Run from the repository root after
`make -C oracle`:

    python tests/golden/make_golden.py [case ...]

Cases: si2_pbe, si2_lda, si2_k (PBE at k = (0.1, 0.2, 0.3)), hfo2_pbe, si8k_k
(8-atom cell at k = (1/6, -1/3, 1/2) of the 6x6x6 mesh), kat, and the
potential-dependent D correction (QDEP, src/half_uspp.F90) on datasets with
PAW partial waves: si2_qdep (k = (0.1, 0.2, 0.3)) and hfo2_qdep (Gamma).
Force cases (the analytic force assembly of app/half_cli.F90:921-962) on the
si2 cell with atom 0 displaced by (0.06, -0.03, 0.02) A (si2d.CHGCAR):
si2d_forces (plain datasets, PBE) and si2d_core_lda / si2d_core_pbe
(datasets with PAW partial waves and a partial core, QDEP on).
The si2 input files are committed next to this script; the hfo2 inputs are
regenerated by the seeded generator and their sha256 is recorded so a test can
refuse to compare against goldens made from different inputs.
"""
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "deepaw-half_amd"))
from half_amd import synth  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "half_ref_driver")
INPUTS = os.path.join(HERE, "inputs")

CASES = {  # workload, xc, k point, bands, QDEP
    "si2_pbe": ("si2", "pbe", (0.0, 0.0, 0.0), 16, False),
    "si2_lda": ("si2", "lda", (0.0, 0.0, 0.0), 16, False),
    "si2_k": ("si2", "pbe", (0.1, 0.2, 0.3), 16, False),
    # 104 bands: the 96-band mixed-precision solve of the same cell is
    # compared band by band, with room above its highest band
    "hfo2_pbe": ("hfo2", "pbe", (0.0, 0.0, 0.0), 104, False),
    "si8k_k": ("si8k", "pbe", (1.0 / 6.0, -1.0 / 3.0, 0.5), 32, False),
    "si2_qdep": ("si2", "pbe", (0.1, 0.2, 0.3), 16, True),
    "hfo2_qdep": ("hfo2", "pbe", (0.0, 0.0, 0.0), 64, True),
}

FORCE_CASES = {  # POTCAR variant, xc, k point, bands, QDEP
    "si2d_forces": ("si2.POTCAR", "pbe", (0.1, 0.2, 0.3), 8, False),
    "si2d_core_lda": ("si2.paw.core.POTCAR", "lda", (0.1, 0.2, 0.3), 8, True),
    "si2d_core_pbe": ("si2.paw.core.POTCAR", "pbe", (0.1, 0.2, 0.3), 8, True),
}
SI2D_SHIFT = (0.06, -0.03, 0.02)  # Cartesian displacement of atom 0 (A)

_FLOAT = re.compile(r"^([-+]?\d+\.\d+)([-+]\d{3})$")


def _num(tok):
    # Fortran ES format drops the exponent letter for 3-digit exponents
    m = _FLOAT.match(tok)
    if m:
        tok = m.group(1) + "E" + m.group(2)
    return float(tok)


def parse(text):
    out = {}
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        out.setdefault(parts[0], []).append([_num(t) for t in parts[1:]])
    return out


def run_case(name):
    wl, xc, k, nb, qdep = CASES[name]
    if wl == "si2":
        chg = os.path.join(INPUTS, "si2.CHGCAR")
        pot = os.path.join(INPUTS, "si2.paw.POTCAR" if qdep else "si2.POTCAR")
        if not os.path.exists(chg) or not os.path.exists(pot):
            os.makedirs(INPUTS, exist_ok=True)
            chg, pot, _ = synth.make_inputs("si2", INPUTS, partial_waves=qdep)
    else:
        chg, pot, _ = synth.make_inputs(wl, tempfile.mkdtemp(prefix="half_golden_"), partial_waves=qdep)
    w = synth.workload(wl)
    veff_path = os.path.join(HERE, f"{name}.veff.bin.tmp")
    cmd = [DRIVER, "solve", chg, pot, repr(w.encut), xc, str(nb), *map(repr, k), veff_path]
    if qdep:
        cmd.append("uspp")
    text = subprocess.run(cmd, check=True, capture_output=True, text=True).stdout
    d = parse(text)
    veff = np.fromfile(veff_path, dtype="<f8")
    os.remove(veff_path)
    g = {
        "workload": wl, "xc": xc, "kpoint": list(k), "nbands": nb, "encut": w.encut, "qdep": qdep,
        "chgcar_sha256": synth.sha256(chg), "potcar_sha256": synth.sha256(pot),
        "ngrid": int(d["ngrid"][0][0]), "npw": int(d["npw"][0][0]),
        "e_hartree": d["energy_terms"][0][0], "e_xc": d["energy_terms"][0][1],
        "e_xc_potential": d["energy_terms"][0][2], "volume": d["volume"][0][0],
        "overlap_min": d["overlap"][0][0], "overlap_max": d["overlap"][0][1],
        "eigenvalues": [r[1] for r in d["eig"]],
        "occupations": [r[1] for r in d["occ"]],
        "band_energy": d["band_energy"][0][0], "fermi": d["fermi"][0][0],
        "ewald": d["ewald"][0][0], "local_g0": d["local_g0"][0][0],
        "atomic_reference": d["atomic_reference"][0][0],
        "harris_energy": d["harris_energy"][0][0],
        "force_nl": [r[1:] for r in d["force_nl"]],
        # (i, j) one-based, Re H, Im H, Re S, Im S
        "hs_samples": d["hs"],
        "h_diag": [r[1] for r in d["hdiag"]],
        "s_diag": [r[2] for r in d["hdiag"]],
        "veff_sum": float(veff.sum()), "veff_sumsq": float((veff * veff).sum()),
        "veff_min": float(veff.min()), "veff_max": float(veff.max()),
    }
    if qdep:
        # per-ion D (one-based ion, a, b) and dD/dR (ion, a, b, alpha)
        g["dij"] = [[int(r[0]), int(r[1]), int(r[2]), r[3]] for r in d["dij"]]
        g["ddij"] = [[int(r[0]), int(r[1]), int(r[2]), int(r[3]), r[4]] for r in d["ddij"]]
    if wl == "si2" and not qdep:
        # full C-order Veff grid and the three-column apply image
        np.save(os.path.join(HERE, f"{name}.veff.npy"), veff)
        ap = np.array(d["apply"])
        n = g["npw"]
        hpsi = (ap[:, 2] + 1j * ap[:, 3]).reshape(3, n).T
        spsi = (ap[:, 4] + 1j * ap[:, 5]).reshape(3, n).T
        np.save(os.path.join(HERE, f"{name}.apply.npy"), np.stack([hpsi, spsi]))
    else:
        idx = np.arange(0, veff.size, max(1, veff.size // 2000))
        g["veff_sample_index"] = idx.tolist()
        g["veff_sample"] = veff[idx].tolist()
    with open(os.path.join(HERE, f"{name}.json"), "w") as fh:
        json.dump(g, fh, indent=1)
    print(name, "npw", g["npw"], "eig0", g["eigenvalues"][0])


def si2d_inputs():
    """si2 cell with atom 0 displaced (no symmetry zeroes the forces), the
    density generated for the displaced atoms; POTCAR variants of si2."""
    import dataclasses
    chg = os.path.join(INPUTS, "si2d.CHGCAR")
    if not os.path.exists(chg):
        w = synth.workload("si2")
        pos = w.positions.copy()
        pos[0] += np.asarray(SI2D_SHIFT) @ np.linalg.inv(w.lattice)
        synth.write_chgcar(chg, dataclasses.replace(w, positions=pos))
    core = os.path.join(INPUTS, "si2.paw.core.POTCAR")
    if not os.path.exists(core):
        synth.write_potcar(core, synth.workload("si2"), True, True)
    return chg


def run_force_case(name):
    potname, xc, k, nb, qdep = FORCE_CASES[name]
    chg = si2d_inputs()
    pot = os.path.join(INPUTS, potname)
    w = synth.workload("si2")
    cmd = [DRIVER, "forces", chg, pot, repr(w.encut), xc, str(nb), *map(repr, k)]
    if qdep:
        cmd.append("uspp")
    d = parse(subprocess.run(cmd, check=True, capture_output=True, text=True).stdout)
    g = {"workload": "si2d", "xc": xc, "kpoint": list(k), "nbands": nb, "encut": w.encut, "qdep": qdep,
         "potcar": potname, "chgcar_sha256": synth.sha256(chg), "potcar_sha256": synth.sha256(pot),
         "npw": int(d["npw"][0][0]),
         "eigenvalues": [r[1] for r in d["eig_occ"]], "occupations": [r[2] for r in d["eig_occ"]],
         "smooth_sum": d["smooth_sum"][0][0], "aug_charge": d["aug_charge"][0][0],
         "rho_out_sum": d["rho_out_sum"][0][0],
         # one-based Fortran-order grid index (x fastest), value (cell volume x density)
         "rho_out_samples": [[int(r[0]), r[1]] for r in d["rho_out"]],
         # one-based ion, a, b, occupancy
         "aug_occupancy": [[int(r[0]), int(r[1]), int(r[2]), r[3]] for r in d["aug_occ"]]}
    for key in ("ewald", "local", "nonlocal", "projector", "augmentation", "nlcc", "harris", "total"):
        g["force_" + key] = [r[1:] for r in d["force_" + key]]
    with open(os.path.join(HERE, f"{name}.json"), "w") as fh:
        json.dump(g, fh, indent=1)
    print(name, "total force atom 1", g["force_total"][0])


def run_kat():
    text = subprocess.run([DRIVER, "kat"], check=True, capture_output=True, text=True).stdout
    d = parse(text)
    g = {
        "occ_zero": d["occ_zero"][0], "occ_fermi": d["occ_fermi"][0],
        "occ_gauss": d["occ_gauss"][0], "ewald_kat": d["ewald_kat"][0],
        "ewald_force_kat": [r[1:] for r in d["ewald_force_kat"]],
        "nonlocal_force_kat": d["nonlocal_force_kat"][0],
    }
    with open(os.path.join(HERE, "kat.json"), "w") as fh:
        json.dump(g, fh, indent=1)
    print("kat", g["nonlocal_force_kat"])


if __name__ == "__main__":
    todo = sys.argv[1:] or list(CASES) + ["kat"] + list(FORCE_CASES)
    for c in todo:
        if c == "kat":
            run_kat()
        elif c in FORCE_CASES:
            run_force_case(c)
        else:
            run_case(c)
