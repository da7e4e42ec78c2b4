"""Seeded synthetic inputs for the PAW-operator workloads.

Writes a CHGCAR-format smooth density and a POTCAR-format synthetic PAW
dataset file that the reference readers accept
(reference: src/half_chgcar.F90:12-98 for the density layout,
src/half_potcar.F90:24-205 for the dataset grammar).

Five workloads are defined (sizes are the plane-wave counts the basis
enumeration gives at Gamma or per k):

  si2      diamond Si, 2-atom FCC cell, a=5.43, 24^3 box, 400 eV -> 725 PW
  hfo2     monoclinic HfO2-like 12-atom cell, 48^3 box, 488.6 eV -> 3407 PW
  hfo2mix  the same cell, solved with 96 bands in the mixed-precision form
  si8k     8-atom cubic Si, 40^3 box, 640 eV, Gamma-centred 6x6x6 k mesh
  si216    3x3x3 supercell of si8k (216 atoms), 96^3 box, 245.5 eV

Species data (per dataset):
  projectors   2 radial functions per angular momentum channel,
               f_l(q) = q^l exp(-q^2 r_c^2 / 2) with r_c = 1.0 A for the first
               and r_c = 0.8 A for the second (kept distinct so the two
               channels of one l are linearly independent), 100-point table
  DION         per-l 2x2 random symmetric block, eigenvalues in [-5, 5] eV
  QPAW         per-l 2x2 random symmetric block with spectral norm 0.1
  local part   Gaussian-screened -Z/r: table value
               4 pi Z felect (1 - exp(-q^2 r_s^2 / 4)) / q^2, r_s = 1.0 A,
               so V_loc(q) = table - 4 pi Z felect / q^2 (1000-point table)
  no partial core

The density is a sum of normalised Gaussians (sigma 0.9 A) over periodic
images weighted by the species valence, times (1 + U[-0.01, 0.01]) seeded
noise, renormalised so that the grid mean equals the valence electron count
(the CHGCAR convention: stored value = rho * cell volume).
"""
from __future__ import annotations

import hashlib
import math
import os
from dataclasses import dataclass, field

import numpy as np

# Physical constants in the reference's convention (src/half_constants.F90:6-14).
AUTOA = 0.529177249
RYTOEV = 13.605826
FELECT = 2.0 * RYTOEV * AUTOA
HSQDTM = RYTOEV * AUTOA * AUTOA

PSP_GMAX = 60.0     # local-table extent, covers every |G| of the boxes used
PSPNL_GMAX = 16.0   # projector-table extent, above every basis |k+G|
N_LOCAL = 1000
N_NONLOCAL = 100


@dataclass
class Species:
    name: str
    zval: float
    lmax_channels: tuple      # l of each channel, POTCAR order
    eatom: float
    pomass: float
    rc: tuple = (1.0, 0.8)    # radial widths of the two projectors of one l


SPECIES = {
    "Si": Species("Si", 4.0, (0, 0, 1, 1), -102.5, 28.085),
    "Hf": Species("Hf", 4.0, (0, 0, 1, 1, 2, 2), -248.0, 178.49),
    "O": Species("O", 6.0, (0, 0, 1, 1), -432.0, 16.0),
}


@dataclass
class Workload:
    name: str
    lattice: np.ndarray          # rows are lattice vectors (A)
    species: list                # species names, POTCAR order
    counts: list
    positions: np.ndarray        # fractional, grouped by species
    grid: tuple
    encut: float
    kpoints: np.ndarray          # fractional reciprocal coordinates
    nbands: int
    seed: int
    xc: str = "pbe"
    extra: dict = field(default_factory=dict)

    @property
    def nelect(self) -> float:
        return float(sum(SPECIES[s].zval * c for s, c in zip(self.species, self.counts)))


def _min_image_distance(lattice, fa, fb):
    d = fa - fb
    d -= np.round(d)
    best = 1e9
    for s in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).T.reshape(-1, 3):
        best = min(best, float(np.linalg.norm((d + s) @ lattice)))
    return best


def _diamond8():
    fcc = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]], float)
    return np.vstack([fcc, fcc + 0.25])


def gamma_centered_mesh(n):
    """Gamma-centred full mesh, coordinates i/n wrapped to (-1/2, 1/2]
    (reference: src/half_kpoints.F90:52-68)."""
    pts = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                v = np.array([i, j, k], float) / n
                v = np.where(v > 0.5 + 1e-12, v - 1.0, v)
                pts.append(v)
    return np.array(pts)


def workload(name: str) -> Workload:
    if name == "si2":
        a = 5.43
        lat = 0.5 * a * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], float)
        pos = np.array([[0, 0, 0], [.25, .25, .25]], float)
        return Workload("si2", lat, ["Si"], [2], pos, (24, 24, 24), 400.0,
                        np.zeros((1, 3)), 16, 20260001)
    if name in ("hfo2", "hfo2mix"):
        beta = math.radians(99.2)
        lat = np.array([[5.12, 0, 0], [0, 5.17, 0],
                        [5.29 * math.cos(beta), 0, 5.29 * math.sin(beta)]])
        rng = np.random.default_rng(20260002)
        pos = []
        while len(pos) < 12:
            f = rng.random(3)
            if all(_min_image_distance(lat, f, p) >= 1.8 for p in pos):
                pos.append(f)
        nb = 64 if name == "hfo2" else 96
        # the mixed-precision workload solves the same input files
        return Workload(name, lat, ["Hf", "O"], [4, 8], np.array(pos), (48, 48, 48),
                        488.6, np.zeros((1, 3)), nb, 20260002, extra={"inputs": "hfo2"})
    if name == "si8k":
        a = 5.43
        lat = a * np.eye(3)
        return Workload("si8k", lat, ["Si"], [8], _diamond8(), (40, 40, 40), 640.0,
                        gamma_centered_mesh(6), 32, 20260004)
    if name == "si216":
        a = 5.43
        L = 3 * a
        lat = L * np.eye(3)
        base = _diamond8()
        frac = []
        for i in range(3):
            for j in range(3):
                for k in range(3):
                    frac.extend((base + np.array([i, j, k])) / 3.0)
        frac = np.array(frac)
        rng = np.random.default_rng(20260005)
        frac = frac + rng.uniform(-0.03, 0.03, frac.shape) / L
        return Workload("si216", lat, ["Si"], [216], frac, (96, 96, 96), 245.5,
                        np.zeros((1, 3)), 512, 20260005)
    raise KeyError(name)


WORKLOADS = ("si2", "hfo2", "hfo2mix", "si8k", "si216")


def reciprocal(lattice):
    return 2.0 * math.pi * np.linalg.inv(lattice).T


def smooth_density(w: Workload) -> np.ndarray:
    """CHGCAR values (rho * volume) on the C-order (n1, n2, n3) box."""
    n1, n2, n3 = w.grid
    brec = reciprocal(w.lattice)
    f = [np.fft.fftfreq(n, 1.0 / n) for n in w.grid]
    g1, g2, g3 = np.meshgrid(f[0], f[1], f[2], indexing="ij")
    gvec = (np.stack([g1, g2, g3], -1).reshape(-1, 3)) @ brec
    g2n = np.einsum("ij,ij->i", gvec, gvec)
    sigma = 0.9
    env = np.exp(-0.5 * g2n * sigma * sigma)
    rho_g = np.zeros(n1 * n2 * n3, complex)
    cart = w.positions @ w.lattice
    ion = 0
    for s, c in zip(w.species, w.counts):
        z = SPECIES[s].zval
        for _ in range(c):
            rho_g += z * np.exp(-1j * (gvec @ cart[ion]))
            ion += 1
    rho_g *= env
    rho = np.fft.ifftn(rho_g.reshape(n1, n2, n3)).real * (n1 * n2 * n3)
    rng = np.random.default_rng(w.seed)
    rho = rho * (1.0 + rng.uniform(-0.01, 0.01, rho.shape))
    rho *= w.nelect / rho.mean()
    return rho


def _fmt_block(values, per_line=5, fmt="%.12E"):
    out = []
    vals = list(values)
    for i in range(0, len(vals), per_line):
        out.append(" " + " ".join(fmt % v for v in vals[i:i + per_line]))
    return "\n".join(out)


def write_chgcar(path: str, w: Workload, header_fmt: str = "%.12f", rho=None) -> None:
    """header_fmt formats the lattice and position numbers of the header;
    rho (C-order box of CHGCAR values) defaults to smooth_density(w)."""
    rho = smooth_density(w) if rho is None else rho
    lines = [f"synthetic {w.name}", "1.0"]
    for v in w.lattice:
        lines.append("  " + "  ".join(header_fmt % x for x in v))
    lines.append("  " + "  ".join(w.species))
    lines.append("  " + "  ".join(str(c) for c in w.counts))
    lines.append("Direct")
    for p in w.positions:
        lines.append("  " + "  ".join(header_fmt % x for x in p))
    lines.append("")
    lines.append(" %d %d %d" % w.grid)
    # x-fastest order, as CHGCAR stores the grid
    flat = rho.ravel(order="F")
    body = np.array(flat).reshape(-1)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
        full = (len(body) // 5) * 5
        if full:
            np.savetxt(fh, body[:full].reshape(-1, 5), fmt="%.12E", delimiter=" ")
        if full < len(body):
            fh.write(" ".join("%.12E" % v for v in body[full:]) + "\n")


def species_tables(sp: Species, seed: int):
    """DION and QPAW (channel x channel) plus the radial tables of one dataset."""
    rng = np.random.default_rng(seed)
    nch = len(sp.lmax_channels)
    dion = np.zeros((nch, nch))
    qpaw = np.zeros((nch, nch))
    for l in sorted(set(sp.lmax_channels)):
        idx = [i for i, ll in enumerate(sp.lmax_channels) if ll == l]
        k = len(idx)
        rot, _ = np.linalg.qr(rng.standard_normal((k, k)))
        lam = rng.uniform(-5.0, 5.0, k)
        dblk = rot @ np.diag(lam) @ rot.T
        qr, _ = np.linalg.qr(rng.standard_normal((k, k)))
        mu = rng.uniform(-1.0, 1.0, k)
        mu *= 0.1 / np.max(np.abs(mu))
        qblk = qr @ np.diag(mu) @ qr.T
        for a, ia in enumerate(idx):
            for b, ib in enumerate(idx):
                dion[ia, ib] = 0.5 * (dblk[a, b] + dblk[b, a])
                qpaw[ia, ib] = 0.5 * (qblk[a, b] + qblk[b, a])
    hq = PSPNL_GMAX / N_NONLOCAL
    q = np.arange(N_NONLOCAL) * hq
    radial = []
    seen = {}
    for l in sp.lmax_channels:
        j = seen.get(l, 0)
        seen[l] = j + 1
        rc = sp.rc[j % len(sp.rc)]
        radial.append(q ** l * np.exp(-0.5 * q * q * rc * rc))
    hl = PSP_GMAX / N_LOCAL
    ql = np.arange(N_LOCAL) * hl
    rs = 1.0
    with np.errstate(divide="ignore", invalid="ignore"):
        loc = 4.0 * math.pi * sp.zval * FELECT * (-np.expm1(-ql * ql * rs * rs / 4.0)) / (ql * ql)
    loc[0] = 4.0 * math.pi * sp.zval * FELECT * rs * rs / 4.0
    rho_atom = sp.zval * np.exp(-0.5 * ql * ql * 0.9 * 0.9)
    return dion, qpaw, radial, loc, rho_atom


PAW_NMAX = 256      # radial grid points of the partial-wave sets
PAW_RMAX = 1.5      # augmentation sphere radius (A)


def partial_waves(sp: Species, seed: int):
    """Radial grid and r-weighted partial waves of one dataset.

    Logarithmic grid from 1e-3 A to 2 A.  Pseudo wave of a channel with
    angular momentum l: r^(l+1) exp(-r^2 / (2 rc^2)); the all-electron wave
    adds b r^(l+1) (1 - (r/PAW_RMAX)^2)^3 inside the augmentation sphere
    (b uniform in [-0.5, 0.5], seeded), so their difference vanishes smoothly
    at PAW_RMAX."""
    rng = np.random.default_rng(seed + 7919)
    r = 1e-3 * (2.0 / 1e-3) ** (np.arange(PAW_NMAX) / (PAW_NMAX - 1))
    wps, wae = [], []
    seen = {}
    for l in sp.lmax_channels:
        j = seen.get(l, 0)
        seen[l] = j + 1
        rc = sp.rc[j % len(sp.rc)]
        ps = r ** (l + 1) * np.exp(-0.5 * r * r / (rc * rc))
        bump = np.where(r < PAW_RMAX, r ** (l + 1) * (1.0 - (r / PAW_RMAX) ** 2) ** 3, 0.0)
        wps.append(ps)
        wae.append(ps + rng.uniform(-0.5, 0.5) * bump)
    return r, wps, wae


def core_table(zcore: float = 1.0, rc: float = 0.7) -> np.ndarray:
    """Partial core charge table on the local-part grid q_i = gmax i / 1000:
    zcore exp(-q^2 rc^2 / 4), the transform of a normalised Gaussian core
    holding zcore electrons (the table's q = 0 value is the core electron
    count in the CHGCAR convention)."""
    q = PSP_GMAX * np.arange(1000) / 1000.0
    return zcore * np.exp(-q * q * rc * rc / 4.0)


def write_potcar(path: str, w: Workload, with_partial_waves: bool = False, partial_core: bool = False) -> None:
    """POTCAR-format datasets; with_partial_waves adds the radial grid and the
    pseudo / all-electron partial waves the potential-dependent D correction
    (QDEP) integrates, otherwise the PAW section carries an 8-point stub;
    partial_core adds a Gaussian partial core charge (core_table)."""
    chunks = []
    for it, s in enumerate(w.species):
        sp = SPECIES[s]
        dion, qpaw, radial, loc, rho_atom = species_tables(sp, w.seed * 10 + it)
        nch = len(sp.lmax_channels)
        c = []
        c.append(f"  PAW_PBE {s} 01Jan2026")
        c.append(f"  {sp.zval:.14f}")
        c.append(" parameters from PSCTR are:")
        c.append(f"   VRHFIN ={s}: synthetic")
        c.append("   LEXCH  = PE")
        c.append(f"   EATOM  = {sp.eatom:12.4f} eV,  {sp.eatom / RYTOEV:12.4f} Ry")
        c.append(f"   POMASS = {sp.pomass:8.3f}; ZVAL   = {sp.zval:8.3f}    mass and valenz")
        c.append("   ENMAX  =  250.000; ENMIN  =  200.000 eV")
        c.append(" END of PSCTR-controll parameters")
        c.append(" local part")
        c.append(f"  {PSP_GMAX:.10f}")
        c.append(_fmt_block(loc))
        c.append(" gradient corrections used for XC")
        c.append("         4")
        if partial_core:
            c.append(" core charge-density (partial)")
            c.append(_fmt_block(core_table()))
        c.append(" atomic pseudo charge-density")
        c.append(_fmt_block(rho_atom))
        c.append(f"  {PSPNL_GMAX:.10f}  T")
        ls = list(sp.lmax_channels)
        ich = 0
        while ich < nch:
            l = ls[ich]
            # consecutive channels with the same l form one block
            npro = 0
            while ich + npro < nch and ls[ich + npro] == l:
                npro += 1
            c.append(" Non local part")
            c.append(f"  {l}  {npro}  1.50000000")
            blk = dion[ich:ich + npro, ich:ich + npro]
            c.append(_fmt_block(blk.ravel()))
            for j in range(npro):
                c.append(" Reciprocal Space Part")
                c.append(_fmt_block(radial[ich + j]))
                c.append(" Real Space Part")
                c.append(_fmt_block(np.zeros(N_NONLOCAL)))
            ich += npro
        c.append(" PAW radial sets")
        if with_partial_waves:
            rgrid, wps, wae = partial_waves(sp, w.seed * 10 + it)
        else:
            rgrid = np.linspace(0.01, 1.5, 8)
            wps = wae = [np.zeros(8)] * nch
        nmax = rgrid.size
        c.append(f"  {nmax}  {PAW_RMAX:.8f}")
        c.append(" (5E20.12)")
        c.append(" augmentation charges (non sperical)")
        c.append(_fmt_block(qpaw.ravel()))
        c.append(" grid")
        c.append(_fmt_block(rgrid))
        c.append(" aepotential")
        c.append(_fmt_block(np.zeros(nmax)))
        c.append(" core charge-density")
        c.append(_fmt_block(np.zeros(nmax)))
        for j in range(nch):
            c.append(" pseudo wavefunction")
            c.append(_fmt_block(wps[j]))
            c.append(" ae wavefunction")
            c.append(_fmt_block(wae[j]))
        c.append(" End of Dataset")
        chunks.append("\n".join(c))
    with open(path, "w") as fh:
        fh.write("\n".join(chunks) + "\n")


def sha256(path: str) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as fh:
        for blk in iter(lambda: fh.read(1 << 20), b""):
            h.update(blk)
    return h.hexdigest()


def make_inputs(name: str, outdir: str, partial_waves: bool = False, partial_core: bool = False):
    """Write CHGCAR/POTCAR for a workload into outdir (cached by content).
    partial_waves selects the datasets with PAW partial waves (file
    <name>.paw.POTCAR), needed by the potential-dependent D correction;
    partial_core adds a partial core charge (<name>[.paw].core.POTCAR)."""
    w = workload(name)
    os.makedirs(outdir, exist_ok=True)
    tag = w.extra.get("inputs", w.name)
    chg = os.path.join(outdir, f"{tag}.CHGCAR")
    pot = os.path.join(outdir, tag + (".paw" if partial_waves else "") + (".core" if partial_core else "") + ".POTCAR")
    if not os.path.exists(chg):
        tmp = chg + f".tmp{os.getpid()}"
        write_chgcar(tmp, w)
        os.replace(tmp, chg)
    if not os.path.exists(pot):
        tmp = pot + f".tmp{os.getpid()}"
        write_potcar(tmp, w, partial_waves, partial_core)
        os.replace(tmp, pot)
    return chg, pot, w


# --------------------------------------------------------------------------
# band folding: a perfect supercell and its primitive cell
# --------------------------------------------------------------------------

def write_folded_inputs(outdir: str, xc_seed: int = 20260006):
    """Inputs of a band-folding pair at the 216-atom workload's sizes: the
    3x3x3 supercell of the 8-atom Si cell without positional noise (216
    atoms, L = 16.29 A, 96^3 box) and the 8-atom cell (a = 5.43 A, 32^3 box),
    the supercell's density the 3x3x3 tiling of the primitive cell's
    (Gaussian sum with the seeded +-1% noise), one POTCAR for both.  At
    Gamma the supercell's spectrum is the union of the primitive cell's
    spectra at the 27 k points (i, j, l)/3, i, j, l in {0, 1, 2}: supercell
    G = 3 G' + (i, j, l) in the primitive reciprocal basis, and both boxes
    enumerate the same vectors.  Returns (supercell CHGCAR, primitive
    CHGCAR, POTCAR, k points, supercell workload)."""
    a = 5.43
    base = _diamond8()
    prim = Workload("si8f", a * np.eye(3), ["Si"], [8], base, (32, 32, 32), 245.5, np.zeros((1, 3)), 40, xc_seed)
    frac = []
    for i in range(3):
        for j in range(3):
            for k in range(3):
                frac.extend((base + np.array([i, j, k])) / 3.0)
    sup = Workload("si216f", 3 * a * np.eye(3), ["Si"], [216], np.array(frac), (96, 96, 96), 245.5,
                   np.zeros((1, 3)), 512, xc_seed)
    os.makedirs(outdir, exist_ok=True)
    rho_p = smooth_density(prim)
    # CHGCAR values are rho x cell volume: the supercell holds 27 cells
    rho_s = np.tile(rho_p, (3, 3, 3)) * 27.0
    paths = []
    for w, rho in ((sup, rho_s), (prim, rho_p)):
        path = os.path.join(outdir, f"{w.name}.CHGCAR")
        if not os.path.exists(path):
            tmp = path + f".tmp{os.getpid()}"
            write_chgcar(tmp, w, rho=rho)
            os.replace(tmp, path)
        paths.append(path)
    pot = os.path.join(outdir, "si216f.POTCAR")
    if not os.path.exists(pot):
        tmp = pot + f".tmp{os.getpid()}"
        write_potcar(tmp, sup)
        os.replace(tmp, pot)
    ks = np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(3)], float) / 3.0
    return paths[0], paths[1], pot, ks, sup


# --------------------------------------------------------------------------
# lattice-precision boundary case (request geometry)
# --------------------------------------------------------------------------

def _kinetic(lattice, grid, kpoint):
    """G vectors of the box (n_i in [-n/2, (n-1)/2], n1 outermost) and
    hsqdtm |(G + k) Brec|^2 on `lattice` (reference: src/half_basis.F90:9-55)."""
    r = [np.arange(-n // 2, (n - 1) // 2 + 1) for n in grid]
    a, b, c = np.meshgrid(*r, indexing="ij")
    g = np.stack([a.ravel(), b.ravel(), c.ravel()], 1)
    q = (g + np.asarray(kpoint, float)) @ reciprocal(lattice)
    return g, HSQDTM * np.einsum("ij,ij->i", q, q)


def boundary_lattice(lattice, grid, encut, kpoint, digits=6, delta=1e-9):
    """A uniformly scaled copy s * lattice, kept at full double precision, on
    which one plane wave sits just above the cutoff (hsqdtm |k+G|^2 =
    encut (1 + delta), so the strict test excludes it) while the same lattice
    rounded to `digits` decimals, the precision of a CHGCAR header, puts it
    below: the rounded cell's basis holds extra G vectors at the unchanged
    ENCUT.  Every other vector stays at least delta/2 (full) and delta/10
    (rounded) away from the cutoff, relative, so both counts are exact in any
    evaluation order.  Returns (full lattice, rounded lattice, count on the
    full lattice, count on the rounded lattice)."""
    g, e1 = _kinetic(np.asarray(lattice, float), grid, kpoint)
    order = np.argsort(np.abs(e1 - encut))
    for c in order[:256]:
        s = math.sqrt(e1[c] / (encut * (1.0 + delta)))
        full = s * np.asarray(lattice, float)
        rounded = np.round(full, digits)
        _, ef = _kinetic(full, grid, kpoint)
        _, er = _kinetic(rounded, grid, kpoint)
        rel_f = np.abs(ef / encut - 1.0)
        rel_f[c] = np.inf
        if rel_f.min() < 0.5 * delta or np.abs(er / encut - 1.0).min() < 0.1 * delta:
            continue
        nf, nr = int((ef < encut).sum()), int((er < encut).sum())
        if nr > nf:
            return full, rounded, nf, nr
    raise ValueError("no boundary scaling found")


def write_boundary_inputs(outdir: str, name: str = "hfo2", encut: float = 520.0, kpoint=(0.0, 0.5, 0.5),
                          digits: int = 6):
    """A cell whose CHGCAR header precision changes its plane-wave basis.

    The workload's lattice is scaled by boundary_lattice and its fractional
    positions displaced by up to 1e-7 (seeded), both kept at full precision.
    Two CHGCAR files hold the same density values: <name>.round.CHGCAR with
    the lattice and positions written to `digits` decimals (as a DeePAW
    density file stores them) and <name>.full.CHGCAR with them written to 17
    significant digits (they read back as the same doubles).  The POTCAR is
    the workload's.  Returns a dict with both files, the full-precision
    lattice and positions, the one-based type of every ion, the grid, ENCUT,
    k point and the basis sizes on the full and rounded lattices."""
    w = workload(name)
    full, rounded, nf, nr = boundary_lattice(w.lattice, w.grid, encut, kpoint, digits)
    rng = np.random.default_rng(w.seed + 17)
    pos = w.positions + rng.uniform(-1e-7, 1e-7, w.positions.shape)
    wf = Workload(w.name, full, w.species, w.counts, pos, w.grid, encut, np.array([kpoint], float), w.nbands,
                  w.seed, w.xc, dict(w.extra))
    os.makedirs(outdir, exist_ok=True)
    paths = {}
    for tag, fmt in (("round", "%%.%df" % digits), ("full", "%.17g")):
        p = os.path.join(outdir, f"{name}.{tag}.CHGCAR")
        tmp = p + f".tmp{os.getpid()}"
        write_chgcar(tmp, wf, header_fmt=fmt)
        os.replace(tmp, p)
        paths[tag] = p
    pot = os.path.join(outdir, f"{name}.boundary.POTCAR")
    if not os.path.exists(pot):
        tmp = pot + f".tmp{os.getpid()}"
        write_potcar(tmp, w)
        os.replace(tmp, pot)
    types = np.concatenate([np.full(c, i + 1, np.int32) for i, c in enumerate(w.counts)])
    return {"chgcar_round": paths["round"], "chgcar_full": paths["full"], "potcar": pot, "lattice": full,
            "lattice_round": rounded, "positions": pos, "species": types, "grid": w.grid, "encut": encut,
            "kpoint": tuple(kpoint), "npw_full": nf, "npw_round": nr}
