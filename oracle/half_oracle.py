"""CPU restatement (numpy, FP64) of the reference's PAW-operator path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library imports, links or
calls this module; it is the checker used by tests/, by the smoke check and by
the CPU-baseline leg of bench.py.

Every function cites the reference routine it restates (path:line into the
reference tree).  Arithmetic recipes are kept exactly (spline end conditions,
cut-offs, finite-difference PBE derivatives, FFT sign and normalisation);
loops are vectorised with numpy where that does not change the arithmetic
beyond summation order.

Pinning (what checks each function besides retyping its formula):
  read_chgcar/read_potcar   golden inputs parse to the reference's volume,
                            electron count and energies (tests/golden)
  build_basis               brute-force count + golden npw at two k points
  spline_*, real_ylm        closed form f_l(q) (spline error bound), the
                            addition theorem sum_m Y_lm^2 = (2l+1)/4pi,
                            and golden H/S elements built from them
  build_projectors          golden H/S elements and golden apply images
  build_veff (LDA, PBE)     golden Veff grids and e_H/e_xc/e_xc_pot
  assemble_dense, apply_hs  golden H/S samples, golden apply images, and
                            matrix-free vs dense agreement
  solve_dense               golden eigenvalues; constant-Veff shift and
                            Veff=0/D=Q=0 kinetic-spectrum special cases
  nonlocal_forces           golden forces and the reference KAT
                            (app/half_energy_check.F90:58-81) + central FD
  occupations, ewald_*      reference KAT values (kat.json) + eta invariance
  acc_solve                 converged eigenvalues vs solve_dense
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass

import numpy as np

# src/half_constants.F90:6-14
PI = math.pi
AUTOA = 0.529177249
RYTOEV = 13.605826
HATOEV = 2.0 * RYTOEV
FELECT = 2.0 * RYTOEV * AUTOA
EDEPS = 4.0 * PI * FELECT
HSQDTM = RYTOEV * AUTOA * AUTOA


@dataclass
class Crystal:
    """src/half_types.F90:13-25 (crystal_t) with update_geometry :59-63."""
    lattice: np.ndarray
    species: list
    counts: list
    positions: np.ndarray  # fractional, nions x 3, grouped by species

    @property
    def volume(self):
        return abs(float(np.linalg.det(self.lattice)))

    @property
    def reciprocal(self):
        return 2.0 * PI * np.linalg.inv(self.lattice).T

    @property
    def nions(self):
        return int(sum(self.counts))


# --------------------------------------------------------------------------
# readers (src/half_chgcar.F90:12-98, src/half_potcar.F90:24-205)
# --------------------------------------------------------------------------

def read_chgcar(path):
    """src/half_chgcar.F90:12-98: POSCAR header then NGX*NGY*NGZ values
    (x fastest).  Returns (crystal, grid, values)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    scale = float(lines[1].split()[0])
    lat = np.array([[float(x) for x in lines[2 + i].split()[:3]] for i in range(3)]) * scale
    words = lines[5].split()
    pos = 6
    if all(re.fullmatch(r"[+-]?\d+", w) for w in words):
        species = [f"Type{i+1}" for i in range(len(words))]
        counts = [int(w) for w in words]
    else:
        species = words
        counts = [int(w) for w in lines[6].split()[:len(words)]]
        pos = 7
    mode = lines[pos].strip()
    if mode[:1].lower() == "s":
        pos += 1
        mode = lines[pos].strip()
    cart = mode[:1].lower() in ("c", "k")
    pos += 1
    nions = sum(counts)
    frac = np.array([[float(x) for x in lines[pos + i].split()[:3]] for i in range(nions)])
    if cart:
        frac = frac @ np.linalg.inv(lat)
    pos += nions
    while not lines[pos].strip():
        pos += 1
    grid = tuple(int(x) for x in lines[pos].split()[:3])
    n = grid[0] * grid[1] * grid[2]
    vals = np.array(" ".join(lines[pos + 1:]).split()[:n], dtype=float)
    return Crystal(lat, species, counts, frac), grid, vals


def read_potcar(path):
    """src/half_potcar.F90:24-205, restricted to the fields the hot path uses."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    out = []
    i = 0
    n = len(lines)

    def nonblank(j):
        while j < n and not lines[j].strip():
            j += 1
        return j

    def floats(j, count):
        vals = []
        while len(vals) < count:
            vals.extend(float(t) for t in lines[j].split())
            j += 1
        return np.array(vals[:count]), j

    while True:
        i = nonblank(i)
        if i >= n:
            break
        if "End of Dataset" in lines[i]:
            i += 1
            continue
        head = lines[i].split()
        p = {"element": head[1] if len(head) > 1 else "", "eatom": 0.0, "has_core": False}
        i += 1
        p["zval"] = float(lines[i].split()[0])
        i += 1
        while "end of psctr" not in lines[i].lower():
            ln = lines[i]
            m = re.search(r"EATOM\s*=\s*([-+0-9.Ee]+)", ln)
            if m:
                p["eatom"] = float(m.group(1))
            m = re.search(r"ZVAL\s*=\s*([-+0-9.Ee]+)", ln)
            if m:
                p["zval"] = float(m.group(1))
            m = re.search(r"VRHFIN\s*=\s*([^\s:;]+)", ln)
            if m:
                p["element"] = m.group(1)
            i += 1
        i = nonblank(i + 1)  # 'local part'
        i += 1
        p["psp_gmax"] = float(lines[i].split()[0])
        p["psp_local"], i = floats(i + 1, 1000)
        while True:
            i = nonblank(i)
            key = lines[i].lower()
            if "gradient" in key:
                i += 2
            elif "core charge-density (partial)" in key:
                p["pspcor"], i = floats(i + 1, 1000)
                p["has_core"] = True
            elif "kinetic energy density" in key:
                _, i = floats(i + 1, 1000)
            elif "atomic pseudo charge-density" in key:
                _, i = floats(i + 1, 1000)
                break
            else:
                raise ValueError("unknown POTCAR local selector: " + lines[i])
        i = nonblank(i)
        p["pspnl_gmax"] = float(lines[i].split()[0])
        i += 1
        lps, dion_blocks, tables = [], [], []
        while True:
            i = nonblank(i)
            key = lines[i].lower()
            if "paw radial sets" in key:
                break
            if "non local part" not in key:
                raise ValueError("expected non-local part")
            t = lines[i + 1].split()
            lval, npro = int(t[0]), int(t[1])
            flat, i = floats(i + 2, npro * npro)
            dion_blocks.append(flat.reshape(npro, npro))
            for _ in range(npro):
                i = nonblank(i)
                tab, i = floats(i + 1, 100)
                tables.append(tab)
                i = nonblank(i)
                _, i = floats(i + 1, 100)
                lps.append(lval)
        nch = len(lps)
        dion = np.zeros((nch, nch))
        o = 0
        for blk in dion_blocks:
            k = blk.shape[0]
            dion[o:o + k, o:o + k] = blk
            o += k
        i += 1
        t = lines[i].split()  # nmax, paw_rmax
        nmax, paw_rmax = int(t[0]), float(t[1])
        i += 1  # format line
        i = nonblank(i + 1)
        flat, i = floats(i + 1, nch * nch)
        p.update(lps=lps, dion=dion, pspnl=np.array(tables).T, qpaw=flat.reshape(nch, nch),
                 paw_rmax=paw_rmax, rgrid=np.zeros(nmax), wae=np.zeros((nmax, nch)), wps=np.zeros((nmax, nch)))
        # PAW radial selectors (src/half_potcar.F90:171-205): keep the grid and
        # the partial waves, skip the rest; the partial waves end the section
        while True:
            i = nonblank(i)
            key = lines[i].strip().lower()
            if "occupancies in atom" in key or "uccopancies in atom" in key:
                _, i = floats(i + 1, nch * nch)
            elif key == "grid":
                p["rgrid"], i = floats(i + 1, nmax)
            elif key in ("aepotential",) or "core charge-density" in key or "local pseudopotential core" in key \
                    or "pspotential valence only" in key or "kinetic energy-density" in key:
                _, i = floats(i + 1, nmax)
            elif "pseudo wavefunction" in key:
                for ich in range(nch):
                    i = nonblank(i)
                    p["wps"][:, ich], i = floats(i + 1, nmax)
                    i = nonblank(i)
                    p["wae"][:, ich], i = floats(i + 1, nmax)
                break
            elif key.startswith("t"):
                _, i = floats(i + 1, nch * nch)
            else:
                raise ValueError("unknown POTCAR PAW selector: " + lines[i])
        while i < n and "End of Dataset" not in lines[i]:
            i += 1
        out.append(p)
    return out


# --------------------------------------------------------------------------
# basis (src/half_basis.F90:9-55)
# --------------------------------------------------------------------------

@dataclass
class Basis:
    grid: tuple
    gvec: np.ndarray      # N x 3 int
    q: np.ndarray         # N x 3 (k+G) cartesian
    kinetic: np.ndarray   # N
    fft_index: np.ndarray  # N, zero-based C-order box index


def build_basis(crystal, grid, encut, kpoint):
    """src/half_basis.F90:9-55: n_i in [-n/2, (n-1)/2], n1 outermost, keep
    hsqdtm |q|^2 < ENCUT (strict)."""
    r = [np.arange(-n // 2, (n - 1) // 2 + 1) for n in grid]
    a, b, c = np.meshgrid(*r, indexing="ij")
    nvec = np.stack([a.ravel(), b.ravel(), c.ravel()], 1)
    q = (nvec + np.asarray(kpoint, float)) @ crystal.reciprocal
    kin = HSQDTM * np.einsum("ij,ij->i", q, q)
    keep = kin < encut
    g = nvec[keep]
    idx = (np.mod(g[:, 0], grid[0]) * grid[1] * grid[2] + np.mod(g[:, 1], grid[1]) * grid[2]
           + np.mod(g[:, 2], grid[2]))
    return Basis(tuple(grid), g.astype(np.int32), q[keep], kin[keep], idx.astype(np.int64))


def vasp_order(basis):
    """src/half_vaspwave.F90:142-155: basis indices in the order VASP stores
    plane-wave coefficients: a box-sized lookup (C-order index -> basis
    index) walked with the z box coordinate slowest and x fastest."""
    n1, n2, n3 = basis.grid
    lookup = np.full(n1 * n2 * n3, -1, np.int64)
    lookup[basis.fft_index] = np.arange(len(basis.fft_index))
    gz, gy, gx = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    walk = lookup[(gx * n2 * n3 + gy * n3 + gz).ravel()]
    return walk[walk >= 0]


# --------------------------------------------------------------------------
# splines and real spherical harmonics (src/half_paw.F90:78-137)
# --------------------------------------------------------------------------

def spline_second(y, xmax):
    """src/half_paw.F90:78-99 (same recipe as src/half_potential.F90:260-278)."""
    y = np.asarray(y, float)
    n = y.size
    h = xmax / n
    diag = np.zeros(n + 1)
    rhs = np.zeros(n + 1)
    upper = np.zeros(n + 1)
    # one-based arrays
    yy = np.concatenate([[0.0], y])
    diag[2] = 1.0
    rhs[2] = ((yy[3] - yy[2]) / h - (yy[2] - yy[1]) / h) / h
    for i in range(3, n - 1):
        diag[i] = 4 * h
        upper[i] = h
        rhs[i] = 6 * ((yy[i + 1] - yy[i]) / h - (yy[i] - yy[i - 1]) / h)
    diag[n - 1] = 1.0
    rhs[n - 1] = ((yy[n] - yy[n - 1]) / h - (yy[n - 1] - yy[n - 2]) / h) / h
    for i in range(3, n):
        w = (h if i <= n - 2 else 0.0) / diag[i - 1]
        diag[i] = diag[i] - w * upper[i - 1]
        rhs[i] = rhs[i] - w * rhs[i - 1]
    m2 = np.zeros(n + 1)
    m2[n - 1] = rhs[n - 1] / diag[n - 1]
    for i in range(n - 2, 1, -1):
        m2[i] = (rhs[i] - upper[i] * m2[i + 1]) / diag[i]
    m2[1] = 2 * m2[2] - m2[3]
    m2[n] = 2 * m2[n - 1] - m2[n - 2]
    return m2[1:]


def spline_eval(y, m2, xmax, x):
    """src/half_paw.F90:101-107, vectorised over x."""
    y = np.asarray(y, float)
    n = y.size
    h = xmax / n
    x = np.asarray(x, float)
    k = np.clip(np.floor(x / h).astype(np.int64) + 1, 1, n - 1)  # one-based
    t = (x - (k - 1) * h) / h
    yk, yk1 = y[k - 1], y[k]
    mk, mk1 = m2[k - 1], m2[k]
    return (1 - t) * yk + t * yk1 + h * h * (((1 - t) ** 3 - (1 - t)) * mk + (t ** 3 - t) * mk1) / 6


def _factorial(n):
    v = 1.0
    for i in range(2, n + 1):
        v *= i
    return v


def assoc_legendre(l, m, x):
    """src/half_paw.F90:124-137 (Condon-Shortley sign)."""
    x = np.asarray(x, float)
    pmm = np.ones_like(x)
    if m > 0:
        fact = 1.0
        s = np.sqrt(np.maximum(0.0, (1 - x) * (1 + x)))
        for _ in range(m):
            pmm = -pmm * fact * s
            fact += 2
    if l == m:
        return pmm
    pmmp1 = x * (2 * m + 1) * pmm
    if l == m + 1:
        return pmmp1
    for ll in range(m + 2, l + 1):
        pll = ((2 * ll - 1) * x * pmmp1 - (ll + m - 1) * pmm) / (ll - m)
        pmm, pmmp1 = pmmp1, pll
    return pmmp1


def real_ylm(l, m, theta, phi):
    """src/half_paw.F90:109-123."""
    am = abs(m)
    x = np.cos(theta)
    p = assoc_legendre(l, am, x)
    norm = math.sqrt((2 * l + 1) / (4 * PI) * _factorial(l - am) / _factorial(l + am))
    if m > 0:
        return math.sqrt(2.0) * (-1.0) ** m * norm * p * np.cos(am * phi)
    if m < 0:
        return math.sqrt(2.0) * (-1.0) ** m * norm * p * np.sin(am * phi)
    return norm * p


# --------------------------------------------------------------------------
# projectors and lm-expanded D/Q (src/half_paw.F90:18-76)
# --------------------------------------------------------------------------

def lm_expand(p):
    """src/half_paw.F90:39-51: D(ilm,jlm)=DION(ich,jch) iff l and m match."""
    lps = p["lps"]
    nlm = sum(2 * l + 1 for l in lps)
    d = np.zeros((nlm, nlm))
    q = np.zeros((nlm, nlm))
    lm = []
    for ich, l in enumerate(lps):
        for m in range(-l, l + 1):
            lm.append((ich, l, m))
    for a, (ich, l, m) in enumerate(lm):
        for b, (jch, lj, mj) in enumerate(lm):
            if l == lj and m == mj:
                d[a, b] = p["dion"][ich, jch]
                q[a, b] = p["qpaw"][ich, jch]
    return d, q, lm


def build_projectors(potcars, crystal, basis):
    """src/half_paw.F90:18-76: beta(iat,lm,G) = Omega^-1/2 (-i)^l f_l(|q|)
    Y_lm(q) exp(-i q.R), f_l = spline of the pspnl table, zero above
    0.99 gmax; theta=phi=0 for |q| < 1e-10."""
    pref = 1.0 / math.sqrt(crystal.volume)
    q = basis.q
    qn = np.sqrt(np.einsum("ij,ij->i", q, q))
    big = qn > 1e-10
    theta = np.where(big, np.arccos(np.clip(np.where(big, q[:, 2] / np.where(big, qn, 1.0), 1.0), -1, 1)), 0.0)
    phi = np.where(big, np.arctan2(q[:, 1], q[:, 0]), 0.0)
    out = []
    ion0 = 0
    for it, p in enumerate(potcars):
        d, qq, lm = lm_expand(p)
        nat = crystal.counts[it]
        nlm = len(lm)
        rpos = crystal.positions[ion0:ion0 + nat] @ crystal.lattice
        phase = q @ rpos.T  # N x nat
        ephase = np.cos(phase) - 1j * np.sin(phase)
        beta = np.zeros((nat, nlm, q.shape[0]), complex)
        for ilm, (ich, l, m) in enumerate(lm):
            tab = p["pspnl"][:, ich]
            m2 = spline_second(tab, p["pspnl_gmax"])
            rval = spline_eval(tab, m2, p["pspnl_gmax"], qn)
            rval = np.where(qn > 0.99 * p["pspnl_gmax"], 0.0, rval)
            ylm = real_ylm(l, m, theta, phi)
            ilfac = (-1j) ** l
            beta[:, ilm, :] = (pref * ilfac * rval * ylm)[None, :] * ephase.T
        out.append({"natoms": nat, "nlm": nlm, "beta": beta, "dij": d, "qij": qq,
                    "dij_atom": np.repeat(d[None], nat, 0), "ddij_atom": np.zeros((nat, nlm, nlm, 3)), "lm": lm})
        ion0 += nat
    return out


# --------------------------------------------------------------------------
# potential-dependent D correction, QDEP (src/half_uspp.F90:112-397)
# --------------------------------------------------------------------------

def radial_weights(r):
    """src/half_uspp.F90:296-311: composite Simpson weights on a non-uniform
    grid (pairs of intervals), with a three-point end correction for an
    even point count."""
    n = r.size
    w = np.zeros(n)
    last = n if n % 2 == 1 else n - 1
    for i in range(0, last - 2, 2):
        h0, h1 = r[i + 1] - r[i], r[i + 2] - r[i + 1]
        sc = (h0 + h1) / 6.0
        w[i] += sc * (2 - h1 / h0)
        w[i + 1] += sc * (h0 + h1) ** 2 / (h0 * h1)
        w[i + 2] += sc * (2 - h0 / h1)
    if n % 2 == 0:
        h0, h1 = r[n - 2] - r[n - 3], r[n - 1] - r[n - 2]
        w[n - 1] += (2 * h1 * h1 + 3 * h0 * h1) / (6 * (h0 + h1))
        w[n - 2] += (h1 * h1 + 3 * h0 * h1) / (6 * h0)
        w[n - 3] -= h1 ** 3 / (6 * h0 * (h0 + h1))
    return w


def sph_bessel(l, x):
    """src/half_uspp.F90:313-323: power series for |x| < l + 2, upward
    recurrence from j0, j1 otherwise (scalar)."""
    if abs(x) < l + 2:
        den = 1.0
        for i in range(1, l + 1):
            den *= 2 * i + 1
        x2 = x * x
        term = x ** l / den
        v = term
        for k in range(31):
            term = -term * x2 / (2.0 * (k + 1) * (2 * l + 2 * k + 3))
            v += term
            if abs(term) < 1e-16 * max(1.0, abs(v)):
                break
        return v
    jm = math.sin(x) / x
    if l == 0:
        return jm
    j = math.sin(x) / (x * x) - math.cos(x) / x
    for nn in range(1, l):
        jm, j = j, (2 * nn + 1) * j / x - jm
    return j


def bessel_root(l, which):
    """src/half_uspp.F90:325-336: scan in steps of 0.05 from 1e-4 for the
    which-th sign change, then bisect to 1e-13."""
    x0 = 1e-4
    f0 = sph_bessel(l, x0)
    found = 0
    while found < which:
        x1 = x0 + 0.05
        f1 = sph_bessel(l, x1)
        if f0 * f1 < 0:
            found += 1
            if found == which:
                break
        x0, f0 = x1, f1
    while x1 - x0 > 1e-13:
        root = 0.5 * (x0 + x1)
        if f0 * sph_bessel(l, root) <= 0:
            x1 = root
        else:
            x0 = root
            f0 = sph_bessel(l, x0)
    return 0.5 * (x0 + x1)


def gauss_legendre(n):
    """src/half_uspp.F90:364-376 (Newton iteration on P_n)."""
    x, w = np.zeros(n), np.zeros(n)
    for i in range(1, (n + 1) // 2 + 1):
        z = math.cos(math.pi * (i - 0.25) / (n + 0.5))
        while True:
            p1, p2 = 1.0, 0.0
            for j in range(1, n + 1):
                p3, p2 = p2, p1
                p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j
            pp = n * (z * p1 - p2) / (z * z - 1)
            z1 = z
            z = z1 - p1 / pp
            if abs(z - z1) < 1e-15:
                break
        x[i - 1], x[n - i] = -z, z
        w[i - 1] = w[n - i] = 2 / ((1 - z * z) * pp * pp)
    return x, w


def compensation_coefficients(l, rc, z1, z2):
    """src/half_uspp.F90:338-345: g_l = c1 j_l(z1 r/rc) + c2 j_l(z2 r/rc)
    with zero slope at rc and unit l-th moment (32-point Gauss-Legendre)."""
    def djl(x):
        return -sph_bessel(1, x) if l == 0 else sph_bessel(l - 1, x) - (l + 1) / x * sph_bessel(l, x)
    a11, a12 = z1 / rc * djl(z1), z2 / rc * djl(z2)
    xg, wg = gauss_legendre(32)
    rg, wg = 0.5 * rc * (xg + 1), 0.5 * rc * wg
    a21 = sum(wg[i] * sph_bessel(l, z1 * rg[i] / rc) * rg[i] ** (l + 2) for i in range(32))
    a22 = sum(wg[i] * sph_bessel(l, z2 * rg[i] / rc) * rg[i] ** (l + 2) for i in range(32))
    det = a11 * a22 - a12 * a21
    return -a12 / det, a11 / det


def _ylm_x(l, m, x, phi):
    """Real Ylm of src/half_uspp.F90:378-385 with x = cos(theta)."""
    am = abs(m)
    p = assoc_legendre(l, am, x)
    norm = math.sqrt((2 * l + 1) / (4 * PI) * _factorial(l - am) / _factorial(l + am))
    sgn = -1.0 if am % 2 else 1.0
    if m > 0:
        return math.sqrt(2.0) * sgn * norm * p * np.cos(am * phi)
    if m < 0:
        return math.sqrt(2.0) * sgn * norm * p * np.sin(am * phi)
    return norm * p


def gaunt_numeric(l1, m1, l2, m2, l3, m3):
    """src/half_uspp.F90:354-362: 12 x 24 Gauss-Legendre x uniform-phi
    quadrature of a product of three real Ylm."""
    xg, wg = gauss_legendre(12)
    g = 0.0
    for it in range(12):
        for ip in range(24):
            phi = 2 * PI * ip / 24
            g += wg[it] * 2 * PI / 24 * float(_ylm_x(l1, m1, xg[it], phi) * _ylm_x(l2, m2, xg[it], phi)
                                               * _ylm_x(l3, m3, xg[it], phi))
    return g


def _bspline_weights(t):
    """Cubic B-spline weights and derivatives (src/half_uspp.F90:278-294)."""
    w = np.stack([(1 - t) ** 3 / 6, (3 * t ** 3 - 6 * t * t + 4) / 6, (-3 * t ** 3 + 3 * t * t + 3 * t + 1) / 6,
                  t ** 3 / 6])
    dw = np.stack([-0.5 * (1 - t) ** 2, 1.5 * t * t - 2 * t, -1.5 * t * t + t + 0.5, 0.5 * t * t])
    return w, dw


def _spline_sample(coeff, grid, frac, invlat):
    """Periodic cubic B-spline value and Cartesian gradient at fractional
    points (src/half_uspp.F90:244-276), vectorised over points."""
    shape = np.array(grid)
    coord = frac * shape
    ix = np.floor(coord).astype(np.int64)
    t = coord - ix
    w, dw = zip(*(_bspline_weights(t[:, d]) for d in range(3)))
    value = np.zeros(frac.shape[0])
    df = np.zeros((frac.shape[0], 3))
    for a in range(4):
        ia = np.mod(ix[:, 0] + a - 1, grid[0])
        for b in range(4):
            ib = np.mod(ix[:, 1] + b - 1, grid[1])
            for c in range(4):
                ic = np.mod(ix[:, 2] + c - 1, grid[2])
                v = coeff[ia, ib, ic]
                value += w[0][a] * w[1][b] * w[2][c] * v
                df[:, 0] += dw[0][a] * w[1][b] * w[2][c] * v * grid[0]
                df[:, 1] += w[0][a] * dw[1][b] * w[2][c] * v * grid[1]
                df[:, 2] += w[0][a] * w[1][b] * dw[2][c] * v * grid[2]
    return value, df @ invlat.T


def build_uspp_dij(veff, grid, potcars, crystal, paw):
    """src/half_uspp.F90:112-242 (build_uspp_dij_cpu): for every atom
    D_ab = DION_ab + sum_k M_ab,k int r^2 g_l(r) [sum_Omega Y_k V(R + r Omega)] dr
    and dD_ab/dR from the spline gradient, both symmetrised; V is the
    periodic cubic B-spline of Veff (FFT prefilter).  Updates paw in place
    ("dij_atom", "ddij_atom")."""
    n1, n2, n3 = grid
    b = [(2.0 + np.cos(2 * PI * np.arange(nn) / nn)) / 3.0 for nn in grid]
    bsym = b[0][:, None, None] * b[1][None, :, None] * b[2][None, None, :]
    coeff = np.fft.ifftn(np.fft.fftn(np.asarray(veff, float).reshape(grid)) / bsym).real
    invlat = np.linalg.inv(crystal.lattice)
    xg, wg = gauss_legendre(12)
    ion0 = 0
    for it, (pc, sp) in enumerate(zip(potcars, paw)):
        lm = [(ich, l, m) for ich, l in enumerate(pc["lps"]) for m in range(-l, l + 1)]
        nlm = len(lm)
        lmax = 2 * max(pc["lps"])
        kl = [(l, m) for l in range(lmax + 1) for m in range(-l, l + 1)]
        r = pc["rgrid"]
        rc = pc["paw_rmax"]
        rw = radial_weights(r)
        gshape = np.zeros((r.size, lmax + 1))
        for l in range(lmax + 1):
            z1, z2 = bessel_root(l, 1), bessel_root(l, 2)
            c1, c2 = compensation_coefficients(l, rc, z1, z2)
            for ir in range(r.size):
                if r[ir] <= rc:
                    gshape[ir, l] = c1 * sph_bessel(l, z1 * r[ir] / rc) + c2 * sph_bessel(l, z2 * r[ir] / rc)
        mult = np.zeros((nlm, nlm, len(kl)))
        for a, (cha, la, ma) in enumerate(lm):
            for bb, (chb, lb, mb) in enumerate(lm):
                for k, (l, m) in enumerate(kl):
                    if abs(la - lb) <= l <= la + lb and (la + lb + l) % 2 == 0:
                        mom = np.sum(rw * (pc["wae"][:, cha] * pc["wae"][:, chb] - pc["wps"][:, cha] * pc["wps"][:, chb])
                                     * r ** l)
                        mult[a, bb, k] = gaunt_numeric(la, ma, lb, mb, l, m) * mom
        dirs, yw = [], []
        for itt in range(12):
            for m in range(24):
                phi = 2 * PI * m / 24
                st = math.sqrt(max(0.0, 1 - xg[itt] * xg[itt]))
                dirs.append([st * math.cos(phi), st * math.sin(phi), xg[itt]])
                yw.append([float(_ylm_x(l, mm, xg[itt], phi)) * wg[itt] * 2 * PI / 24 for l, mm in kl])
        dirs, yw = np.array(dirs), np.array(yw)  # (nang, 3), (nang, naug)
        lk = np.array([l for l, _ in kl])
        radial = rw * r * r  # (nrad,)
        for iat in range(sp["natoms"]):
            pos = crystal.positions[ion0 + iat] @ crystal.lattice
            pts = pos[None, None, :] + r[:, None, None] * dirs[None, :, :]  # (nrad, nang, 3)
            frac = np.mod(pts.reshape(-1, 3) @ invlat, 1.0)
            val, grad = _spline_sample(coeff, grid, frac, invlat)
            val = val.reshape(r.size, -1)
            grad = grad.reshape(r.size, -1, 3)
            vlm = val @ yw  # (nrad, naug)
            rv = np.einsum("r,rk,rk->k", radial, gshape[:, lk], vlm)
            d = sp["dij"] + mult @ rv
            sp["dij_atom"][iat] = 0.5 * (d + d.T)
            for al in range(3):
                glm = grad[:, :, al] @ yw
                rg = np.einsum("r,rk,rk->k", radial, gshape[:, lk], glm)
                dd = mult @ rg
                sp["ddij_atom"][iat, :, :, al] = 0.5 * (dd + dd.T)
        ion0 += sp["natoms"]
    return paw


# --------------------------------------------------------------------------
# XC (src/half_xc.F90:10-116)
# --------------------------------------------------------------------------

def perdew_zunger_xc(rho, omega):
    """src/half_xc.F90:10-32."""
    rho = np.asarray(rho, float)
    vxc = np.zeros_like(rho)
    dvau = (omega / rho.size) / AUTOA ** 3
    dens = rho / omega
    rhoau = dens * AUTOA ** 3
    ok = rhoau > 1e-30
    r = np.where(ok, rhoau, 1.0)
    rs = (3.0 / (4 * PI * r)) ** (1.0 / 3.0)
    ex = -0.4582 / rs
    vx = -(4.0 / 3.0) * 0.4582 / rs
    sr = np.sqrt(rs)
    denom = 1 + 1.0529 * sr + 0.3334 * rs
    ec_hi = -0.1423 / denom
    vc_hi = ec_hi * (1 + (7.0 / 6.0) * 1.0529 * sr + (4.0 / 3.0) * 0.3334 * rs) / denom
    lnrs = np.log(rs)
    ec_lo = 0.0311 * lnrs - 0.048 + 0.002 * rs * lnrs - 0.0116 * rs
    vc_lo = 0.0311 * lnrs + (-0.048 - 0.0311 / 3) + (2.0 / 3) * 0.002 * rs * lnrs + (-0.0232 - 0.002) * rs / 3
    hi = rs >= 1
    ec = np.where(hi, ec_hi, ec_lo)
    vc = np.where(hi, vc_hi, vc_lo)
    vxc = np.where(ok, (vx + vc) * HATOEV, 0.0)
    exc = float(np.sum(np.where(ok, (ex + ec) * rhoau * dvau * HATOEV, 0.0)))
    return vxc, exc


def pw92(rs):
    """src/half_xc.F90:108-116."""
    a, a1, b1, b2, b3, b4 = .031091, .21370, 7.5957, 3.5876, 1.6382, .49294
    sr = np.sqrt(rs)
    q = 2 * a * (b1 * sr + b2 * rs + b3 * rs * sr + b4 * rs * rs)
    la = np.log(1 + 1 / q)
    return -2 * a * (1 + a1 * rs) * la


def pbe_energy_density(density, sigma):
    """src/half_xc.F90:94-107 (vectorised)."""
    mu, kappa, beta = 0.2195149727645171, 0.804, 0.06672455060314922
    gamma = (1.0 - math.log(2.0)) / (PI * PI)
    n = np.maximum(density, 1e-30)
    s = np.maximum(sigma, 0.0)
    rs = (3 / (4 * PI * n)) ** (1.0 / 3.0)
    kf = (3 * PI * PI * n) ** (1.0 / 3.0)
    exlda = -3 * kf / (4 * PI)
    s2 = s / (4 * kf * kf * n * n)
    fx = 1 + kappa - kappa / (1 + mu * s2 / kappa)
    ec = pw92(rs)
    ks2 = 4 * kf / PI
    t2 = s / (4 * ks2 * n * n)
    aa = beta / gamma / (np.exp(-ec / gamma) - 1)
    at2 = aa * t2
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        r1 = t2 * (1 + at2) / (1 + at2 + at2 * at2)
        r2 = (1 / aa) * (1 + 1 / at2) / (1 + 1 / at2 + (1 / at2) ** 2)
    ratio = np.where(t2 == 0, 0.0, np.where(at2 < 1e100, r1, r2))
    hh = gamma * np.log(1 + beta / gamma * ratio)
    return n * (exlda * fx + ec + hh)


def _gvec_grid(grid, crystal):
    f = [np.where(np.arange(n) <= (n - 1) // 2, np.arange(n), np.arange(n) - n) for n in grid]
    a, b, c = np.meshgrid(*f, indexing="ij")
    return (np.stack([a, b, c], -1).reshape(-1, 3).astype(float)) @ crystal.reciprocal


def fft3_forward(x, grid):
    """src/half_fft.F90:25-30: sign -1, unnormalised, C order."""
    return np.fft.fftn(np.asarray(x).reshape(grid)).ravel()


def fft3_backward(x, grid):
    """src/half_fft.F90:31-36: sign +1, unnormalised, C order."""
    n = grid[0] * grid[1] * grid[2]
    return (np.fft.ifftn(np.asarray(x).reshape(grid)) * n).ravel()


def pbe_xc(rho, grid, crystal):
    """src/half_xc.F90:34-92."""
    n = rho.size
    vol = crystal.volume
    density = np.maximum(rho / vol * AUTOA ** 3, 1e-20)
    g = _gvec_grid(grid, crystal)
    freq = fft3_forward(rho / vol + 0j, grid)
    grad = np.zeros((3, n))
    for al in range(3):
        work = fft3_backward(1j * g[:, al] * freq, grid)
        grad[al] = work.real / n * AUTOA ** 4
    sigma = (grad ** 2).sum(0)
    low = rho / vol * AUTOA ** 3 <= 1e-20
    density = np.where(low, 1e-20, density)
    sigma = np.where(low, 0.0, sigma)
    dn = np.maximum(np.abs(density) * 2e-5, 1e-12)
    nm = np.maximum(density - dn, 1e-30)
    npl = density + dn
    fn = (pbe_energy_density(npl, sigma) - pbe_energy_density(nm, sigma)) / (npl - nm)
    ds = np.maximum(np.abs(sigma) * 2e-5, 1e-14)
    sm = np.maximum(sigma - ds, 0.0)
    sp = sigma + ds
    fs = (pbe_energy_density(density, sp) - pbe_energy_density(density, sm)) / (sp - sm)
    edens = pbe_energy_density(density, sigma)
    fs = np.where(low, 0.0, fs)
    vxc = fn.copy()
    for al in range(3):
        flux = 2 * fs * grad[al]
        freq2 = fft3_forward(flux + 0j, grid)
        work = fft3_backward(1j * g[:, al] * freq2, grid)
        vxc = vxc - AUTOA * work.real / n
    vxc = np.where(low, 0.0, vxc)
    vxc = vxc * HATOEV
    dvau = (vol / n) / AUTOA ** 3
    exc = float(np.sum(np.where(low, 0.0, edens * dvau * HATOEV)))
    return vxc, exc


# --------------------------------------------------------------------------
# Veff (src/half_potential.F90:175-254)
# --------------------------------------------------------------------------

def reorder_chgcar_f_to_c(values, grid):
    """src/half_potential.F90:240-254: x-fastest file order to C order."""
    return np.asarray(values).reshape(grid[2], grid[1], grid[0]).transpose(2, 1, 0).ravel()


def build_veff(values, grid, potcars, crystal, use_pbe, with_vxc=False):
    """src/half_potential.F90:175-238.  Returns (veff C-order, e_H, e_xc,
    e_xc_potential), and the XC potential (C order, eV) as a fifth value when
    with_vxc (the reference's vxc_out)."""
    n = grid[0] * grid[1] * grid[2]
    vol = crystal.volume
    charge_c = reorder_chgcar_f_to_c(values, grid)
    rho_g = fft3_forward(charge_c + 0j, grid) / n
    g = _gvec_grid(grid, crystal)
    g2 = np.einsum("ij,ij->i", g, g)
    gn = np.sqrt(g2)
    nz = g2 > 1e-12
    local_g = np.zeros(n, complex)
    local_g[nz] = EDEPS * rho_g[nz] / (g2[nz] * vol)
    e_h = float(np.sum(0.5 * (np.conj(rho_g[nz]) * local_g[nz]).real))
    core_g = np.zeros(n, complex)
    ion0 = 0
    for it, p in enumerate(potcars):
        m2l = spline_second(p["psp_local"], p["psp_gmax"])
        radial = spline_eval(p["psp_local"], m2l, p["psp_gmax"], gn)
        radial = np.where(nz, radial - 4 * PI * p["zval"] * FELECT / np.where(nz, g2, 1.0), radial)
        radial = np.where((~nz) | (gn > p["psp_gmax"] - p["psp_gmax"] / 1000), 0.0, radial)
        core = np.zeros(n)
        if p["has_core"]:
            m2c = spline_second(p["pspcor"], p["psp_gmax"])
            cval = spline_eval(p["pspcor"], m2c, p["psp_gmax"], gn)
            core = np.where(gn <= p["psp_gmax"] - 3 * p["psp_gmax"] / 1000, cval, 0.0)
        for iat in range(crystal.counts[it]):
            rpos = crystal.positions[ion0 + iat] @ crystal.lattice
            ph = g @ rpos
            e = np.cos(ph) - 1j * np.sin(ph)
            local_g += radial * e / vol
            core_g += core * e
        ion0 += crystal.counts[it]
    local_r = fft3_backward(local_g, grid).real
    core_r = fft3_backward(core_g, grid).real
    if use_pbe:
        vxc, exc = pbe_xc(charge_c + core_r, grid, crystal)
    else:
        vxc, exc = perdew_zunger_xc(charge_c + core_r, vol)
    exv = float(np.sum(charge_c * vxc) / n)
    if with_vxc:
        return local_r + vxc, e_h, exc, exv, vxc
    return local_r + vxc, e_h, exc, exv


# --------------------------------------------------------------------------
# dense H/S and matrix-free H.psi (src/half_dense_solver.F90:53-94,
# src/half_cuda_assembly.cuf:468-519)
# --------------------------------------------------------------------------

def veff_g(veff, grid):
    """Veff(G)/N_r in C order (src/half_dense_solver.F90:61)."""
    return fft3_forward(veff + 0j, grid) / veff.size


def assemble_dense(veff, basis, paw):
    """src/half_dense_solver.F90:53-88 (H_ij = Veff(Gi-Gj)/N_r + delta T +
    sum_I beta_a(Gi) D_ab conj beta_b(Gj), S likewise, then hermitised)."""
    grid = basis.grid
    vg = veff_g(veff, grid)
    g = basis.gvec.astype(np.int64)
    a = np.mod(g[:, None, 0] - g[None, :, 0], grid[0])
    b = np.mod(g[:, None, 1] - g[None, :, 1], grid[1])
    c = np.mod(g[:, None, 2] - g[None, :, 2], grid[2])
    h = vg[a * grid[1] * grid[2] + b * grid[2] + c]
    del a, b, c
    n = basis.kinetic.size
    s = np.eye(n, dtype=complex)
    h[np.diag_indices(n)] += basis.kinetic
    for sp in paw:
        for iat in range(sp["natoms"]):
            bt = sp["beta"][iat]  # nlm x N
            h += bt.T @ sp["dij_atom"][iat] @ bt.conj()
            s += bt.T @ sp["qij"] @ bt.conj()
    h = 0.5 * (h + h.conj().T)
    s = 0.5 * (s + s.conj().T)
    return h, s


def apply_hs(veff, basis, paw, psi, workers=None):
    """Matrix-free H.psi, S.psi (src/half_cuda_assembly.cuf:468-519):
    scatter -> inverse FFT -> * Veff/N_r -> forward FFT -> gather + T psi,
    then + B D B^H psi and S psi = psi + B Q B^H psi.  workers: FFT threads
    (scipy.fft over the whole band block); None transforms band by band with
    numpy."""
    grid = basis.grid
    nr = veff.size
    psi = np.asarray(psi, complex)
    if psi.ndim == 1:
        psi = psi[:, None]
    hpsi = np.empty_like(psi)
    if workers:
        import scipy.fft as sf
        nb = psi.shape[1]
        box = np.zeros((nb, nr), complex)
        box[:, basis.fft_index] = psi.T
        # ifftn carries the 1/N_r of Veff/N_r applied after an unnormalised backward transform
        r = sf.ifftn(box.reshape(nb, *grid), axes=(1, 2, 3), workers=workers) * veff.reshape(grid)
        f = sf.fftn(r, axes=(1, 2, 3), workers=workers).reshape(nb, nr)
        hpsi[:] = f[:, basis.fft_index].T + basis.kinetic[:, None] * psi
    else:
        for j in range(psi.shape[1]):
            box = np.zeros(nr, complex)
            box[basis.fft_index] = psi[:, j]
            r = fft3_backward(box, grid) * (veff / nr)
            hpsi[:, j] = fft3_forward(r, grid)[basis.fft_index] + basis.kinetic * psi[:, j]
    spsi = psi.copy()
    for sp in paw:
        nat, nlm = sp["natoms"], sp["nlm"]
        B = sp["beta"].reshape(nat * nlm, -1).T  # N x Np, column (iat, lm)
        c = B.conj().T @ psi
        c3 = c.reshape(nat, nlm, -1)
        dc = np.einsum("aij,ajb->aib", sp["dij_atom"], c3).reshape(nat * nlm, -1)
        qc = np.einsum("ij,ajb->aib", sp["qij"], c3).reshape(nat * nlm, -1)
        hpsi += B @ dc
        spsi += B @ qc
    return hpsi, spsi


def solve_dense(h, s, nbands=None, vectors=True):
    """src/half_dense_solver.F90:29-51: overlap spectrum (zheev 'N'), then the
    generalized problem zhegv(itype=1, 'U')."""
    import scipy.linalg as sl
    sv = sl.eigvalsh(s)
    w, v = sl.eigh(h, s, lower=False, type=1, driver="gv")
    if nbands is not None:
        w, v = w[:nbands], v[:, :nbands]
    return w, (v if vectors else None), float(sv[0]), float(sv[-1])


# --------------------------------------------------------------------------
# forces, occupations, Ewald, Harris energy
# --------------------------------------------------------------------------

def nonlocal_forces(basis_q, paw, eigenvalues, occupations, kweight, vectors, nions=None):
    """src/half_forces.F90:9-73:
    F_I,alpha = -2 w sum_n f_n Re[(M c_n)^H dc_n,alpha] - w sum_n f_n Re[c_n^H dD_I/dR_alpha c_n],
    M = D_I - eps_n Q, c_a = sum_G conj(beta_a) psi_G,
    dc_a,alpha = sum_G i q_alpha conj(beta_a) psi_G (dD/dR is zero without QDEP)."""
    nions = nions or sum(sp["natoms"] for sp in paw)
    forces = np.zeros((nions, 3))
    ion0 = 0
    for sp in paw:
        for iat in range(sp["natoms"]):
            bc = sp["beta"][iat].conj()  # nlm x N
            c = bc @ vectors  # nlm x nb
            for ib, (e, f) in enumerate(zip(eigenvalues, occupations)):
                if f == 0.0:
                    continue
                mc = (sp["dij_atom"][iat] - e * sp["qij"]) @ c[:, ib]
                for al in range(3):
                    dc = bc @ (1j * basis_q[:, al] * vectors[:, ib])
                    forces[ion0 + iat, al] += -2.0 * kweight * f * float(np.vdot(mc, dc).real)
                    dd = sp.get("ddij_atom")
                    if dd is not None:
                        forces[ion0 + iat, al] += -kweight * f * float(np.vdot(c[:, ib], dd[iat, :, :, al] @ c[:, ib]).real)
        ion0 += sp["natoms"]
    return forces


# --------------------------------------------------------------------------
# analytic force set (app/half_cli.F90:921-962, src/half_local_forces.F90,
# src/half_uspp.F90:16-111)
# --------------------------------------------------------------------------

def reorder_c_to_f(values, grid):
    """C-order box (i1 slowest) to the CHGCAR order (x fastest)."""
    return np.asarray(values).reshape(grid).transpose(2, 1, 0).ravel()


def smooth_density(basis, vectors, occupations, kweight, grid):
    """src/half_local_forces.F90 accumulate_smooth_density: sum_n w f_n
    |psi_n(r)|^2 with psi(r) the unnormalised backward transform of the
    plane-wave coefficients (so the grid mean is the electron count, the
    CHGCAR convention); returned in CHGCAR order."""
    n = grid[0] * grid[1] * grid[2]
    dens = np.zeros(n)
    for ib, f in enumerate(occupations):
        if f == 0.0:
            continue
        box = np.zeros(n, complex)
        box[basis.fft_index] = vectors[:, ib]
        dens += kweight * f * np.abs(fft3_backward(box, grid)) ** 2
    return reorder_c_to_f(dens, grid)


def augmentation_occupancy(paw, vectors, occupations, kweight):
    """src/half_uspp.F90:26-45: rho_ab^I = sum_n w f_n Re[conj(c_a) c_b],
    c = B^H psi; one nlm x nlm matrix per atom, species order."""
    out = []
    for sp in paw:
        for iat in range(sp["natoms"]):
            c = sp["beta"][iat].conj() @ vectors  # nlm x nb
            m = np.zeros((sp["nlm"], sp["nlm"]))
            for ib, f in enumerate(occupations):
                if f != 0.0:
                    m += kweight * f * np.real(np.outer(c[:, ib].conj(), c[:, ib]))
            out.append(m)
    return out


def augmentation_density(potcars, crystal, grid, occupancy):
    """src/half_uspp.F90:47-111 (add_augmentation_density): inside each
    atom's PAW sphere, Omega sum_k q_k g_l(r) Y_k(r_hat) with
    q_k = sum_ab rho_ab M_ab,k (M: Gaunt x (AE - PS) partial-wave moment,
    g_l: two-Bessel compensation shape).  Returns the CHGCAR-order field
    and its grid mean (the augmentation charge)."""
    n1, n2, n3 = grid
    out = np.zeros(n1 * n2 * n3)  # CHGCAR order: index i1 + n1 (i2 + n2 i3)
    i3, i2, i1 = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    frac = np.stack([i1.ravel() / n1, i2.ravel() / n2, i3.ravel() / n3], 1)
    ion0 = 0
    k_at = 0
    for it, pc in enumerate(potcars):
        lm = [(ich, l, m) for ich, l in enumerate(pc["lps"]) for m in range(-l, l + 1)]
        nlm = len(lm)
        lmax = 2 * max(pc["lps"])
        kl = [(l, m) for l in range(lmax + 1) for m in range(-l, l + 1)]
        r = pc["rgrid"]
        rc = pc["paw_rmax"]
        rw = radial_weights(r)
        mult = np.zeros((nlm, nlm, len(kl)))
        for a, (cha, la, ma) in enumerate(lm):
            for bb, (chb, lb, mb) in enumerate(lm):
                for k, (l, m) in enumerate(kl):
                    if abs(la - lb) <= l <= la + lb and (la + lb + l) % 2 == 0:
                        mom = np.sum(rw * (pc["wae"][:, cha] * pc["wae"][:, chb] - pc["wps"][:, cha] * pc["wps"][:, chb])
                                     * r ** l)
                        mult[a, bb, k] = gaunt_numeric(la, ma, lb, mb, l, m) * mom
        roots = [(bessel_root(l, 1), bessel_root(l, 2)) for l in range(lmax + 1)]
        cc = [compensation_coefficients(l, rc, *roots[l]) for l in range(lmax + 1)]
        for iat in range(crystal.counts[it]):
            q = np.einsum("ab,abk->k", occupancy[k_at], mult)
            k_at += 1
            df = frac - crystal.positions[ion0 + iat]
            df -= np.rint(df)
            cart = df @ crystal.lattice
            rr = np.sqrt((cart * cart).sum(1))
            inside = rr <= rc
            if not inside.any():
                continue
            ci, ri = cart[inside], rr[inside]
            big = ri > 1e-14
            theta = np.where(big, np.arccos(np.clip(ci[:, 2] / np.where(big, ri, 1.0), -1, 1)), 0.0)
            phi = np.where(big, np.arctan2(ci[:, 1], ci[:, 0]), 0.0)
            val = np.zeros(ri.size)
            for k, (l, m) in enumerate(kl):
                if q[k] == 0.0:
                    continue
                z1, z2 = roots[l]
                c1, c2 = cc[l]
                g = c1 * np.array([sph_bessel(l, z1 * x / rc) for x in ri]) + \
                    c2 * np.array([sph_bessel(l, z2 * x / rc) for x in ri])
                val += q[k] * g * real_ylm(l, m, theta, phi)
            out[np.nonzero(inside)[0]] += crystal.volume * val
        ion0 += crystal.counts[it]
    return out, float(out.sum() / out.size)


def _radial_table_forces(field_g, grid, potcars, crystal, table_key, cut, scale_by_volume, subtract_coulomb):
    """F_I = -sum_G q t(|G|) Im[conj(field_G) e^{-i G.R_I}] over the box
    (src/half_local_forces.F90:11-150); t is the spline of the species'
    table (minus 4 pi Z felect / G^2 for the local part), zero at G = 0 and
    beyond gmax (1 - cut / n_table); divided by the cell volume for the local
    part."""
    g = _gvec_grid(grid, crystal)
    g2 = np.einsum("ij,ij->i", g, g)
    gn = np.sqrt(g2)
    f = np.zeros((crystal.nions, 3))
    ion0 = 0
    for it, p in enumerate(potcars):
        tab = p.get(table_key)
        if tab is None or (table_key == "pspcor" and not p["has_core"]):
            ion0 += crystal.counts[it]
            continue
        m2 = spline_second(tab, p["psp_gmax"])
        ok = (g2 > 1e-12) & ~(gn > p["psp_gmax"] - cut * p["psp_gmax"] / len(tab))
        rad = np.where(ok, spline_eval(tab, m2, p["psp_gmax"], gn), 0.0)
        if subtract_coulomb:
            rad = np.where(ok, rad - 4 * PI * p["zval"] * FELECT / np.where(ok, g2, 1.0), 0.0)
        if scale_by_volume:
            rad = rad / crystal.volume
        for iat in range(crystal.counts[it]):
            rpos = crystal.positions[ion0 + iat] @ crystal.lattice
            ph = g @ rpos
            term = rad * np.imag(np.conj(field_g) * (np.cos(ph) - 1j * np.sin(ph)))
            f[ion0 + iat] = -(g * term[:, None]).sum(0)
        ion0 += crystal.counts[it]
    return f


def local_ionic_forces(rho_f, grid, potcars, crystal):
    """src/half_local_forces.F90:11-52 on a CHGCAR-order density."""
    n = grid[0] * grid[1] * grid[2]
    rho_g = fft3_forward(reorder_chgcar_f_to_c(rho_f, grid) + 0j, grid) / n
    return _radial_table_forces(rho_g, grid, potcars, crystal, "psp_local", 1, True, True)


def nlcc_forces(vxc_c, grid, potcars, crystal):
    """src/half_local_forces.F90:54-98 on a C-order XC potential."""
    n = grid[0] * grid[1] * grid[2]
    v_g = fft3_forward(np.asarray(vxc_c, float) + 0j, grid) / n
    return _radial_table_forces(v_g, grid, potcars, crystal, "pspcor", 3, False, False)


def harris_correction_forces(rho_in_f, rho_out_f, grid, potcars, crystal, use_pbe, t=1e-4):
    """src/half_local_forces.F90:124-158: the XC kernel on (n_out - n_in) by
    central differences of v_xc, contracted with the core density
    derivative."""
    delta = np.asarray(rho_out_f) - np.asarray(rho_in_f)
    vp = build_veff(np.asarray(rho_in_f) + t * delta, grid, potcars, crystal, use_pbe, with_vxc=True)[4]
    vm = build_veff(np.asarray(rho_in_f) - t * delta, grid, potcars, crystal, use_pbe, with_vxc=True)[4]
    return nlcc_forces((vp - vm) / (2 * t), grid, potcars, crystal)


def compute_occupations(eig, weights, nelect, sigma, ismear=0):
    """src/half_energy.F90:41-120.  eig is (nk, nb).  Returns (occ, mu,
    band_energy, entropy_term)."""
    from math import erfc
    eig = np.asarray(eig, float)
    weights = np.asarray(weights, float)
    nk, nb = eig.shape
    occ = np.zeros((nk, nb))
    if sigma == 0.0:
        states = [(eig[ik, ib], ik, ib) for ik in range(nk) for ib in range(nb)]
        order = sorted(range(len(states)), key=lambda s: states[s][0])  # stable like insertion sort
        remaining = nelect
        mu = float(eig.min())
        pos = 0
        ns = len(states)
        while pos < ns and remaining > 1e-13:
            level = states[order[pos]][0]
            stop = pos + 1
            tol = 1e-10 * max(1.0, abs(level))
            while stop < ns and abs(states[order[stop]][0] - level) <= tol:
                stop += 1
            cap = sum(2.0 * weights[states[order[i]][1]] for i in range(pos, stop))
            frac = min(1.0, max(0.0, remaining / cap))
            for i in range(pos, stop):
                _, ik, ib = states[order[i]]
                occ[ik, ib] = 2.0 * frac
            filled = frac * cap
            if filled > 0:
                mu = level
            remaining -= filled
            pos = stop
        if remaining > 1e-9:
            raise ValueError("not enough bands for electron count")
        return occ, mu, float(np.sum(weights[:, None] * occ * eig)), 0.0

    def count(mu):
        v = 0.0
        for jk in range(nk):
            for jb in range(nb):
                xx = (eig[jk, jb] - mu) / sigma
                if ismear == 0:
                    v += weights[jk] * erfc(xx)
                else:
                    xx = max(-700.0, min(700.0, xx))
                    v += weights[jk] * 2.0 / (1.0 + math.exp(xx))
        return v
    lo = eig.min() - max(10.0, 100.0 * sigma)
    hi = eig.max() + max(10.0, 100.0 * sigma)
    for _ in range(256):
        mid = 0.5 * (lo + hi)
        if count(mid) < nelect:
            lo = mid
        else:
            hi = mid
        if hi - lo < 1e-13 * max(1.0, abs(mid)):
            break
    mu = 0.5 * (lo + hi)
    ent = 0.0
    for ik in range(nk):
        for ib in range(nb):
            x = (eig[ik, ib] - mu) / sigma
            if ismear == 0:
                occ[ik, ib] = erfc(x)
                if abs(x) < 20.0:
                    ent -= sigma * weights[ik] * math.exp(-x * x) / math.sqrt(PI)
            else:
                x = max(-700.0, min(700.0, x))
                p = 1.0 / (1.0 + math.exp(x))
                occ[ik, ib] = 2.0 * p
                p = max(1e-300, min(1.0 - 1e-16, p))
                ent += 2.0 * sigma * weights[ik] * (p * math.log(p) + (1 - p) * math.log(1 - p))
    return occ, mu, float(np.sum(weights[:, None] * occ * eig)), ent


def _smallest_singular(mat):
    return float(np.sqrt(max(0.0, np.linalg.eigvalsh(mat @ mat.T).min())))


def _ewald_setup(crystal, eta, tolerance):
    alpha = math.sqrt(PI) / crystal.volume ** (1.0 / 3.0) if eta is None else eta
    scale = math.sqrt(-math.log(tolerance))
    rcut, gcut = scale / alpha, 2 * alpha * scale
    rb = math.ceil(rcut / _smallest_singular(crystal.lattice)) + 1
    gb = math.ceil(gcut / _smallest_singular(crystal.reciprocal)) + 1
    return alpha, rcut, gcut, rb, gb


def ewald_energy(crystal, charges, eta=None, tolerance=1e-11):
    """src/half_energy.F90:122-162."""
    from scipy.special import erfc
    charges = np.asarray(charges, float)
    alpha, rcut, gcut, rb, gb = _ewald_setup(crystal, eta, tolerance)
    pos = crystal.positions @ crystal.lattice
    r = np.arange(-rb, rb + 1)
    nv = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3) @ crystal.lattice
    re_ = 0.0
    for i in range(len(charges)):
        for j in range(len(charges)):
            d = pos[i] - pos[j] + nv
            dist = np.sqrt((d * d).sum(1))
            ok = (dist > 1e-13) & (dist <= rcut)
            re_ += charges[i] * charges[j] * np.sum(erfc(alpha * dist[ok]) / dist[ok])
    re_ *= 0.5 * FELECT
    gg = np.arange(-gb, gb + 1)
    gv = np.stack(np.meshgrid(gg, gg, gg, indexing="ij"), -1).reshape(-1, 3) @ crystal.reciprocal
    gn = np.sqrt((gv * gv).sum(1))
    ok = (gn > 1e-13) & (gn <= gcut)
    gv, gn = gv[ok], gn[ok]
    ph = gv @ pos.T
    sr = (charges * np.cos(ph)).sum(1)
    si = -(charges * np.sin(ph)).sum(1)
    ge = 2 * PI * FELECT / crystal.volume * np.sum(np.exp(-gn * gn / (4 * alpha * alpha)) * (sr * sr + si * si) / (gn * gn))
    se = -FELECT * alpha / math.sqrt(PI) * np.sum(charges * charges)
    be = -FELECT * PI * charges.sum() ** 2 / (2 * alpha * alpha * crystal.volume)
    return float(re_ + ge + se + be)


def ewald_forces(crystal, charges, eta=None, tolerance=1e-11):
    """src/half_energy.F90:164-210."""
    from scipy.special import erfc
    charges = np.asarray(charges, float)
    alpha, rcut, gcut, rb, gb = _ewald_setup(crystal, eta, tolerance)
    pos = crystal.positions @ crystal.lattice
    r = np.arange(-rb, rb + 1)
    nv = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3) @ crystal.lattice
    f = np.zeros((len(charges), 3))
    for i in range(len(charges)):
        for j in range(len(charges)):
            d = pos[i] - pos[j] + nv
            dist = np.sqrt((d * d).sum(1))
            ok = (dist > 1e-13) & (dist <= rcut)
            d, dist = d[ok], dist[ok]
            scr = erfc(alpha * dist) / dist ** 3 + 2 * alpha * np.exp(-(alpha * dist) ** 2) / (math.sqrt(PI) * dist ** 2)
            f[i] += FELECT * charges[i] * charges[j] * (scr[:, None] * d).sum(0)
    gg = np.arange(-gb, gb + 1)
    gv = np.stack(np.meshgrid(gg, gg, gg, indexing="ij"), -1).reshape(-1, 3) @ crystal.reciprocal
    gn = np.sqrt((gv * gv).sum(1))
    ok = (gn > 1e-13) & (gn <= gcut)
    gv, gn = gv[ok], gn[ok]
    ph = gv @ pos.T
    sr = (charges * np.cos(ph)).sum(1)
    si = -(charges * np.sin(ph)).sum(1)
    pre = 4 * PI * FELECT / crystal.volume * np.exp(-gn * gn / (4 * alpha * alpha)) / (gn * gn)
    for i in range(len(charges)):
        fac = pre * charges[i] * (sr * np.sin(ph[:, i]) + si * np.cos(ph[:, i]))
        f[i] += (fac[:, None] * gv).sum(0)
    return f


def harris_energy(eigenvalues, nelect, e_h, e_xc, e_xc_pot, potcars, crystal):
    """app/half_cli.F90:896-920 at one k point (weight 1), zero temperature,
    use_uspp off: band - e_H - e_xc,pot + e_xc + Ewald + local_g0 + sum EATOM."""
    occ, mu, band, _ = compute_occupations(np.asarray(eigenvalues)[None, :], [1.0], nelect, 0.0)
    charges = []
    for it, p in enumerate(potcars):
        charges += [p["zval"]] * crystal.counts[it]
    ew = ewald_energy(crystal, charges)
    atomic = sum(p["eatom"] * crystal.counts[it] for it, p in enumerate(potcars))
    lg0 = sum(p["psp_local"][0] * crystal.counts[it] for it, p in enumerate(potcars)) * nelect / crystal.volume
    total = band - e_h - e_xc_pot + e_xc + ew + lg0 + atomic
    return {"band_energy": band, "fermi": mu, "ewald": ew, "local_g0": lg0,
            "atomic_reference": atomic, "harris_energy": total, "occupations": occ[0]}


# --------------------------------------------------------------------------
# ACC all-band solver (src/half_cuda_iterative_solver.cuf:79-310)
# --------------------------------------------------------------------------

def kinetic_order(kinetic):
    """src/half_cuda_iterative_solver.cuf:277-310: heap sort of indices by
    kinetic energy (ties resolved as the reference's heap does)."""
    idx = list(range(1, len(kinetic) + 1))
    kin = kinetic
    n = len(idx)
    for i in range(2, n + 1):
        child = i
        while child > 1:
            parent = child // 2
            if kin[idx[parent - 1] - 1] >= kin[idx[child - 1] - 1]:
                break
            idx[parent - 1], idx[child - 1] = idx[child - 1], idx[parent - 1]
            child = parent
    for last in range(n, 1, -1):
        idx[0], idx[last - 1] = idx[last - 1], idx[0]
        parent = 1
        while True:
            child = 2 * parent
            if child > last - 1:
                break
            if child + 1 <= last - 1 and kin[idx[child] - 1] > kin[idx[child - 1] - 1]:
                child += 1
            if kin[idx[parent - 1] - 1] >= kin[idx[child - 1] - 1]:
                break
            idx[parent - 1], idx[child - 1] = idx[child - 1], idx[parent - 1]
            parent = child
    return np.array(idx) - 1


def acc_solve(apply, kinetic, m, tol=1e-4, maxiter=40):
    """src/half_cuda_iterative_solver.cuf:79-207 restated with numpy/scipy.
    apply(X) -> (HX, SX).  Returns (eigenvalues, X, iterations, max residual)."""
    import scipy.linalg as sl
    n = kinetic.size
    if m < 1 or 2 * m > n:
        raise ValueError("ACC requires 1 <= m <= N/2")
    order = kinetic_order(kinetic)
    x = np.zeros((n, m), complex)
    x[order[:m], np.arange(m)] = 1.0
    hx, sx = apply(x)

    def rr(z, hz, sz, keep):
        hs = z.conj().T @ hz
        ss = z.conj().T @ sz
        w, v = sl.eigh(hs, ss, lower=False, type=1)
        v = v[:, :keep]
        return w[:keep], z @ v, hz @ v, sz @ v

    w, x, hx, sx = rr(x, hx, sx, m)
    it = 0
    maxres = float("inf")
    for it in range(1, maxiter + 1):
        r = hx - sx * w[None, :]
        gram = (r.conj() * r).sum(0).real
        res = np.sqrt(np.maximum(0.0, gram))
        maxres = float(res.max())
        if maxres <= tol:
            break
        seeds = np.where(res < max(tol * 1e-2, 1e-12), np.mod(m + np.arange(m), n), -1)
        p = -r / (np.abs(kinetic[:, None] - w[None, :]) + 1.0)
        for j in np.nonzero(seeds >= 0)[0]:
            p[:, j] = 0
            p[seeds[j], j] = 1.0
        hp, sp = apply(p)
        cross = x.conj().T @ sp
        p = p - x @ cross
        hp = hp - hx @ cross
        sp = sp - sx @ cross
        met = p.conj().T @ sp
        mu, u = np.linalg.eigh(met)
        u = u / np.sqrt(np.maximum(mu, 1e-14))[None, :]
        p, hp, sp = p @ u, hp @ u, sp @ u
        w, x, hx, sx = rr(np.hstack([x, p]), np.hstack([hx, hp]), np.hstack([sx, sp]), m)
    return w, x, it, maxres


# --------------------------------------------------------------------------
# convenience: a whole context from files
# --------------------------------------------------------------------------

class Context:
    """Mirror of the CPU context (src/half_library.F90:51-110, 267-388)."""

    def __init__(self, chgcar, potcar, encut, xc="pbe"):
        self.crystal, self.grid, self.values = read_chgcar(chgcar)
        self.potcars = read_potcar(potcar)
        self.encut = float(encut)
        self.veff, self.e_h, self.e_xc, self.e_xc_pot = build_veff(
            self.values, self.grid, self.potcars, self.crystal, xc == "pbe")

    @property
    def nelect(self):
        return float(sum(p["zval"] * c for p, c in zip(self.potcars, self.crystal.counts)))

    def prepare(self, kpoint=(0.0, 0.0, 0.0)):
        basis = build_basis(self.crystal, self.grid, self.encut, kpoint)
        paw = build_projectors(self.potcars, self.crystal, basis)
        return basis, paw
