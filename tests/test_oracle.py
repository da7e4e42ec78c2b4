"""The numpy restatement (oracle/half_oracle.py) pinned against the
reference's own CPU path (golden fixtures written by oracle/_ref, the
reference Fortran compiled with flang + MKL), the reference's known-answer
tests (app/half_energy_check.F90) and closed-form properties."""
import itertools
import math

import numpy as np
import pytest
from conftest import golden, psi_block, require_same_inputs, workload_files

import half_oracle as O

SI2 = ("tests/golden/inputs/si2.CHGCAR", "tests/golden/inputs/si2.POTCAR")


@pytest.fixture(scope="module", params=["si2_pbe", "si2_lda", "si2_k"])
def si2_case(request):
    g = golden(request.param)
    ctx = O.Context(*SI2, g["encut"], g["xc"])
    basis, paw = ctx.prepare(g["kpoint"])
    h, s = O.assemble_dense(ctx.veff, basis, paw)
    return request.param, g, ctx, basis, paw, h, s


def test_veff_and_energy_terms(si2_case):
    name, g, ctx, *_ = si2_case
    ref = np.load(f"tests/golden/{name}.veff.npy")
    # PBE uses finite-difference derivatives with a 2e-5 relative step, which
    # amplifies last-bit differences of the energy density; LDA does not.
    tol = 1e-7 if g["xc"] == "pbe" else 1e-11
    assert np.abs(ctx.veff - ref).max() < tol
    assert abs(ctx.e_h - g["e_hartree"]) < 1e-10
    assert abs(ctx.e_xc - g["e_xc"]) < 1e-9
    assert abs(ctx.e_xc_pot - g["e_xc_potential"]) < 1e-9
    assert abs(ctx.crystal.volume - g["volume"]) < 1e-12


def test_hs_elements(si2_case):
    name, g, ctx, basis, paw, h, s = si2_case
    assert basis.kinetic.size == g["npw"]
    hs = np.array(g["hs_samples"])
    i, j = hs[:, 0].astype(int) - 1, hs[:, 1].astype(int) - 1
    assert np.abs(h[i, j] - (hs[:, 2] + 1j * hs[:, 3])).max() < 1e-9
    assert np.abs(s[i, j] - (hs[:, 4] + 1j * hs[:, 5])).max() < 1e-12
    assert np.abs(h.diagonal().real - g["h_diag"]).max() < 1e-9


def test_apply_images(si2_case):
    name, g, ctx, basis, paw, h, s = si2_case
    psi = psi_block(g["npw"])
    hp, sp = O.apply_hs(ctx.veff, basis, paw, psi)
    ref = np.load(f"tests/golden/{name}.apply.npy")
    assert np.abs(hp - ref[0]).max() / np.abs(ref[0]).max() < 1e-11
    assert np.abs(sp - ref[1]).max() < 1e-12
    # matrix-free and dense forms are the same circular convolution
    assert np.abs(h @ psi - hp).max() / np.abs(hp).max() < 1e-12


def test_eigen_energy_forces(si2_case):
    name, g, ctx, basis, paw, h, s = si2_case
    w, v, smin, smax = O.solve_dense(h, s, g["nbands"])
    assert np.abs(w - g["eigenvalues"]).max() < 1e-9
    assert abs(smin - g["overlap_min"]) < 1e-12 and abs(smax - g["overlap_max"]) < 1e-12
    he = O.harris_energy(w, ctx.nelect, ctx.e_h, ctx.e_xc, ctx.e_xc_pot, ctx.potcars, ctx.crystal)
    for key in ("band_energy", "ewald", "local_g0", "atomic_reference", "harris_energy"):
        assert abs(he[key] - g[key]) < 1e-9, key
    f = O.nonlocal_forces(basis.q, paw, w, he["occupations"], 1.0, v)
    assert np.abs(f - np.array(g["force_nl"])).max() < 1e-9


def test_hfo2_dense_golden(inputs_dir):
    g = golden("hfo2_pbe")
    chg, pot = workload_files("hfo2", inputs_dir)
    require_same_inputs(g, chg, pot)
    ctx = O.Context(chg, pot, g["encut"], "pbe")
    idx = np.array(g["veff_sample_index"])
    assert np.abs(ctx.veff[idx] - g["veff_sample"]).max() < 1e-7
    assert abs(ctx.e_h - g["e_hartree"]) < 1e-9 and abs(ctx.e_xc - g["e_xc"]) < 1e-9
    basis, paw = ctx.prepare()
    assert basis.kinetic.size == g["npw"] == 3407
    h, s = O.assemble_dense(ctx.veff, basis, paw)
    hs = np.array(g["hs_samples"])
    i, j = hs[:, 0].astype(int) - 1, hs[:, 1].astype(int) - 1
    assert np.abs(h[i, j] - (hs[:, 2] + 1j * hs[:, 3])).max() < 1e-9
    import scipy.linalg as sl
    w = sl.eigh(h, s, eigvals_only=True, subset_by_index=[0, g["nbands"] - 1])
    assert np.abs(w - g["eigenvalues"]).max() < 1e-9


# --------------------------------------------------------------------------
# known-answer tests of app/half_energy_check.F90
# --------------------------------------------------------------------------

def test_occupation_kats():
    k = golden("kat")
    occ, mu, band, ent = O.compute_occupations(np.array([[0.0, 2.0], [1.0, 3.0]]), [0.5, 0.5], 2.0, 0.0)
    ref = np.array(k["occ_zero"])
    assert np.abs(occ.T.ravel() - ref[:4]).max() < 1e-14  # Fortran column order
    assert abs(mu - ref[4]) < 1e-14 and abs(band - ref[5]) < 1e-14 and abs(ent) < 1e-14
    eig = np.array([[-1.0, 0.5, 2.0], [-0.8, 0.7, 2.2]])
    for key, ismear in (("occ_fermi", -1), ("occ_gauss", 0)):
        occ, mu, band, ent = O.compute_occupations(eig, [0.25, 0.75], 2.0, 0.1, ismear)
        ref = np.array(k[key])
        assert np.abs(occ.T.ravel() - ref[:6]).max() < 1e-12
        assert abs(mu - ref[6]) < 1e-12 and abs(band - ref[7]) < 1e-12 and abs(ent - ref[8]) < 1e-12
        assert abs((np.array([0.25, 0.75])[:, None] * occ).sum() - 2.0) < 1e-12 and ent < 0


def _kat_crystal():
    return O.Crystal(np.diag([4.0, 4.0, 4.0]), ["X"], [2], np.array([[0.07, 0.11, 0.19], [0.31, 0.46, 0.73]]))


def test_ewald_kats():
    k = golden("kat")
    c = _kat_crystal()
    e1 = O.ewald_energy(c, [1.0, 1.0], eta=0.35, tolerance=1e-12)
    e2 = O.ewald_energy(c, [1.0, 1.0], eta=0.80, tolerance=1e-12)
    assert abs(e1 - k["ewald_kat"][0]) < 1e-10 and abs(e2 - k["ewald_kat"][1]) < 1e-10
    assert abs(e1 - e2) < 1e-10  # splitting-parameter invariance
    f = O.ewald_forces(c, [1.0, 1.0], eta=0.35, tolerance=1e-12)
    assert np.abs(f - np.array(k["ewald_force_kat"])).max() < 1e-10
    # central finite difference, h = 1e-5 A (app/half_energy_check.F90:47-57)
    inv = np.linalg.inv(c.lattice)
    h = 1e-5
    for i, j in itertools.product(range(2), range(3)):
        d = np.zeros(3)
        d[j] = h
        cp, cm = _kat_crystal(), _kat_crystal()
        cp.positions[i] += d @ inv
        cm.positions[i] -= d @ inv
        num = -(O.ewald_energy(cp, [1, 1], eta=0.35, tolerance=1e-12) -
                O.ewald_energy(cm, [1, 1], eta=0.35, tolerance=1e-12)) / (2 * h)
        assert abs(num - f[i, j]) < 2e-7


def _kat_nonlocal():
    q = np.array([[0.2, 0.7, -0.4], [-0.3, 0.5, 0.8], [0.6, -0.2, 0.9]]).T  # (ig, axis)
    p0 = np.array([[0.4 + 0.1j, -0.2 + 0.3j], [0.1 - 0.2j, 0.5 + 0.4j], [-0.3 + 0.2j, 0.2 + 0.6j]]).T  # (lm, ig)
    waves = np.array([0.3 - 0.1j, -0.2 + 0.4j, 0.5 + 0.2j])[:, None]
    d = np.array([[1.2, 0.15], [0.15, 0.8]])
    qq = np.array([[0.1, 0.02], [0.02, 0.07]])
    return q, p0, waves, d, qq


def test_nonlocal_force_kat():
    """app/half_energy_check.F90:58-81: analytic vs central FD, 2e-9 eV/A,
    and the reference's analytic value."""
    q, p0, waves, d, qq = _kat_nonlocal()
    r0 = np.array([0.13, -0.21, 0.08])

    def paw_at(r):
        beta = p0 * np.exp(-1j * (q @ r))[None, :]
        return [{"natoms": 1, "nlm": 2, "beta": beta[None], "dij": d, "qij": qq, "dij_atom": d[None]}]

    def energy(r):
        c = paw_at(r)[0]["beta"][0].conj() @ waves[:, 0]
        m = d - 0.37 * qq
        return 0.4 * 1.6 * float(np.vdot(c, m @ c).real)

    f = O.nonlocal_forces(q, paw_at(r0), [0.37], [1.6], 0.4, waves)[0]
    h = 1e-5
    num = np.array([-(energy(r0 + h * e) - energy(r0 - h * e)) / (2 * h) for e in np.eye(3)])
    assert np.abs(f - num).max() < 2e-9
    assert np.abs(f - np.array(golden("kat")["nonlocal_force_kat"])).max() < 1e-13


# --------------------------------------------------------------------------
# closed forms and invariants
# --------------------------------------------------------------------------

def test_basis_bruteforce():
    ctx_c, grid, _ = O.read_chgcar(SI2[0])
    b = O.build_basis(ctx_c, grid, 400.0, (0, 0, 0))
    count = 0
    for n in itertools.product(*[range(-g // 2, (g - 1) // 2 + 1) for g in grid]):
        q = np.array(n, float) @ ctx_c.reciprocal
        count += O.HSQDTM * q @ q < 400.0
    assert count == b.kinetic.size == 725
    # fft_index is the C-order position of G mod n
    g = b.gvec
    assert np.all(b.fft_index == np.mod(g[:, 0], 24) * 576 + np.mod(g[:, 1], 24) * 24 + np.mod(g[:, 2], 24))


def test_spline_against_closed_form():
    gmax, n = 16.0, 100
    qq = np.arange(n) * gmax / n
    for l, rc in ((0, 1.0), (1, 1.0), (2, 0.8)):
        y = qq ** l * np.exp(-0.5 * qq * qq * rc * rc)
        m2 = O.spline_second(y, gmax)
        x = np.linspace(0, 0.9 * gmax, 777)
        exact = x ** l * np.exp(-0.5 * x * x * rc * rc)
        assert np.abs(O.spline_eval(y, m2, gmax, x) - exact).max() < 5e-4
        # interpolating at the nodes
        assert np.abs(O.spline_eval(y, m2, gmax, qq[:-1]) - y[:-1]).max() < 1e-13


def test_ylm_addition_theorem():
    rng = np.random.default_rng(3)
    th, ph = np.arccos(rng.uniform(-1, 1, 50)), rng.uniform(-np.pi, np.pi, 50)
    for l in range(4):
        tot = sum(O.real_ylm(l, m, th, ph) ** 2 for m in range(-l, l + 1))
        assert np.abs(tot - (2 * l + 1) / (4 * np.pi)).max() < 1e-13


def test_fft_sign_and_direct_dft():
    rng = np.random.default_rng(5)
    grid = (4, 6, 5)
    x = rng.standard_normal(grid) + 1j * rng.standard_normal(grid)
    idx = np.array(list(itertools.product(*[range(n) for n in grid])))
    ph = np.exp(-2j * np.pi * (idx[:, None, :] * idx[None, :, :] / np.array(grid)).sum(-1))
    direct = ph @ x.ravel()
    assert np.abs(O.fft3_forward(x, grid) - direct).max() < 1e-12
    assert np.abs(O.fft3_backward(x, grid) - ph.conj() @ x.ravel()).max() < 1e-12


def test_spectrum_special_cases():
    crystal, grid, vals = O.read_chgcar(SI2[0])
    pots = O.read_potcar(SI2[1])
    basis = O.build_basis(crystal, grid, 400.0, (0, 0, 0))
    nopaw = []
    # Veff = 0 and no projectors: eigenvalues are the sorted kinetic energies
    h, s = O.assemble_dense(np.zeros(np.prod(grid)), basis, nopaw)
    w, _, _, _ = O.solve_dense(h, s, 40)
    assert np.abs(w - np.sort(basis.kinetic)[:40]).max() < 1e-10
    rng = np.random.default_rng(1)
    v = rng.standard_normal(np.prod(grid))
    # without projectors a constant shift is exact
    h0, s0 = O.assemble_dense(v, basis, nopaw)
    h0b, s0b = O.assemble_dense(v + 1.75, basis, nopaw)
    assert np.abs(O.solve_dense(h0b, s0b, 20)[0] - O.solve_dense(h0, s0, 20)[0] - 1.75).max() < 1e-10


def test_planted_phase_sign_is_caught():
    """A plausible mistake (wrong sign of the projector phase) must fail the
    golden H elements: the check is not a tautology."""
    g = golden("si2_pbe")
    ctx = O.Context(*SI2, g["encut"], g["xc"])
    basis, paw = ctx.prepare()
    cart = ctx.crystal.positions @ ctx.crystal.lattice
    for sp in paw:  # exp(+i q.R) instead of exp(-i q.R)
        for iat in range(sp["natoms"]):
            sp["beta"][iat] *= np.exp(2j * (basis.q @ cart[iat]))[None, :]
    h, s = O.assemble_dense(ctx.veff, basis, paw)
    hs = np.array(g["hs_samples"])
    i, j = hs[:, 0].astype(int) - 1, hs[:, 1].astype(int) - 1
    assert np.abs(h[i, j] - (hs[:, 2] + 1j * hs[:, 3])).max() > 1e-6


def test_acc_restatement_matches_dense():
    g = golden("si2_pbe")
    ctx = O.Context(*SI2, g["encut"], g["xc"])
    basis, paw = ctx.prepare()
    w, x, it, res = O.acc_solve(lambda X: O.apply_hs(ctx.veff, basis, paw, X), basis.kinetic, 16, tol=1e-7,
                                maxiter=200)
    assert res <= 1e-7
    assert np.abs(w - np.array(g["eigenvalues"])).max() < 1e-9
