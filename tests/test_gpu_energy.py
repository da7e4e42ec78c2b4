"""Harris energy from the library (half_amd_harris_energy) against the
reference CLI's assembly (app/half_cli.F90:896-920) as the reference's own
Fortran computed it for the golden fixtures: occupations, Fermi level, band
energy, Ewald energy, local_g0, the atomic reference and the total, from the
library's own dense eigenvalues, <= 1e-9 eV.  Smearing against the oracle's
occupations (the oracle is pinned to the reference's occupation
known-answer tests)."""
import numpy as np
import pytest
from conftest import golden, require_same_inputs, workload_files

import half_amd as H
import half_oracle as O

pytestmark = pytest.mark.gpu

SI2 = ("tests/golden/inputs/si2.CHGCAR", "tests/golden/inputs/si2.POTCAR")


def _context(case, tmp_path_factory):
    g = golden(case)
    if g["workload"] == "si2":
        chg, pot = SI2
    else:
        chg, pot = workload_files(g["workload"], str(tmp_path_factory.mktemp(case)))
        require_same_inputs(g, chg, pot)
    print(f"[gpu] create {case}", flush=True)
    return g, chg, pot, H.Context(chg, pot, g["encut"], g["xc"], H.BACKEND_CUDA, H.SOLVER_EVD)


@pytest.mark.parametrize("case", ["si2_pbe", "si2_lda", "si2_k", "hfo2_pbe", "si8k_k"])
def test_harris_energy_matches_reference(case, tmp_path_factory):
    g, _, _, ctx = _context(case, tmp_path_factory)
    try:
        w, _, _, _ = ctx.solve(g["nbands"], g["kpoint"], vectors=False)
        assert np.abs(w - g["eigenvalues"]).max() < 1e-9
        t, occ = ctx.harris_energy(w)
        print(f"[gpu] {case} harris {t['harris_energy']:.12f} golden {g['harris_energy']:.12f}", flush=True)
        assert np.array_equal(occ[0], np.array(g["occupations"]))
        assert abs(t["fermi"] - g["fermi"]) < 1e-9
        assert abs(t["band_energy"] - g["band_energy"]) < 1e-9
        assert abs(t["ewald"] - g["ewald"]) < 1e-9
        assert abs(t["local_g0"] - g["local_g0"]) < 1e-9
        assert t["atomic_reference"] == g["atomic_reference"]
        assert abs(t["e_hartree"] - g["e_hartree"]) < 1e-9
        assert abs(t["harris_energy"] - g["harris_energy"]) < 1e-9
        assert t["entropy"] == 0.0 and t["free_energy"] == t["harris_energy"]
        # the golden eigenvalues themselves give the golden total
        tg, _ = ctx.harris_energy(np.array(g["eigenvalues"]))
        assert abs(tg["harris_energy"] - g["harris_energy"]) < 1e-9
    finally:
        ctx.close()


def test_harris_smearing_and_errors(tmp_path_factory):
    g, chg, pot, ctx = _context("si2_pbe", tmp_path_factory)
    try:
        w = np.array(g["eigenvalues"])
        nelect = 8.0
        # two k points with the same bands shifted: weights 0.25 / 0.75
        eig = np.vstack([w, w + 0.37])
        kw = np.array([0.25, 0.75])
        for ismear in (0, -1):
            t, occ = ctx.harris_energy(eig, kw, sigma=0.2, ismear=ismear)
            ro, mu, band, ent = O.compute_occupations(eig, kw, nelect, 0.2, ismear)
            assert abs((kw[:, None] * occ).sum() - nelect) < 1e-10
            assert np.abs(occ - ro).max() < 1e-12
            assert abs(t["fermi"] - mu) < 1e-12 and abs(t["band_energy"] - band) < 1e-10
            assert abs(t["entropy"] - ent) < 1e-12 and t["entropy"] != 0.0
            assert abs(t["free_energy"] - t["harris_energy"] - t["entropy"]) < 1e-12
        # zero temperature over k points: the oracle's degeneracy-grouped filling
        t, occ = ctx.harris_energy(eig, kw)
        ro, mu, band, _ = O.compute_occupations(eig, kw, nelect, 0.0)
        assert np.array_equal(occ, ro) and t["fermi"] == mu and abs(t["band_energy"] - band) < 1e-12
        for kwargs, msg in (({"kweights": [0.5, 0.6]}, "sum to one"), ({"sigma": -1.0}, "sigma"),
                            ({"ismear": 1}, "ISMEAR")):
            with pytest.raises(H.HalfError) as e:
                ctx.harris_energy(eig, **{"kweights": kw, **kwargs})
            assert e.value.status == H.INVALID_ARGUMENT and msg in e.value.message
        with pytest.raises(H.HalfError) as e:  # 3 bands hold 6 of the 8 electrons
            ctx.harris_energy(w[:3])
        assert e.value.status == H.INVALID_ARGUMENT and "electrons" in e.value.message
        with pytest.raises(H.HalfError) as e:  # smearing with the top band occupied
            ctx.harris_energy(w[:5], sigma=0.5)
        assert e.value.status == H.INVALID_ARGUMENT and "increase" in e.value.message
    finally:
        ctx.close()
