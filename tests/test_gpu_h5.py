"""VASP HDF5 files through the library (reference: src/half_chgcar.F90:25,
src/half_potcar.F90:34, src/half_vaspwave.F90:94-155,
src/half_hdf5_bridge.c:10-130).

Inputs: a vaspout.h5 holding the 2-atom Si cell's structure, charge and
POTCAR text (written by the test's own HDF5 access) gives a context whose
Veff and eigenvalues are bit-identical to the CHGCAR/POTCAR context's, for
every layout the reference's reader takes (structure group, fractional or
Cartesian positions, charge with or without the spin axis).

Export: vaspwave.h5 from solved bands at two k points, read back from
outside: structure, charge, ENCUT, Fermi level, eigenvalues, occupations / 2,
plane-wave counts, k points, and the coefficients in the oracle's VASP order
rounded to single precision (bit-exact); a context built from the exported
file solves to the same eigenvalues."""
import numpy as np
import pytest
from conftest import golden

import h5util
import half_amd as H
import half_oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not h5util.available(), reason="HDF5 1.10 C library not available")]

SI2 = ("tests/golden/inputs/si2.CHGCAR", "tests/golden/inputs/si2.POTCAR")
KS = [(0.0, 0.0, 0.0), (0.1, 0.2, 0.3)]


@pytest.fixture(scope="module")
def ref():
    g = golden("si2_k")
    print("[gpu] h5: CHGCAR context", flush=True)
    ctx = H.Context(SI2[0], SI2[1], g["encut"], g["xc"])
    out = {"encut": g["encut"], "xc": g["xc"], "veff": ctx.veff(), "info": ctx.system_info(),
           "solves": [ctx.solve(8, k) for k in KS]}
    ctx.close()
    return out


@pytest.mark.parametrize("layout", ["structure", "input/poscar", "cartesian", "no_spin_axis"])
def test_vaspout_input(tmp_path, ref, layout):
    crystal, grid, vals = O.read_chgcar(SI2[0])
    pot = open(SI2[1]).read()
    path = tmp_path / "vaspout.h5"
    h5util.write_vaspout(path, crystal.lattice, crystal.species, crystal.counts, crystal.positions, grid, vals,
                         pot, group="input/poscar" if layout == "input/poscar" else "structure/positions",
                         direct=layout != "cartesian", spin_axis=layout != "no_spin_axis")
    print(f"[gpu] h5: vaspout input ({layout})", flush=True)
    with H.Context(str(path), str(path), ref["encut"], ref["xc"]) as ctx:
        info = ctx.system_info()
        assert info["nions"] == 2 and info["grid"] == (24, 24, 24)
        assert np.array_equal(info["lattice"], ref["info"]["lattice"])
        if layout == "cartesian":  # fractional positions go through r A^-1: rounding only
            assert np.abs(ctx.veff() - ref["veff"]).max() < 1e-10
            w, _, _, _ = ctx.solve(8, KS[1], vectors=False)
            assert np.abs(w - ref["solves"][1][0]).max() < 1e-10
        else:
            assert np.array_equal(ctx.veff(), ref["veff"])
            for k, r in zip(KS, ref["solves"]):
                # eigenvectors on, as in the reference solves: the same
                # eigensolver path, so the comparison can be bitwise
                w, _, _, _ = ctx.solve(8, k)
                assert np.array_equal(w, r[0])
    # the POTCAR may also come as text beside an HDF5 charge file
    with H.Context(str(path), SI2[1], ref["encut"], ref["xc"]) as ctx:
        if layout != "cartesian":
            assert np.array_equal(ctx.veff(), ref["veff"])


def test_vaspout_without_potcar_is_an_io_error(tmp_path, ref):
    crystal, grid, vals = O.read_chgcar(SI2[0])
    path = tmp_path / "charge_only.h5"
    h5util.write_vaspout(path, crystal.lattice, crystal.species, crystal.counts, crystal.positions, grid, vals)
    with pytest.raises(H.HalfError) as e:
        H.Context(SI2[0], str(path), ref["encut"], ref["xc"])
    assert e.value.status == H.IO and "potcar" in e.value.message


def test_vaspwave_export(tmp_path, ref):
    crystal, grid, vals = O.read_chgcar(SI2[0])
    eig = np.array([s[0] for s in ref["solves"]])
    vecs = [s[1] for s in ref["solves"]]
    occ = np.tile(np.array([2.0, 2.0, 2.0, 2.0, 1.25, 0.0, 0.0, 0.0]), (2, 1))
    path = tmp_path / "vaspwave.h5"
    print("[gpu] h5: export", flush=True)
    with H.Context(SI2[0], SI2[1], ref["encut"], ref["xc"]) as ctx:
        ctx.write_vaspwave(path, KS, eig, occ, vecs, fermi=6.5)
        with pytest.raises(H.HalfError) as e:
            ctx.write_vaspwave(tmp_path / "bad.h5", KS, eig, occ * 1.5, vecs, fermi=6.5)
        assert e.value.status == H.INVALID_ARGUMENT and "occupations" in e.value.message
    r = h5util.read
    assert r(path, "version/major", "i4") == 6 and r(path, "version/minor", "i4") == 6
    assert r(path, "version/patch", "i4") == 0
    assert np.array_equal(r(path, "structure/positions/lattice_vectors"), crystal.lattice)
    assert np.array_equal(r(path, "structure/positions/position_ions"), crystal.positions)
    assert r(path, "structure/positions/number_ion_types", "i4").tolist() == crystal.counts
    assert r(path, "structure/positions/ion_types", "s")[0] == crystal.species[0].encode().ljust(16)
    assert r(path, "structure/positions/direct_coordinates", "i4") == 1
    assert r(path, "structure/positions/scale") == 1.0
    assert r(path, "charge/grid", "i4").tolist() == list(grid)
    assert np.array_equal(r(path, "charge/charge").ravel(), vals)
    assert np.array_equal(r(path, "wave/amat"), crystal.lattice)
    assert r(path, "wave/efermi") == 6.5 and r(path, "wave/enmax") == ref["encut"]
    assert (r(path, "wave/rispin"), r(path, "wave/rnb_tot"), r(path, "wave/rnkpts")) == (1.0, 8.0, 2.0)
    npws = []
    for j, k in enumerate(KS):
        grp = f"wave/spin_1/kpoint_{j + 1}"
        b = O.build_basis(crystal, grid, ref["encut"], k)
        n = len(b.kinetic)
        npws.append(n)
        assert r(path, grp + "/num_planewaves", "i4") == n
        assert np.array_equal(r(path, grp + "/vkpt"), np.array(k))
        cel = r(path, grp + "/celtot")
        assert np.array_equal(cel[:, 0], eig[j]) and not cel[:, 1].any()
        assert np.array_equal(r(path, grp + "/fertot"), 0.5 * occ[j])
        wave = r(path, grp + "/wave", "f4")
        assert wave.shape == (8, n, 2)
        v = vecs[j][O.vasp_order(b)]  # npw x nbands in VASP order
        expect = np.stack([v.real.T.astype(np.float32), v.imag.T.astype(np.float32)], -1)
        assert np.array_equal(wave, expect)
    assert r(path, "wave/rdum") == 8.0 * max(npws)
    # the exported file is itself an input
    with H.Context(str(path), SI2[1], ref["encut"], ref["xc"]) as ctx:
        assert np.array_equal(ctx.veff(), ref["veff"])
        w, _, _, _ = ctx.solve(8, KS[1])  # the eigensolver path of the exported solve
        assert np.array_equal(w, eig[1])
