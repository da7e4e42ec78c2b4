"""Host side of the HDF5 inputs and the vaspwave.h5 export: the oracle's
VASP plane-wave order (reference: src/half_vaspwave.F90:142-155) against an
independent sort and brute force, and the test's own HDF5 access (used to
check the library from outside) on a round trip."""
import numpy as np
import pytest

import h5util
import half_oracle as O

SI2 = ("tests/golden/inputs/si2.CHGCAR", "tests/golden/inputs/si2.POTCAR")


@pytest.mark.parametrize("k", [(0.0, 0.0, 0.0), (0.1, 0.2, 0.3)])
def test_vasp_order(k):
    crystal, grid, _ = O.read_chgcar(SI2[0])
    b = O.build_basis(crystal, grid, 400.0, k)
    order = O.vasp_order(b)
    assert sorted(order.tolist()) == list(range(len(b.fft_index)))
    # the same order from the box coordinates: z slowest, x fastest
    n1, n2, n3 = grid
    i1, i2, i3 = b.fft_index // (n2 * n3), (b.fft_index // n3) % n2, b.fft_index % n3
    assert np.array_equal(order, np.lexsort((i1, i2, i3)))
    # brute force on the first entries: the smallest (z, y, x) box coordinates
    box = [(int(i3[j]), int(i2[j]), int(i1[j])) for j in range(len(i1))]
    assert [box[j] for j in order[:20]] == sorted(box)[:20]


@pytest.mark.skipif(not h5util.available(), reason="HDF5 1.10 C library not available")
def test_h5util_round_trip(tmp_path):
    crystal, grid, vals = O.read_chgcar(SI2[0])
    pot = open(SI2[1]).read()
    path = tmp_path / "v.h5"
    h5util.write_vaspout(path, crystal.lattice, crystal.species, crystal.counts, crystal.positions, grid, vals, pot)
    assert np.array_equal(h5util.read(path, "structure/positions/lattice_vectors"), crystal.lattice)
    assert np.array_equal(h5util.read(path, "structure/positions/position_ions"), crystal.positions)
    assert h5util.read(path, "charge/grid", "i4").tolist() == list(grid)
    ch = h5util.read(path, "charge/charge")
    assert ch.shape == (1, grid[2], grid[1], grid[0]) and np.array_equal(ch.ravel(), vals)
    assert h5util.read(path, "input/potcar/content", "s", len(pot.encode()))[0].decode() == pot
    with open(path, "rb") as fh:
        assert fh.read(8) == b"\x89HDF\r\n\x1a\n"


@pytest.mark.skipif(not h5util.available(), reason="HDF5 1.10 C library not available")
@pytest.mark.parametrize("bad", ["extra_types", "system_array"])
def test_malformed_structure_rejected(tmp_path, bad):
    """The library's vaspout.h5 reader checks every extent before reading:
    an ion_types list longer than number_ion_types, or a system name stored
    as an array, is an I/O error with a message (the file is read before the
    device is touched, so this runs without a GPU)."""
    import ctypes as C
    import half_amd as H
    crystal, grid, vals = O.read_chgcar(SI2[0])
    pot = open(SI2[1]).read()
    path = tmp_path / "bad.h5"
    kw = {"extra_types": 3} if bad == "extra_types" else {"system_count": 4}
    h5util.write_vaspout(path, crystal.lattice, crystal.species, crystal.counts, crystal.positions, grid, vals, pot,
                         **kw)
    lib = H.load()
    h = C.c_int64(0)
    e = C.create_string_buffer(256)
    st = lib.half_create_from_files_backend(str(path).encode(), str(path).encode(), 400.0, H.XC_PBE, 0,
                                            H.BACKEND_AUTO, H.SOLVER_EVD, C.byref(h), e, 256)
    assert st == H.IO and h.value == 0
    msg = e.value.decode()
    assert ("ion_types" in msg) if bad == "extra_types" else ("system" in msg), msg
