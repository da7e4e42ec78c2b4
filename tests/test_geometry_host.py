"""Host side of the request-geometry case: the boundary cell generator
against the oracle's reader and basis enumeration
(reference: src/half_chgcar.F90:12-98, src/half_basis.F90:9-55)."""
import numpy as np
import pytest

import half_oracle as O
from half_amd import synth


@pytest.mark.parametrize("name,encut,k,sizes", [("hfo2", 520.0, (0.0, 0.5, 0.5), (3724, 3728)),
                                                ("si2", 400.0, (0.1, 0.2, 0.3), (723, 729))])
def test_boundary_inputs(tmp_path, name, encut, k, sizes):
    b = synth.write_boundary_inputs(str(tmp_path), name, encut, k)
    assert (b["npw_full"], b["npw_round"]) == sizes
    full, grid, vals_f = O.read_chgcar(b["chgcar_full"])
    rnd, grid_r, vals_r = O.read_chgcar(b["chgcar_round"])
    # 17 significant digits read back as the same doubles; the rounded header
    # is the six-decimal image; the density values are the same
    assert np.array_equal(full.lattice, b["lattice"]) and np.array_equal(full.positions, b["positions"])
    assert np.abs(rnd.lattice - np.round(b["lattice"], 6)).max() < 1e-12
    assert 0 < np.abs(rnd.lattice - full.lattice).max() <= 5e-7
    assert grid == grid_r and np.array_equal(vals_f, vals_r)
    bf = O.build_basis(full, grid, encut, k)
    br = O.build_basis(rnd, grid, encut, k)
    assert (len(bf.kinetic), len(br.kinetic)) == sizes
    # the rounded basis is the full one plus vectors at the cutoff
    sf, sr = set(map(tuple, bf.gvec.tolist())), set(map(tuple, br.gvec.tolist()))
    assert sf < sr
    extra = np.array(sorted(sr - sf))
    q = (extra + np.asarray(k)) @ full.reciprocal
    e_full = O.HSQDTM * np.einsum("ij,ij->i", q, q)
    assert np.all(e_full >= encut) and np.all(e_full < encut * (1 + 1e-6))
