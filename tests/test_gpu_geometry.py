"""The host's request geometry as the context's cell.

A DeePAW density file stores its lattice and positions to six decimals; the
host (VASP) enumerates its plane waves on the full-precision cell.  The
reference keeps building from the file, so at a cutoff that a G vector sits
on, the two bases differ (reference: docs/validation/
HFO2_NWAVE_BOUNDARY_MISMATCH.md: 3672 host vs 3676 HALF plane waves at
520 eV).  Here the synthetic HfO2-like cell is scaled so that one vector lies
1e-9 (relative) above 520 eV at k = (0, 1/2, 1/2): the six-decimal header
gives 3728 plane waves, the full-precision cell 3724.

Checks: the mapped solve's diagnostics on the rounded file (both sizes, the
first G vector in one basis only); after half_set_request_geometry the
basis equals the oracle's enumeration on the full cell, the mapped solve
succeeds at the unchanged ENCUT, and every result is bit-identical to a
context read from a file holding the full-precision cell; on the 2-atom Si
cell the eigenvalues match the oracle's dense solve on the full cell
(<= 1e-10 eV); invalid requests leave the context unchanged."""
import numpy as np
import pytest
from conftest import golden

import half_amd as H
import half_oracle as O
from half_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hfo2b(tmp_path_factory):
    b = synth.write_boundary_inputs(str(tmp_path_factory.mktemp("boundary")), "hfo2", 520.0, (0.0, 0.5, 0.5))
    assert (b["npw_full"], b["npw_round"]) == (3724, 3728)
    crystal = O.read_chgcar(b["chgcar_full"])[0]
    b["host_basis"] = O.build_basis(crystal, b["grid"], b["encut"], b["kpoint"])
    return b


def _request(ctx, b, order=None):
    idx = np.arange(len(b["species"])) if order is None else order
    ctx.set_request_geometry(b["lattice"], b["species"][idx], b["positions"][idx], b["grid"],
                             ntypes=int(b["species"].max()))


def _g_key(g):
    return set(map(tuple, np.asarray(g).tolist()))


def test_rounded_header_names_the_difference(hfo2b):
    b = hfo2b
    print("[gpu] hfo2 boundary: rounded header", flush=True)
    with H.Context(b["chgcar_round"], b["potcar"], b["encut"], "pbe") as ctx:
        assert ctx.basis_size(b["kpoint"]) == b["npw_round"]
        own, _ = ctx.basis(b["kpoint"])
        host = b["host_basis"].gvec
        hk = _g_key(host)
        extra = [tuple(g) for g in own.tolist() if tuple(g) not in hk]
        assert len(extra) == 4 and _g_key(host) <= _g_key(own)
        with pytest.raises(H.HalfError) as e:
            ctx.solve_mapped(8, host, b["kpoint"])
        assert e.value.status == H.INVALID_ARGUMENT
        msg = e.value.message
        assert msg.startswith("host and HALF plane-wave bases have different sizes")
        assert "(host 3724, HALF 3728)" in msg
        assert "first differing G vector (%d, %d, %d) in the HALF basis only" % extra[0] in msg, msg
        # same size, different sets: one host vector swapped for an extra one
        swapped = np.concatenate([host[1:], np.array([extra[0]], np.int32)])
        with H.Context(b["chgcar_full"], b["potcar"], b["encut"], "pbe") as full:
            with pytest.raises(H.HalfError) as e:
                full.solve_mapped(8, swapped, b["kpoint"])
            assert e.value.status == H.INVALID_ARGUMENT
            assert e.value.message.startswith("host and HALF plane-wave bases contain different G vectors")
            assert "(%d, %d, %d) in the HALF basis only" % tuple(host[0]) in e.value.message


def test_request_geometry_restores_host_basis(hfo2b):
    b = hfo2b
    host = b["host_basis"]
    perm = np.random.default_rng(3).permutation(host.gvec.shape[0])
    print("[gpu] hfo2 boundary: full-precision file", flush=True)
    with H.Context(b["chgcar_full"], b["potcar"], b["encut"], "pbe") as full:
        w_ref, v_ref, smin_ref, smax_ref = full.solve(8, b["kpoint"])
        veff_ref = full.veff()
        terms_ref = full.energy_terms()
    print("[gpu] hfo2 boundary: rounded file + request geometry", flush=True)
    with H.Context(b["chgcar_round"], b["potcar"], b["encut"], "pbe") as ctx:
        _request(ctx, b)
        info = ctx.system_info()
        assert np.array_equal(info["lattice"], b["lattice"])
        assert info["volume"] == pytest.approx(abs(np.linalg.det(b["lattice"])), rel=1e-14)
        g, kin = ctx.basis(b["kpoint"])
        assert g.shape[0] == b["npw_full"]
        assert np.array_equal(g, host.gvec)
        assert np.abs(kin - host.kinetic).max() <= 1e-12 * host.kinetic.max()
        assert np.array_equal(ctx.veff(), veff_ref)
        assert ctx.energy_terms() == terms_ref
        w, v, smin, smax = ctx.solve_mapped(8, host.gvec[perm], b["kpoint"])
        assert np.array_equal(w, w_ref) and (smin, smax) == (smin_ref, smax_ref)
        assert np.array_equal(v, v_ref[perm])
        req = ctx.get_request_geometry()
        assert req["nions"] == 12 and req["ntypes"] == 2 and req["grid"] == b["grid"]
        assert np.array_equal(req["positions"], b["positions"])
    print("[gpu] hfo2 boundary: ions interleaved by type", flush=True)
    order = np.array([4, 0, 5, 6, 1, 7, 8, 2, 9, 10, 3, 11])
    with H.Context(b["chgcar_round"], b["potcar"], b["encut"], "pbe") as ctx:
        _request(ctx, b, order)
        w, v, _, _ = ctx.solve_mapped(8, host.gvec, b["kpoint"])
        assert np.array_equal(w, w_ref) and np.array_equal(v, v_ref)
        assert np.array_equal(ctx.get_request_geometry()["species"], b["species"][order])


def test_request_geometry_si2_against_oracle(tmp_path):
    """2-atom Si at a general k point: the six-decimal header adds six plane
    waves at 400 eV; with the request geometry the mapped solve's
    eigenvalues equal the oracle's dense solve on the full cell."""
    k = (0.1, 0.2, 0.3)
    b = synth.write_boundary_inputs(str(tmp_path), "si2", 400.0, k)
    assert (b["npw_full"], b["npw_round"]) == (723, 729)
    ref = O.Context(b["chgcar_full"], b["potcar"], 400.0, "pbe")
    basis, paw = ref.prepare(k)
    h, s = O.assemble_dense(ref.veff, basis, paw)
    w_ref, _, smin, smax = O.solve_dense(h, s, 12, vectors=False)
    print("[gpu] si2 boundary: rounded file + request geometry", flush=True)
    with H.Context(b["chgcar_round"], b["potcar"], 400.0, "pbe") as ctx:
        assert ctx.basis_size(k) == 729
        _request(ctx, b)
        w, v, _, _ = ctx.solve_mapped(12, basis.gvec[::-1], k)
        print(f"[gpu] si2 boundary: max |eig - oracle| {np.abs(w - w_ref).max():.2e} eV", flush=True)
        assert np.abs(w - w_ref).max() <= 1e-10
        # the returned vectors are in host order: S-orthonormal on the oracle's
        # matrices once put back in the oracle's order
        vv = v[::-1]
        assert np.abs(vv.conj().T @ s @ vv - np.eye(12)).max() < 1e-10
        assert np.abs(np.diag(vv.conj().T @ h @ vv).real - w_ref).max() < 1e-9


def test_request_geometry_rejections():
    g = golden("si2_k")
    chg, pot = "tests/golden/inputs/si2.CHGCAR", "tests/golden/inputs/si2.POTCAR"
    print("[gpu] si2 request geometry rejections", flush=True)
    with H.Context(chg, pot, g["encut"], g["xc"]) as ctx:
        w0, _, _, _ = ctx.solve(4, g["kpoint"], vectors=False)
        crystal = O.read_chgcar(chg)[0]
        lat, pos = crystal.lattice, crystal.positions
        cases = {
            "density grid": dict(lattice=lat, species=[1, 1], positions=pos, grid=(24, 24, 20)),
            "datasets": dict(lattice=lat, species=[1, 2], positions=pos, grid=(24, 24, 24)),
            "ion count per type": dict(lattice=lat, species=[1, 1, 1], positions=np.vstack([pos, pos[:1]]),
                                       grid=(24, 24, 24)),
            "invalid request geometry values": dict(lattice=np.where(np.eye(3) > 0, np.inf, lat), species=[1, 1],
                                                    positions=pos, grid=(24, 24, 24)),
            "singular": dict(lattice=np.zeros((3, 3)), species=[1, 1], positions=pos, grid=(24, 24, 24)),
        }
        for text, kw in cases.items():
            with pytest.raises(H.HalfError) as e:
                ctx.set_request_geometry(**kw)
            assert e.value.status == H.INVALID_ARGUMENT and text in e.value.message, (text, e.value.message)
        with pytest.raises(H.HalfError) as e:
            ctx.get_request_geometry()
        assert e.value.status == H.UNAVAILABLE
        w1, _, _, _ = ctx.solve(4, g["kpoint"], vectors=False)
        assert np.array_equal(w0, w1)
        # a geometry change drops the context's operator handles
        op = ctx.operator(g["kpoint"], 64)
        ctx.set_request_geometry(lat, [1, 1], pos + 0.01, (24, 24, 24))
        from half_amd import hip
        x = hip.Buffer.from_numpy(np.ones((1, g["npw"]), np.complex128))
        hp, sp = hip.Buffer(x.shape, np.complex128), hip.Buffer(x.shape, np.complex128)
        with pytest.raises(H.HalfError) as e:
            op.apply_device(1, x.ptr, hp.ptr, sp.ptr)
        assert e.value.status == H.INVALID_HANDLE
        w2, _, _, _ = ctx.solve(4, g["kpoint"], vectors=False)
        assert np.abs(w2 - w0).max() > 1e-6  # the moved atoms are used
