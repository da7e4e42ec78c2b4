"""k-point mesh batching: all-band solves of every k point of a mesh in
lockstep (one launch per operator stage for all k, batched Rayleigh-Ritz
on one-workgroup-per-k kernels, whole iterations replayed from hipGraphs)
against per-k dense solves and the reference's golden eigenvalues.

Tolerances: FP64 <= 1e-9 eV per band, mixed <= 1e-8 eV per band (the
per-k solver's bars); graph replay against eager launches bit-identical
(same kernels, same order)."""
import numpy as np
import pytest
from conftest import golden, require_same_inputs, workload_files

import half_amd as H

pytestmark = pytest.mark.gpu

SI2 = ("tests/golden/inputs/si2.CHGCAR", "tests/golden/inputs/si2.POTCAR")
# Gamma (golden si2_pbe), the golden si2_k point, and two mesh points with
# fewer and more plane waves
SI2_MESH = [(0.0, 0.0, 0.0), (0.1, 0.2, 0.3), (0.25, 0.0, 0.0), (0.5, 0.5, 0.0)]


@pytest.fixture(scope="module")
def si2ctx():
    g = golden("si2_pbe")
    print("[gpu] create si2 kmesh", flush=True)
    ctx = H.Context(SI2[0], SI2[1], g["encut"], g["xc"], H.BACKEND_CUDA, H.SOLVER_EVD)
    yield ctx
    ctx.close()


def _dense(ctx, nb, ks):
    return np.array([ctx.solve(nb, k, vectors=False)[0] for k in ks])


def test_kmesh_ragged_sizes(si2ctx):
    sizes = [si2ctx.basis_size(k) for k in SI2_MESH]
    assert len(set(sizes)) > 1, sizes  # the mesh exercises ragged plane-wave counts


def test_kmesh_fp64_matches_dense_and_golden(si2ctx):
    nb = 16
    ref = _dense(si2ctx, nb, SI2_MESH)
    assert np.abs(ref[0] - np.array(golden("si2_pbe")["eigenvalues"])).max() < 1e-9
    assert np.abs(ref[1] - np.array(golden("si2_k")["eigenvalues"])).max() < 1e-9
    w, st = si2ctx.solve_kmesh(nb, SI2_MESH, tol=1e-7, maxiter=200)
    print("[gpu] si2 kmesh fp64", st, np.abs(w - ref).max(), flush=True)
    assert w.shape == (len(SI2_MESH), nb)
    assert st["max_residual_eV"] <= 1e-7
    assert np.abs(w - ref).max() < 1e-9
    hist = si2ctx.solve_history()
    assert len(hist) == st["iterations"] + 1 or len(hist) == st["iterations"]


def test_kmesh_mixed_matches_dense(si2ctx):
    nb = 16
    ref = _dense(si2ctx, nb, SI2_MESH)
    w, st = si2ctx.solve_kmesh(nb, SI2_MESH, mixed=True, tol=1e-6, maxiter=200)
    print("[gpu] si2 kmesh mixed", st, np.abs(w - ref).max(), flush=True)
    assert st["fp64_iterations"] >= 1
    assert np.abs(w - ref).max() < 1e-8


@pytest.mark.parametrize("mixed", [False, True])
def test_kmesh_graph_replay_identical(si2ctx, mixed):
    nb = 12
    wg, sg = si2ctx.solve_kmesh(nb, SI2_MESH, mixed=mixed, tol=1e-6, maxiter=200, extra=4, graph=True)
    we, se = si2ctx.solve_kmesh(nb, SI2_MESH, mixed=mixed, tol=1e-6, maxiter=200, extra=4, graph=False)
    assert sg["iterations"] == se["iterations"]
    assert sg["band_applications"] == se["band_applications"]
    assert np.array_equal(wg, we)


def test_kmesh_mixed_switch_below_fp32_floor(si2ctx):
    """The mesh solve's FP32 phase also ends on stagnation when switch_tol is
    out of FP32's reach."""
    nb = 16
    ref = _dense(si2ctx, nb, SI2_MESH)
    w, st = si2ctx.solve_kmesh(nb, SI2_MESH, mixed=True, tol=1e-6, switch_tol=1e-12, maxiter=200)
    print("[gpu] si2 kmesh mixed, switch below the FP32 floor", st, flush=True)
    assert 1 <= st["fp64_iterations"] < st["iterations"]
    assert st["max_residual_eV"] <= 1e-6
    assert np.abs(w - ref).max() < 1e-8


@pytest.mark.parametrize("mixed", [False, True])
def test_kmesh_soft_locking(si2ctx, mixed):
    """Soft locking in the mesh solve (one correction width for all k points,
    padded per k point with its converged bands) and the full 2m subspace
    reach the same eigenvalues; locking narrows the subspace problems, and
    graph replay across width changes matches eager launches bit for bit."""
    nb = 16
    ref = _dense(si2ctx, nb, SI2_MESH)
    tol, bar = (1e-6, 1e-8) if mixed else (1e-7, 1e-9)
    saved = H.get_tuning("solver.lock")
    H.set_tuning("solver.lock", 0)
    try:
        wf, sf = si2ctx.solve_kmesh(nb, SI2_MESH, mixed=mixed, tol=tol, maxiter=200, extra=4)
    finally:
        H.set_tuning("solver.lock", saved)
    wl, sl = si2ctx.solve_kmesh(nb, SI2_MESH, mixed=mixed, tol=tol, maxiter=200, extra=4)
    we, se = si2ctx.solve_kmesh(nb, SI2_MESH, mixed=mixed, tol=tol, maxiter=200, extra=4, graph=False)
    print(f"[gpu] si2 kmesh mixed={mixed} full", sf, "locked", sl, flush=True)
    assert sf["rr_order_mean"] == 2 * (nb + 4)
    assert sl["rr_order_mean"] < 2 * (nb + 4)
    for w, st in ((wf, sf), (wl, sl)):
        assert st["max_residual_eV"] <= tol
        assert np.abs(w - ref).max() < bar
    assert sl["iterations"] == se["iterations"] and np.array_equal(wl, we)


@pytest.mark.parametrize("mixed", [False, True])
def test_kmesh_small_kernels_match_library(si2ctx, mixed):
    """The mesh's subspace problems on the one-workgroup kernels with whole
    iterations replayed from hipGraphs (solver.smalleig 1, default) against
    rocSOLVER's batched factorisations and eigensolves with eager launches
    (0), and with the correction-block metric in the eigen-decomposition form
    at every k point and iteration (3): the same eigenvalues at the per-k
    dense solves' bar."""
    nb = 16
    ref = _dense(si2ctx, nb, SI2_MESH)
    tol, bar = (1e-6, 1e-8) if mixed else (1e-7, 1e-9)
    saved = H.get_tuning("solver.smalleig")
    out = {}
    try:
        for mode in (0, 1, 3):
            H.set_tuning("solver.smalleig", mode)
            print(f"[gpu] si2 kmesh mixed={mixed} solver.smalleig={mode}", flush=True)
            out[mode] = si2ctx.solve_kmesh(nb, SI2_MESH, mixed=mixed, tol=tol, maxiter=200, extra=4)
    finally:
        H.set_tuning("solver.smalleig", saved)
    for mode, (w, st) in out.items():
        print(f"[gpu]   mode {mode}: {st} |dw| {np.abs(w - ref).max():.2e}", flush=True)
        assert st["max_residual_eV"] <= tol
        assert np.abs(w - ref).max() < bar
    assert out[0][1]["metric_fallbacks"] == 0 and out[1][1]["metric_fallbacks"] == 0
    # every iteration's metric took the eigen-decomposition form
    assert out[3][1]["metric_fallbacks"] >= out[3][1]["iterations"] - 2


def test_kmesh_overlap_matches_in_order(si2ctx):
    """The mesh operator with its projector stage on the second stream gives
    the eigenvalues of the in-order form."""
    try:
        H.set_tuning("operator.overlap", 0)
        w0, s0 = si2ctx.solve_kmesh(12, SI2_MESH, mixed=True, tol=1e-7, maxiter=200, extra=4)
    finally:
        H.set_tuning("operator.overlap", 1)
    w1, s1 = si2ctx.solve_kmesh(12, SI2_MESH, mixed=True, tol=1e-7, maxiter=200, extra=4)
    assert np.abs(w0 - w1).max() < 1e-10


def test_kmesh_single_point_matches_acc(si2ctx):
    k = SI2_MESH[1]
    w1, st1 = si2ctx.solve_kmesh(16, [k], tol=1e-7, maxiter=200)
    w0, _, _ = si2ctx.solve_acc(16, k, tol=1e-7, maxiter=200)
    assert np.abs(w1[0] - w0).max() < 1e-10


def test_kmesh_argument_errors(si2ctx):
    with pytest.raises(H.HalfError):
        si2ctx.solve_kmesh(16, np.zeros((0, 3)))
    with pytest.raises(H.HalfError):
        si2ctx.solve_kmesh(400, SI2_MESH)  # 2 x 400 > plane waves
    w, _ = si2ctx.solve_kmesh(4, SI2_MESH[:1], tol=1e-6, maxiter=100)  # context still usable
    assert np.all(np.isfinite(w))


def test_kmesh_si8k(tmp_path_factory):
    """8-atom cubic cell, 40^3 box, about 5900 plane waves per k point: the
    golden reference point k = (1/6, -1/3, 1/2) of the 6x6x6 mesh solved
    together with two more mesh points."""
    g = golden("si8k_k")
    chg, pot = workload_files("si8k", str(tmp_path_factory.mktemp("si8kmesh")))
    require_same_inputs(g, chg, pot)
    print("[gpu] create si8k kmesh", flush=True)
    ctx = H.Context(chg, pot, g["encut"], g["xc"], H.BACKEND_CUDA, H.SOLVER_EVD)
    try:
        nb = g["nbands"]
        ks = [tuple(g["kpoint"]), (0.0, 0.0, 0.0), (1 / 6, 1 / 6, 1 / 6)]
        ref0 = np.array(g["eigenvalues"])
        w, st = ctx.solve_kmesh(nb, ks, tol=1e-6, maxiter=300, extra=8)
        print("[gpu] si8k kmesh fp64", st, flush=True)
        assert np.abs(w[0] - ref0).max() < 1e-9
        for j in (1, 2):
            wj, _, _ = ctx.solve_acc(nb, ks[j], tol=1e-6, maxiter=300, extra=8)
            assert np.abs(w[j] - wj).max() < 1e-9
        wm, st = ctx.solve_kmesh(nb, ks, mixed=True, tol=1e-6, maxiter=300, extra=8)
        print("[gpu] si8k kmesh mixed", st, np.abs(wm - w).max(), flush=True)
        assert st["fp64_iterations"] >= 1
        assert np.abs(wm - w).max() < 1e-8
    finally:
        ctx.close()


def test_kmesh_full_216(tmp_path_factory):
    """The whole Gamma-centred 6x6x6 mesh of the 8-atom cell (216 k points,
    5,856-5,910 plane waves, 40^3 box, 32 bands + 8 buffer bands) in one
    batched mixed-precision solve (graph-replayed correction block): the
    golden k point against the reference's dense eigenvalues, and a spread of
    other k points against per-k FP64 all-band solves, <= 1e-8 eV."""
    from half_amd import synth
    g = golden("si8k_k")
    chg, pot = workload_files("si8k", str(tmp_path_factory.mktemp("si8kfull")))
    require_same_inputs(g, chg, pot)
    ks = np.asarray(synth.workload("si8k").kpoints, float)
    assert len(ks) == 216
    jg = int(np.argmin(np.abs(ks - np.array(g["kpoint"])).sum(1)))
    assert np.abs(ks[jg] - np.array(g["kpoint"])).max() < 1e-12
    print("[gpu] create si8k full mesh", flush=True)
    ctx = H.Context(chg, pot, g["encut"], g["xc"], H.BACKEND_CUDA, H.SOLVER_EVD)
    try:
        sizes = [ctx.basis_size(tuple(k)) for k in ks]
        nb = g["nbands"]
        w, st = ctx.solve_kmesh(nb, ks, mixed=True, tol=1e-6, maxiter=300, extra=8)
        print("[gpu] si8k 216-k mixed", st, f"npw {min(sizes)}-{max(sizes)}", flush=True)
        assert w.shape == (216, nb) and st["fp64_iterations"] >= 1 and st["max_residual_eV"] <= 1e-6
        assert np.abs(w[jg] - np.array(g["eigenvalues"])).max() < 1e-8
        for j in (0, 1, 43, int(np.argmin(sizes)), int(np.argmax(sizes)), 215):
            wj, _, _ = ctx.solve_acc(nb, tuple(ks[j]), tol=1e-6, maxiter=300, extra=8)
            print(f"[gpu] k {j} npw {sizes[j]} |batched - per-k| {np.abs(w[j] - wj).max():.2e}", flush=True)
            assert np.abs(w[j] - wj).max() < 1e-8
    finally:
        ctx.close()
