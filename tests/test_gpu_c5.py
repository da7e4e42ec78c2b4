"""The 216-atom Si cell: the 3x3x3 supercell of the 8-atom cell with seeded
+-0.03 A positional noise, Gamma, 245.5 eV, 37,705 plane waves, 96^3 box,
1,728 projectors, PBE, 512 bands (+64 buffer bands) -- the sizes of the
headline workload.  These sizes take code paths no smaller cell reaches:
the 96-point transforms, the five-kernel FP64 local stage (41 active planes),
the LDS-tiled contraction kernels of the projector stage (more than 512
projectors), box batching of a 2 GiB workspace, and Rayleigh-Ritz problems
of order above 256 on rocSOLVER.

Checks:
- Veff and the energy terms against the oracle (PBE pointwise <= 1e-7 eV);
- H.psi / S.psi of a 520-band block, FP64 and FP32, against the oracle's
  matrix-free operator on columns spread over the box batches (FP64
  <= 1e-11 relative, FP32 <= 2e-6), with the x pass (41 planes) as the
  vector-unit convolution and as the Toeplitz product on the matrix cores,
  and (FP64) with the projector contractions' three-MFMA complex products;
- the FP64 all-band solve (tol 1e-6): Ritz pairs satisfy the oracle's
  operator (Rayleigh quotients <= 1e-9 eV, residuals <= 2e-6 eV,
  S-orthonormal); the mixed solve per band <= 1e-8 eV from it;
- band folding: on the supercell without positional noise (LDA) the lowest
  512 eigenvalues of the FP64 all-band solve equal the union of the 8-atom
  cell's dense spectra at the 27 k points (i, j, l)/3 (<= 1e-8 eV), an
  independent pin of the 216-atom solve;
- analytic non-local forces (216 x 3, adjoint projector kernel) against
  central differences (h = 1e-4 A) of the band energy of the lowest 432
  bands under frozen-potential displacements, 3 atoms x 3 axes, <= 1e-5 eV/A.
"""
import numpy as np
import pytest
from conftest import workload_files

import half_amd as H
import half_oracle as O
from half_amd import hip, synth

pytestmark = pytest.mark.gpu

NB, EXTRA = 512, 64
NOCC = 432  # 864 valence electrons


@pytest.fixture(scope="module")
def c5(tmp_path_factory):
    w = synth.workload("si216")
    chg, pot = workload_files("si216", str(tmp_path_factory.mktemp("si216")))
    print("[gpu] create si216", flush=True)
    ctx = H.Context(chg, pot, w.encut, w.xc, H.BACKEND_CUDA, H.SOLVER_EVD)
    print("[gpu] oracle si216 (CPU)", flush=True)
    ref = O.Context(chg, pot, w.encut, w.xc)
    basis, paw = ref.prepare()
    state = {}
    yield w, ctx, ref, basis, paw, state
    ctx.close()


def test_c5_sizes_and_potential(c5):
    w, ctx, ref, basis, paw, _ = c5
    gv, kin = ctx.basis()
    assert gv.shape[0] == basis.kinetic.size == 37705
    assert np.array_equal(gv, basis.gvec) and np.abs(kin - basis.kinetic).max() < 1e-10
    assert ctx.system_info()["grid"] == (96, 96, 96)
    op = ctx.operator((0, 0, 0), 64)
    assert op.nproj == 1728
    op.close()
    v = ctx.veff()
    assert np.abs(v - ref.veff).max() < 1e-7
    eh, exc, exv = ctx.energy_terms()
    print(f"[gpu] si216 e_H {eh - ref.e_h:.2e} e_xc {exc - ref.e_xc:.2e} e_xcpot {exv - ref.e_xc_pot:.2e}", flush=True)
    assert abs(eh - ref.e_h) < 1e-8 and abs(exc - ref.e_xc) < 1e-8 and abs(exv - ref.e_xc_pot) < 1e-8


def test_c5_operator_520_bands(c5):
    """One application to 520 bands, the FP64 form split into box batches by
    its 2 GiB workspace: columns in every batch against the oracle."""
    w, ctx, ref, basis, paw, _ = c5
    n = basis.kinetic.size
    nb = 520
    rng = np.random.default_rng(2026)
    psi = rng.standard_normal((nb, n)) + 1j * rng.standard_normal((nb, n))  # row b = column b
    cols = [0, 1, 137, 260, 333, 450, 518, 519]
    rh, rs = O.apply_hs(ref.veff, basis, paw, psi[cols].T)
    stream = hip.Stream()
    keys = ["local.fp64.xconv", "local.fp32.xconv", "contract.fp64.gram.gauss", "contract.fp64.tall.gauss"]
    saved = {k: H.get_tuning(k) for k in keys}
    try:
        for prec, dt, tol in ((64, np.complex128, 1e-11), (32, np.complex64, 2e-6)):
            op = ctx.operator((0, 0, 0), prec)
            x = hip.Buffer.from_numpy(psi.astype(dt))
            hp, sp = hip.Buffer(x.shape, dt), hip.Buffer(x.shape, dt)
            # x pass on the vector units and as the Toeplitz product on the matrix
            # cores; FP64 projector contractions with four and three MFMAs per
            # complex product
            for xconv, gauss in ((1, 0), (2, 0), (2, 1), (3, 1)) if prec == 64 else ((1, 0), (2, 0), (3, 0)):
                H.set_tuning(f"local.fp{prec}.xconv", xconv)
                H.set_tuning("contract.fp64.gram.gauss", gauss)
                H.set_tuning("contract.fp64.tall.gauss", gauss)
                op.apply_device(nb, x.ptr, hp.ptr, sp.ptr, stream=stream.ptr)
                stream.synchronize()
                h = hp.numpy()[cols].T.astype(np.complex128)
                s = sp.numpy()[cols].T.astype(np.complex128)
                eh = np.abs(h - rh).max() / np.abs(rh).max()
                es = np.abs(s - rs).max() / np.abs(rs).max()
                print(f"[gpu] si216 fp{prec} xconv {xconv} gauss {gauss} 520-band apply: rel |dH psi| {eh:.2e} "
                      f"|dS psi| {es:.2e}", flush=True)
                assert eh < tol and es < tol
            op.close()
            x.free(), hp.free(), sp.free()
    finally:
        for k, v in saved.items():
            H.set_tuning(k, v)


def test_c5_acc_fp64_and_mixed(c5):
    w, ctx, ref, basis, paw, state = c5
    w64, v64, st = ctx.solve_acc(NB, tol=1e-6, maxiter=300, extra=EXTRA, vectors=True)
    print("[gpu] si216 fp64 acc", st, flush=True)
    assert st["max_residual_eV"] <= 1e-6
    assert np.all(np.diff(w64) >= 0)
    state["w64"], state["v64"] = w64, v64
    # the Ritz pairs against the oracle's own operator
    sample = [0, 1, 2, 3, 200, 201, 300, NOCC - 1, NOCC, 505, 506, 507, 508, 509, 510, 511]
    x = v64[:, sample]
    rh, rs = O.apply_hs(ref.veff, basis, paw, x)
    rq = np.einsum("ij,ij->j", x.conj(), rh).real / np.einsum("ij,ij->j", x.conj(), rs).real
    res = np.linalg.norm(rh - rs * w64[sample][None, :], axis=0)
    gram = x.conj().T @ rs
    print(f"[gpu] si216 oracle check: |rq - eps| {np.abs(rq - w64[sample]).max():.2e} "
          f"residual {res.max():.2e} |X^H S X - I| {np.abs(gram - np.eye(len(sample))).max():.2e}", flush=True)
    assert np.abs(rq - w64[sample]).max() < 1e-9
    assert res.max() < 2e-6
    assert np.abs(gram - np.eye(len(sample))).max() < 1e-10
    wm, _, stm = ctx.solve_acc(NB, mixed=True, tol=1e-6, maxiter=300, extra=EXTRA)
    err = np.abs(wm - w64)
    print("[gpu] si216 mixed acc", stm, flush=True)
    print(f"[gpu] si216 mixed vs fp64 per band: max {err.max():.2e} mean {err.mean():.2e} "
          f"bands > 1e-10: {int((err > 1e-10).sum())}", flush=True)
    assert stm["fp64_iterations"] >= 1 and stm["max_residual_eV"] <= 1e-6
    assert err.max() < 1e-8


def test_c5_band_folding(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("fold"))
    sup_chg, prim_chg, pot, ks, sup = synth.write_folded_inputs(d)
    print("[gpu] create folded pair (LDA)", flush=True)
    prim = H.Context(prim_chg, pot, sup.encut, "lda", H.BACKEND_CUDA, H.SOLVER_EVD)
    try:
        sizes = [prim.basis_size(tuple(k)) for k in ks]
        per_k = [prim.solve(40, tuple(k), vectors=False)[0] for k in ks]
    finally:
        prim.close()
    union = np.sort(np.concatenate(per_k))[:NB]
    top40 = min(p[-1] for p in per_k)
    sctx = H.Context(sup_chg, pot, sup.encut, "lda", H.BACKEND_CUDA, H.SOLVER_EVD)
    try:
        assert sctx.basis_size() == sum(sizes)
        ws, _, st = sctx.solve_acc(NB, tol=1e-6, maxiter=300, extra=EXTRA)
    finally:
        sctx.close()
    print("[gpu] folded supercell acc", st, flush=True)
    # every k point's 40th band lies above the 512th of the union, so the
    # union's lowest 512 are complete
    assert union[-1] < top40
    keep = NB - 16  # leave the top of the block out: it may cut a degenerate group
    err = np.abs(ws[:keep] - union[:keep])
    print(f"[gpu] folding: max |eps_216 - eps_8(k)| {err.max():.2e} over {keep} bands", flush=True)
    assert err.max() < 1e-8


def test_c5_forces_finite_difference(c5):
    """Full non-local force vector (216 x 3) from the adjoint kernel against
    central differences of 2 sum_{n < 432} eps_n under frozen-potential
    projector displacements on 3 atoms, all axes."""
    w, ctx, ref, basis, paw, state = c5
    tol = 1e-7
    wv, v, st = ctx.solve_acc(NB, tol=tol, maxiter=400, extra=EXTRA, vectors=True)
    print("[gpu] si216 base solve", st, f"gap eps_{NOCC} - eps_{NOCC - 1} = {wv[NOCC] - wv[NOCC - 1]:.4f} eV",
          flush=True)
    occ = np.zeros(NB)
    occ[:NOCC] = 2.0
    f = ctx.nonlocal_forces(wv, occ, v)
    assert f.shape == (216, 3) and np.all(np.isfinite(f))
    crystal = O.read_chgcar(ctx.chgcar)[0]
    base = crystal.positions.copy()
    inv = np.linalg.inv(crystal.lattice)
    h = 1e-4
    worst = 0.0
    try:
        for atom in (0, 107, 215):
            for axis in range(3):
                e = []
                for sgn in (1, -1):
                    pos = base.copy()
                    pos[atom] += sgn * h * inv[axis]
                    ctx.set_projector_positions(pos)
                    wd, _, std = ctx.solve_acc(NB, tol=tol, maxiter=400, extra=EXTRA)
                    assert std["max_residual_eV"] <= tol
                    e.append(2.0 * wd[:NOCC].sum())
                fd = -(e[0] - e[1]) / (2 * h)
                print(f"[gpu] si216 atom {atom} axis {axis}: analytic {f[atom, axis]:+.8f} fd {fd:+.8f} "
                      f"diff {f[atom, axis] - fd:.2e}", flush=True)
                worst = max(worst, abs(f[atom, axis] - fd))
    finally:
        ctx.set_projector_positions(base)
    assert worst < 1e-5
