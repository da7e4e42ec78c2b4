"""Band energy at a frozen potential as a torch.autograd.Function.

    E(R) = w * sum_n f_n eps_n(R)

where eps_n(R) are the lowest eigenvalues of H(R) psi = eps S(R) psi at one
k point, the projectors beta_I(k+G) e^{-i (k+G).R_I} follow the atom
positions R (Cartesian, Angstrom, atoms in species order) and Veff stays
that of the context's density.  Forward solves through the C ABI (dense
EVD, or the all-band solver in FP64 or mixed precision); backward is the
Hellmann-Feynman gradient dE/dR_I = -F_I with F from the adjoint projector
kernel (half_amd_nonlocal_forces):

    F_I,a = -2 w sum_n f_n Re[((D_I - eps_n Q) c_n) . conj(dc_n,a)],
    c = B^H psi,   dc_a = B^H (i q_a psi)

plus -w sum_n f_n c^H (dD/dR) c when the context builds D from the
potential (reference: src/half_forces.F90:9-73).  The gradient is exact for
the frozen-potential band energy as long as the occupied bands are
separated from the empty ones.

The positions are device-side state of the context: each forward call moves
the projectors of `hctx` to the given positions and leaves them there.

torch is used only for its autograd graph; every number comes from the
HIP library.
"""
from __future__ import annotations

import numpy as np
import torch


class BandEnergy(torch.autograd.Function):
    """BandEnergy.apply(positions, hctx, kpoint, occupations, kweight, solver, tol)

    positions: (natoms, 3) float64 tensor, Cartesian Angstrom.
    hctx: half_amd.Context.  kpoint: 3 fractional coordinates.
    occupations: per-band occupations (len = number of bands solved).
    solver: "dense" (EVD), "acc" (FP64 all-band) or "mixed" (FP32 H.psi
    iterations, FP64 correction sweeps).  tol: residual tolerance of the
    all-band forms (eV).
    """

    @staticmethod
    def forward(ctx, positions, hctx, kpoint, occupations, kweight=1.0, solver="dense", tol=1e-7):
        if positions.dim() != 2 or positions.shape[1] != 3:
            raise ValueError("positions must be (natoms, 3)")
        info = hctx.system_info()
        if positions.shape[0] != info["nions"]:
            raise ValueError(f"positions hold {positions.shape[0]} atoms, the context {info['nions']}")
        pos = positions.detach().to("cpu", torch.float64).numpy()
        frac = pos @ np.linalg.inv(info["lattice"])
        hctx.set_projector_positions(frac)
        occ = np.ascontiguousarray(occupations, np.float64)
        nb = len(occ)
        k = tuple(float(x) for x in kpoint)
        if solver == "dense":
            w, v, _, _ = hctx.solve(nb, k, vectors=True)
        elif solver in ("acc", "mixed"):
            w, v, _ = hctx.solve_acc(nb, k, mixed=(solver == "mixed"), tol=tol, maxiter=400,
                                     extra=max(4, nb // 8), vectors=True)
        else:
            raise ValueError(f"unknown solver {solver!r}")
        forces = hctx.nonlocal_forces(w, occ, v, k, kweight)
        ctx.save_for_backward(torch.from_numpy(-forces))
        ctx.eigenvalues = w
        return positions.new_tensor(kweight * float(np.dot(occ, w)))

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out * grad.to(grad_out), None, None, None, None, None, None


def band_energy(positions, hctx, kpoint, occupations, kweight=1.0, solver="dense", tol=1e-7):
    """Functional form of BandEnergy.apply."""
    return BandEnergy.apply(positions, hctx, kpoint, occupations, kweight, solver, tol)
