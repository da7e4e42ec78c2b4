"""Host logic of the autograd harness (half_amd.autograd) against a
stand-in context with a closed-form spectrum, no GPU: Cartesian <->
fractional conversion, the Hellmann-Feynman sign convention (gradient =
-force) and composition inside a torch graph, checked with
torch.autograd.gradcheck."""
import numpy as np
import torch

from half_amd.autograd import BandEnergy, band_energy


class HarmonicContext:
    """eps_n(R) = c_n + (kappa_n / 2) sum_I |R_I - R0_I|^2 with R = frac A,
    and the matching analytic force -sum_n w f_n kappa_n (R - R0)."""

    def __init__(self, lattice, r0):
        self.lattice = np.asarray(lattice, float)
        self.r0 = np.asarray(r0, float)
        self.frac = None
        self.c = np.array([-3.0, -1.0, 0.5, 2.0])
        self.kappa = np.array([1.0, 2.5, 0.3, 4.0])

    def system_info(self):
        return {"nions": len(self.r0), "lattice": self.lattice}

    def set_projector_positions(self, frac):
        self.frac = np.array(frac, float)

    def _disp(self):
        return self.frac @ self.lattice - self.r0

    def solve(self, nb, k, vectors=True):
        d2 = float((self._disp() ** 2).sum())
        w = self.c[:nb] + 0.5 * self.kappa[:nb] * d2
        return w, np.zeros((5, nb), complex), 0.0, 0.0

    def nonlocal_forces(self, w, occ, v, k, kweight=1.0):
        return -kweight * float(np.dot(occ, self.kappa[: len(occ)])) * self._disp()


def _setup():
    lat = np.array([[5.1, 0.0, 0.0], [0.3, 5.2, 0.0], [-0.8, 0.1, 5.3]])
    r0 = np.array([[0.1, 0.2, 0.3], [1.4, 1.2, 1.1], [2.5, 0.3, 3.9]])
    return HarmonicContext(lat, r0)


def test_gradcheck_closed_form():
    ctx = _setup()
    pos = torch.tensor(ctx.r0 + 0.07, dtype=torch.float64, requires_grad=True)
    occ = np.array([2.0, 2.0, 1.0])
    f = lambda p: BandEnergy.apply(p, ctx, (0, 0, 0), occ, 0.5, "dense", 1e-7)
    assert torch.autograd.gradcheck(f, (pos,), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_value_and_positions_round_trip():
    ctx = _setup()
    pos = torch.tensor(ctx.r0 + 0.1, dtype=torch.float64)
    occ = np.array([2.0, 2.0])
    e = band_energy(pos, ctx, (0, 0, 0), occ, kweight=0.25)
    assert np.allclose(ctx.frac @ ctx.lattice, pos.numpy(), atol=1e-13)
    d2 = 0.01 * pos.numel()
    assert abs(e.item() - 0.25 * float(np.dot(occ, ctx.c[:2] + 0.5 * ctx.kappa[:2] * d2))) < 1e-12


def test_inside_graph():
    ctx = _setup()
    pos = torch.tensor(ctx.r0 - 0.05, dtype=torch.float64, requires_grad=True)
    occ = np.array([2.0, 1.0])
    loss = 3.0 * band_energy(pos, ctx, (0, 0, 0), occ) + (pos ** 3).sum()
    loss.backward()
    expect = 3.0 * float(np.dot(occ, ctx.kappa[:2])) * (pos.detach().numpy() - ctx.r0) + 3 * pos.detach().numpy() ** 2
    assert np.abs(pos.grad.numpy() - expect).max() < 1e-12


def test_rejects_bad_shapes():
    ctx = _setup()
    import pytest
    with pytest.raises(ValueError):
        band_energy(torch.zeros(2, 3, dtype=torch.float64), ctx, (0, 0, 0), [2.0])
    with pytest.raises(ValueError):
        band_energy(torch.zeros(3, 3, dtype=torch.float64), ctx, (0, 0, 0), [2.0], solver="lobpcg")
