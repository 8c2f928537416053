"""Differentiable leapfrog runs for PyTorch: a torch.autograd.Function over clmi_adjoint_lf.

    phiT, piT, bgT = leapfrog_run(coeffs, phi0, pi0, spec, N, dx, dt, nsteps)

runs, on the GPU through the C ABI, the model initialiser's background step (setAllAverages and
aDot from the Friedmann constraint, initializers/modelinitializer.h:71-80), the LeapFrog
constructor, `nsteps` leapfrog steps and a final sync (evolvers/leapfrog.h:29-116).  `coeffs` are the
coefficients of the model's potential monomials (ModelSpec.terms order), `phi0`/`pi0` the initial
fields, shape (n_scalars, N, N, N), float64.  bgT = (aI, aDotI) at the end.  The backward pass is the
library's hand-written reverse-mode leapfrog with checkpointed recomputation; gradients flow to
coeffs, phi0 and pi0.

`chi_energy_fraction_and_band` is the observable used for the lphi4 sensitivity runs: the final
fraction of the total energy held by field 1 (kinetic + gradient, definitions/energies.h:37-45 over
Energies::rho, energies.h:217-239) plus the variance of field 1 carried by modes with k in a band.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch

from .clmi import CLMI_F64, FLD_S, PI_S, Lattice, ModelSpec


def _spec_with(spec: ModelSpec, coeffs) -> ModelSpec:
    terms = [(float(c), list(e)) for c, (_, e) in zip(coeffs, spec.terms)]
    return dataclasses.replace(spec, terms=terms)


def _prepare(spec, coeffs, phi0, pi0, N, dx, dt, precision, expansion, device):
    lat = Lattice(_spec_with(spec, coeffs), N, dx, dt, precision, expansion=expansion, device=device)
    for f in range(spec.n_scalars):
        lat.upload(FLD_S, f, 0, phi0[f])
        lat.upload(PI_S, f, 0, pi0[f])
    return lat


class LeapfrogRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coeffs, phi0, pi0, spec, N, dx, dt, nsteps, precision, expansion, checkpoint_every, device):
        c = coeffs.detach().cpu().numpy().astype(np.float64)
        p0 = phi0.detach().cpu().numpy().astype(np.float64)
        q0 = pi0.detach().cpu().numpy().astype(np.float64)
        lat = _prepare(spec, c, p0, q0, N, dx, dt, precision, expansion, device)
        try:
            if expansion:
                _, _, rhs = lat.hubble_constraint()
                lat.set_background(aDotI=math.sqrt(rhs))
            else:
                lat.set_all_averages()
            lat.lf_init()
            lat.lf_evolve(nsteps)
            lat.lf_sync()
            phiT = np.stack([lat.download(FLD_S, f, 0) for f in range(spec.n_scalars)])
            piT = np.stack([lat.download(PI_S, f, 0) for f in range(spec.n_scalars)])
            bg = lat.background()
        finally:
            lat.close()
        ctx.save_for_backward(coeffs, phi0, pi0)
        ctx.args = (spec, N, dx, dt, nsteps, precision, expansion, checkpoint_every, device)
        out_dev = phi0.device
        return (torch.from_numpy(phiT).to(out_dev), torch.from_numpy(piT).to(out_dev),
                torch.tensor([bg["aI"], bg["aDotI"]], dtype=torch.float64, device=out_dev))

    @staticmethod
    def backward(ctx, g_phi, g_pi, g_bg):
        coeffs, phi0, pi0 = ctx.saved_tensors
        spec, N, dx, dt, nsteps, precision, expansion, checkpoint_every, device = ctx.args
        nf = spec.n_scalars
        zeros = np.zeros((nf, N, N, N))
        sf = g_phi.detach().cpu().numpy() if g_phi is not None else zeros
        sp = g_pi.detach().cpu().numpy() if g_pi is not None else zeros
        sb = g_bg.detach().cpu().numpy() if g_bg is not None else np.zeros(2)
        lat = _prepare(spec, coeffs.detach().cpu().numpy(), phi0.detach().cpu().numpy(),
                       pi0.detach().cpu().numpy(), N, dx, dt, precision, expansion, device)
        try:
            gf, gp, gc, _ = lat.adjoint_lf(nsteps, list(sf), list(sp), (float(sb[0]), float(sb[1])),
                                           checkpoint_every=checkpoint_every, bg_from_fields=True)
        finally:
            lat.close()
        dev = phi0.device
        return (torch.tensor(gc, dtype=torch.float64, device=coeffs.device), torch.from_numpy(np.stack(gf)).to(dev),
                torch.from_numpy(np.stack(gp)).to(dev), None, None, None, None, None, None, None, None, None)


def leapfrog_run(coeffs, phi0, pi0, spec: ModelSpec, N: int, dx: float, dt: float, nsteps: int,
                 precision: int = CLMI_F64, expansion: bool = True, checkpoint_every: int = 0, device: int = 0):
    return LeapfrogRun.apply(coeffs, phi0, pi0, spec, N, dx, dt, nsteps, precision, expansion, checkpoint_every,
                             device)


def potential_torch(spec: ModelSpec, coeffs, phi):
    """V(phi) = sum_t coeffs[t] prod_f phi_f^e_tf for phi of shape (n_scalars, ...)."""
    v = torch.zeros_like(phi[0])
    for t, (_, e) in enumerate(spec.terms):
        p = coeffs[t] * torch.ones_like(phi[0])
        for f, ef in enumerate(e):
            if ef:
                p = p * phi[f] ** ef
        v = v + p
    return v


def grad2_torch(f, dx):
    """Sum over directions of the squared forward differences (definitions/fieldfunctionals.h:45-56)."""
    return sum(((torch.roll(f, -1, dims=d) - f) / dx) ** 2 for d in range(3))


def chi_energy_fraction_and_band(phiT, piT, bgT, coeffs, spec: ModelSpec, dx: float, kIR: float,
                                 band=(0.5, 1.5), field: int = 1):
    """E_field / rho + sum over modes with band[0] <= |n| kIR <= band[1] of |f_n|^2 / N^6."""
    a = bgT[0]
    nf = spec.n_scalars
    kin = [0.5 * torch.mean(piT[f] ** 2) * a ** -6 for f in range(nf)]
    grad = [0.5 * torch.mean(grad2_torch(phiT[f], dx)) * a ** -2 for f in range(nf)]
    rho = sum(kin) + sum(grad) + torch.mean(potential_torch(spec, coeffs, phiT))
    frac = (kin[field] + grad[field]) / rho
    N = phiT.shape[-1]
    fk = torch.fft.fftn(phiT[field])
    n = torch.fft.fftfreq(N, d=1.0 / N, dtype=torch.float64)
    nx, ny, nz = torch.meshgrid(n, n, n, indexing="ij")
    k = torch.sqrt(nx ** 2 + ny ** 2 + nz ** 2) * kIR
    mask = ((k >= band[0]) & (k <= band[1])).to(torch.float64)
    power = torch.sum(mask * (fk.real ** 2 + fk.imag ** 2)) / N ** 6
    return frac + power


def mode_gradients(grad_field: np.ndarray, modes) -> np.ndarray:
    """dL/dRe and dL/dIm of the half-space Fourier amplitudes f_n (irfftn convention) of one initial
    field, from dL/dphi(x): perturbing Re f_n (with its conjugate partner) moves phi by
    (2/N^3) cos(n.x), Im f_n by -(2/N^3) sin(n.x).  modes: (M, 3) integer array of n with
    0 < n_z < N/2."""
    N = grad_field.shape[0]
    g = np.fft.fftn(grad_field)
    out = []
    for n in np.asarray(modes):
        c = g[tuple(int(x) % N for x in n)]
        out.append((2.0 / N ** 3 * c.real, 2.0 / N ** 3 * c.imag))
    return np.array(out)
