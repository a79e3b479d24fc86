"""Gradients of the flux or the power density with respect to undulator pole errors,
through torch.autograd.

The forward model is the reference's: a pole-scaled ideal undulator (each of the 2N
interior half-period poles scaled by 1 + delta_p), the RK4 trajectory
(OSCARSSR.cc:562-874) and the flux integral (OSCARSSR.cc:2054-2092) or the power
density (OSCARSSR.cc:1477-1598), in FP64 on the GPU for a batch of error realisations
at once.  The loss of realisation r is L_r = sum_j (F_rj - target_j)^2 over the
observation points (F = flux or power density).  The backward pass is the
hand-written adjoint of include/srs_hip.h (srs_hip_pole_backward): the integral's
adjoint kernel and the RK4's, which recomputes each checkpoint segment.

    prob = PoleProblem(undulators, poled=0, particle=row, obs_xyz=xyz, energy_ev=e,
                       charge=q, current=i)
    prob.target = prob.flux(np.zeros((1, 2 * N)))[0]          # ideal-undulator map
    deltas = torch.tensor(d, dtype=torch.float64, requires_grad=True)
    loss = pole_loss(deltas, prob)                              # [nreal]
    loss.sum().backward()                                       # deltas.grad

    # power density on a screen with normals (energy unused):
    prob = PoleProblem(undulators, 0, row, xyz, None, q, i, quantity="power_density", normals=nrm)
"""
import numpy as np
import torch

from . import abi


class PoleProblem:
    """Field, particle, screen and energy of a pole-error gradient problem.

    undulators: list of undulator dicts (field, period, nperiods[, phase, taper, center]);
    poled: index of the undulator whose poles carry the errors; particle: dict with
    x0, beta0, qmg, t0, ct_start, ct_stop, npoints (the srs_particle fields);
    quantity: "flux" (at energy_ev) or "power_density" (normals [ns, 3], directional
    as calculate_power_density's normal handling); precision: "fp64", or "mixed" for the
    flux (mixed flux terms forward, mixed adjoint: FP64 phase, FP32 amplitudes)."""

    def __init__(self, undulators, poled, particle, obs_xyz, energy_ev, charge, current, target=None,
                 checkpoint_every=1000, quantity="flux", normals=None, directional=False, precision="fp64"):
        self.undulators = [dict(u) for u in undulators]
        for u in self.undulators:
            u.pop("pole_delta", None)
        self.poled = int(poled)
        self.npoles = 2 * int(self.undulators[self.poled]["nperiods"])
        self.particle = dict(particle)
        self.obs_xyz = np.ascontiguousarray(obs_xyz, dtype=np.float64).reshape(-1, 3)
        if quantity not in ("flux", "power_density"):
            raise ValueError("quantity must be 'flux' or 'power_density'")
        self.quantity = quantity
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if precision == "mixed" and quantity != "flux":
            raise ValueError("the mixed-precision gradient is for the flux loss")
        self.precision = precision
        self.energy_ev = float(energy_ev) if energy_ev is not None else None
        if quantity == "flux" and self.energy_ev is None:
            raise ValueError("the flux loss needs energy_ev")
        if quantity == "power_density":
            if normals is None:
                raise ValueError("the power-density loss needs normals")
            self.normals = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
            if len(self.normals) != len(self.obs_xyz):
                raise ValueError("normals need one row per observation point")
        self.directional = bool(directional)
        self.charge = float(charge)
        self.current = float(current)
        self.checkpoint_every = int(checkpoint_every)
        self.target = None if target is None else np.ascontiguousarray(target, dtype=np.float64)

    def _particle(self):
        return abi.particles([self.particle])[0]

    def forward(self, deltas, want_flux=False):
        """(loss [nreal], map [nreal, ns] or None, tape) for deltas [nreal, 2N]; the map is
        the flux or the power density (want_flux asks for it in either case)."""
        if self.target is None:
            raise ValueError("PoleProblem.target is not set")
        d = np.ascontiguousarray(np.atleast_2d(deltas), dtype=np.float64)
        if d.shape[1] != self.npoles:
            raise ValueError("deltas need %d pole values per realisation" % self.npoles)
        if self.quantity == "power_density":
            return abi.pole_pd_forward(self.undulators, self.poled, self._particle(), d, self.charge, self.current,
                                       self.obs_xyz, self.normals, self.target, self.directional,
                                       self.checkpoint_every, want_flux)
        return abi.pole_forward(self.undulators, self.poled, self._particle(), d, self.charge, self.current,
                                self.obs_xyz, self.energy_ev, self.target, self.checkpoint_every, want_flux,
                                abi.MIXED if self.precision == "mixed" else abi.FP64)

    def flux(self, deltas):
        """Maps [nreal, ns] of the realisations (flux, or power density for that quantity)."""
        saved = self.target
        if saved is None:
            self.target = np.zeros(len(self.obs_xyz))
        try:
            return self.forward(deltas, want_flux=True)[1]
        finally:
            self.target = saved

    def backward(self, tape, nreal, dloss=None):
        return abi.pole_backward(tape, nreal, self.npoles, dloss)


class PoleFluxLoss(torch.autograd.Function):
    """deltas [nreal, 2N] (float64) -> L [nreal]; backward through the adjoint kernels."""

    @staticmethod
    def forward(ctx, deltas, problem):
        d = deltas.detach().to("cpu", torch.float64).numpy()
        loss, _, tape = problem.forward(d)
        ctx.problem = problem
        ctx.tape = tape
        ctx.nreal = d.shape[0]
        ctx.device = deltas.device
        return torch.from_numpy(loss).to(deltas.device)

    @staticmethod
    def backward(ctx, grad_loss):
        g = ctx.problem.backward(ctx.tape, ctx.nreal, grad_loss.detach().to("cpu", torch.float64).numpy())
        return torch.from_numpy(g).to(ctx.device), None


def pole_loss(deltas, problem):
    """Differentiable L_r(delta_r) for a [nreal, 2N] float64 tensor."""
    return PoleFluxLoss.apply(deltas, problem)
