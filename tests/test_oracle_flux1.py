"""CalculateFlux1 (OSCARSSR.cc:2209-2346), the Lienard-Wiechert form of the flux with
the stored A/c: the oracle's restatement against the reference engine's own output
and against the integration-by-parts identity it must converge to."""
import hashlib

import numpy as np
import pytest

import cases
from conftest import golden

C = 299792458.0
CASES = ["flux1_c1_small", "flux1_epu_offresonance"]


def _rect(oracle, name):
    r = cases.rect(cases.args(name))
    return oracle.rect_surface(r["plane"], r["nx1"], r["nx2"], r["w1"], r["w2"], r["rot"], r["trans"], r["normal"])


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.mark.parametrize("name", CASES)
def test_flux1_oracle_matches_reference_bitwise(oracle, name):
    """The reference computed its own trajectory inside CalculateFlux1; the oracle's
    trajectory is that one bit for bit (the first A/c reads 0 there: the harness zeroes
    the stack under the call, where the reference leaves dxdt uninitialised), and the
    oracle's flux on it equals the reference's container values bit for bit."""
    F, p, t, dt = cases.oracle_setup(oracle, name)
    g = golden(name)
    assert np.array_equal(g["a_first"], np.zeros(3))
    assert dt == float(g["dt"]) and _digest(t) == str(g["digest_traj"])
    xyz, nrm, x12 = _rect(oracle, name)
    c = g["container"]
    assert np.array_equal(c[:, :2], x12)  # 2-D container points (CalculateFlux1's Dimension)
    f = oracle.flux1(t, dt, xyz, float(g["energy_ev"]), p.q, p.current)
    assert np.array_equal(f, c[:, 3])


def _identity_flux(t, dt, obs, energy_ev, q, current):
    """The same flux from the identity that integration by parts gives for the
    reference's integrand (its acceleration term carries no 1/c, so the physical
    E(omega) is scaled by c there and not in the velocity term):
      S = int [c n/D^2 + i W (n - b)/D] e^{-i phi} dt + [(n - b)/(D (1 - n.b)) e^{-i phi}]_ends
          - (c - 1) sum (1 - b^2)(n - b)/(D^2 (1 - n.b)^2) e^{-i phi} dt,
    phi = W (t + D/c), W = 2 pi E / 4.1357e-15; trapezoid rule for the integral."""
    n = t.shape[1]
    X, B = t[0:3], t[3:6]
    W = energy_ev * 2 * np.pi / 4.1357e-15
    c0 = q / (4 * np.pi * C * 8.854187817e-12 * np.sqrt(2 * np.pi))
    c2 = 4 * np.pi * current / (6.62606876e-34 * abs(q) * 1.2566370614e-6 * C) * 1e-6 * 1e-3
    tt = dt * np.arange(n)
    out = []
    for o in obs:
        R = o[:, None] - X
        D = np.sqrt((R * R).sum(0))
        N = R / D
        e = np.exp(-1j * W * (tt + D / C))
        omb = 1 - (N * B).sum(0)
        f = (C * N / D**2 + 1j * W * (N - B) / D) * e
        integral = (f[:, 1:] + f[:, :-1]).sum(1) * dt / 2
        g = (N - B) / (D * omb) * e
        vel = ((1 - (B * B).sum(0)) * (N - B) / (D**2 * omb**2) * e).sum(1) * dt
        S = c0 * (integral + g[:, -1] - g[:, 0] - (C - 1) * vel)
        out.append(c2 * np.sum(np.abs(S)**2))
    return np.array(out)


@pytest.mark.parametrize("energy_ev", [143.52, 300.0])
def test_flux1_converges_to_integration_by_parts_identity(oracle, energy_ev):
    """CalculateFlux1's sum (lagged A/c, rectangle rule) approaches the identity as the
    step shrinks: measured 4.8e-3, 1.2e-3, 2.8e-4 at 20001, 40001, 80001 points at the
    first harmonic (3.0e-3, 9.7e-4, 3.5e-4 at 300 eV).  A wrong cross product, a dropped
    or mis-signed term or a wrong point index breaks the convergence."""
    F = oracle.FieldSet([dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)])
    p = oracle.ideal_electron(3.0, [0, 0, -1.5], [0, 0, 1], 0.5)
    xyz, nrm, x12 = oracle.rect_surface("XY", 5, 3, 0.01, 0.006, (0, 0, 0), (0, 0, 30), 0)
    errs = []
    for npts in (20001, 80001):
        t, dt = oracle.trajectory(F, p, 0.0, 3.0, npts)
        f1 = oracle.flux1(t, dt, xyz, energy_ev, p.q, p.current)
        fi = _identity_flux(t, dt, xyz, energy_ev, p.q, p.current)
        errs.append(float(np.max(np.abs(f1 - fi)) / np.max(fi)))
    assert errs[0] < 6e-3 and errs[1] < 5e-4 and errs[1] < errs[0] / 8, errs
