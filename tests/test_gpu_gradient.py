"""GPU tests of the hand-written adjoint (flux integral + RK4 with checkpoints) for
d(loss)/d(pole errors), through the C ABI and torch.autograd.

References: the oracle's forward-mode (dual number) FP64 gradient, which shares no
derivative formula with the adjoint, and central differences of the GPU forward.
Central differences of this FP64 model carry ~7e-9/h relative noise (phase rounding of
the flux), so they are taken at h = 1e-4 and compared at 1e-3."""
import numpy as np
import pytest

import cases
from conftest import golden
from gpu_util import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grad():
    from srsoriginal_amd import abi, grad as G
    assert abi.device_count() >= 1
    return G


def _fixture_problem(G, oracle, ct_start=0.0, npoints=None):
    name = "poles_flux_and_fd_gradient"
    a = cases.args(name)
    g = golden(name)
    pole = cases.pole(a)
    b = cases.beam(a)
    p = oracle.ideal_electron(b["energy"], b["x0"], b["d0"], b["current"])
    r = cases.rect(a)
    xyz, _, _ = oracle.rect_surface(r["plane"], r["nx1"], r["nx2"], r["w1"], r["w2"], r["rot"], r["trans"], r["normal"])
    c0, c1, n = cases.ct(a)
    n = npoints or n
    row = dict(x0=list(p.x0), beta0=list(p.beta0), qmg=p.qmg, t0=0.0, ct_start=ct_start, ct_stop=c1, npoints=n)
    und = [dict(field=pole["field"], period=pole["period"], nperiods=pole["nperiods"])]
    prob = G.PoleProblem(und, 0, row, xyz, float(a["energy_ev"]), p.q, p.current, target=g["flux_ideal"][:len(xyz)])
    return prob, p, pole, xyz, (ct_start, c1, n), g


def _deltas(g, nreal):
    d = [g["deltas"]] + [np.random.default_rng(100 + k).normal(0, 0.005, len(g["deltas"])) for k in range(nreal - 1)]
    return np.array(d)


def test_pole_gradient_matches_oracle_dual_numbers(grad, oracle):
    say("pole-error gradient: GPU adjoint vs oracle dual numbers (5x5, 2001 points, 3 realisations)")
    prob, p, pole, xyz, (c0, c1, n), g = _fixture_problem(grad, oracle)
    D = _deltas(g, 3)
    loss, flux, tape = prob.forward(D, want_flux=True)
    G = prob.backward(tape, len(D))
    for r in range(len(D)):
        F = oracle.FieldSet([], pole, D[r])
        lo, go = oracle.pole_loss_gradient(F, p, c0, c1, n, xyz, prob.energy_ev, prob.target, nthreads=8)
        el = abs(loss[r] - lo) / lo
        eg = np.max(np.abs(G[r] - go)) / np.max(np.abs(go))
        say("  realisation %d: loss rel %.3e  gradient normwise %.3e" % (r, el, eg))
        assert el <= 1e-9 and eg <= 1e-8
    assert abs(loss[0] - float(g["loss"])) <= 1e-9 * float(g["loss"])


def test_pole_gradient_backward_leg_and_checkpoint_intervals(grad, oracle):
    say("pole-error gradient with a backward RK4 leg (ct from -0.3 m); checkpoints every 1000 / 7 / whole")
    prob, p, pole, xyz, (c0, c1, n), g = _fixture_problem(grad, oracle, ct_start=-0.3, npoints=2201)
    D = _deltas(g, 2)
    grads = []
    for every in (1000, 7, 100000):
        prob.checkpoint_every = every
        loss, _, tape = prob.forward(D)
        grads.append(prob.backward(tape, len(D)))
    # segments are swept independently and chained, so intervals differ only by rounding
    for gk in grads[1:]:
        assert np.max(np.abs(gk - grads[0])) <= 1e-12 * np.max(np.abs(grads[0]))
    for r in range(len(D)):
        F = oracle.FieldSet([], pole, D[r])
        lo, go = oracle.pole_loss_gradient(F, p, c0, c1, n, xyz, prob.energy_ev, prob.target, nthreads=8)
        eg = np.max(np.abs(grads[0][r] - go)) / np.max(np.abs(go))
        say("  realisation %d: gradient normwise %.3e" % (r, eg))
        assert eg <= 1e-8


def test_pole_gradient_central_differences_config4_shape(grad, oracle):
    say("pole-error gradient vs GPU central differences, 21x21 over 5 mm at 30 m, 20001 points, 2 realisations")
    name = "poles_flux_and_fd_gradient"
    a = cases.args(name)
    b = cases.beam(a)
    p = oracle.ideal_electron(b["energy"], b["x0"], b["d0"], b["current"])
    xyz, _, _ = oracle.rect_surface("XY", 21, 21, 0.005, 0.005, (0, 0, 0), (0, 0, 30), 0)
    row = dict(x0=list(p.x0), beta0=list(p.beta0), qmg=p.qmg, t0=0.0, ct_start=0.0, ct_stop=3.0, npoints=20001)
    und = [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)]
    prob = grad.PoleProblem(und, 0, row, xyz, 143.52176088044024, p.q, p.current)
    prob.target = prob.flux(np.zeros((1, 82)))[0]
    rng = np.random.default_rng(7)
    D = rng.normal(0, 0.005, (2, 82))
    loss, _, tape = prob.forward(D)
    G = prob.backward(tape, 2)
    h = 1e-4
    for k in range(3):
        v = rng.normal(0, 1, (2, 82))
        lp = prob.forward(D + h * v)[0]
        lm = prob.forward(D - h * v)[0]
        fd = (lp - lm) / (2 * h)
        ad = np.sum(G * v, axis=1)
        # relative to |grad| |v|: a directional derivative can be small by cancellation
        # while the difference noise keeps the scale of the whole gradient
        err = np.abs(fd - ad) / (np.linalg.norm(G, axis=1) * np.linalg.norm(v, axis=1))
        say("  direction %d: adjoint %s  fd %s  err %s" % (k, ad, fd, err))
        assert np.all(err <= 1e-3)


# The mixed gradient (mixed flux terms forward, mixed adjoint: FP64 phase, FP32
# amplitudes, their derivatives and products) is held to the mixed forward's bound
# against the FP64 result.
TOL_MIXED_GRAD = 1e-5


def test_mixed_pole_gradient_against_oracle_and_fp64(grad, oracle):
    say("mixed-precision pole-error gradient (mixed flux forward, mixed adjoint) vs oracle dual numbers "
        "and the FP64 GPU gradient (5x5, 2001 points, 3 realisations)")
    prob, p, pole, xyz, (c0, c1, n), g = _fixture_problem(grad, oracle)
    D = _deltas(g, 3)
    l64, f64, tape = prob.forward(D, want_flux=True)
    G64 = prob.backward(tape, len(D))
    prob.precision = "mixed"
    lmx, fmx, tape = prob.forward(D, want_flux=True)
    Gmx = prob.backward(tape, len(D))
    ef = np.max(np.abs(fmx - f64)) / np.max(np.abs(f64))
    say("  flux mixed vs FP64 %.3e" % ef)
    assert ef <= 1e-5
    for r in range(len(D)):
        F = oracle.FieldSet([], pole, D[r])
        lo, go = oracle.pole_loss_gradient(F, p, c0, c1, n, xyz, prob.energy_ev, prob.target, nthreads=8)
        el = abs(lmx[r] - lo) / lo
        eg = np.max(np.abs(Gmx[r] - go)) / np.max(np.abs(go))
        e64 = np.max(np.abs(Gmx[r] - G64[r])) / np.max(np.abs(G64[r]))
        say("  realisation %d: loss rel %.3e  gradient normwise vs oracle %.3e, vs FP64 GPU %.3e" % (r, el, eg, e64))
        assert el <= TOL_MIXED_GRAD and eg <= TOL_MIXED_GRAD and e64 <= TOL_MIXED_GRAD


def test_mixed_pole_gradient_config4_shape(grad, oracle):
    say("mixed-precision pole-error gradient vs the FP64 one, 21x21 over 5 mm at 30 m, 20001 points, 2 realisations")
    name = "poles_flux_and_fd_gradient"
    a = cases.args(name)
    b = cases.beam(a)
    p = oracle.ideal_electron(b["energy"], b["x0"], b["d0"], b["current"])
    xyz, _, _ = oracle.rect_surface("XY", 21, 21, 0.005, 0.005, (0, 0, 0), (0, 0, 30), 0)
    row = dict(x0=list(p.x0), beta0=list(p.beta0), qmg=p.qmg, t0=0.0, ct_start=0.0, ct_stop=3.0, npoints=20001)
    und = [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)]
    prob = grad.PoleProblem(und, 0, row, xyz, 143.52176088044024, p.q, p.current)
    prob.target = prob.flux(np.zeros((1, 82)))[0]
    D = np.random.default_rng(7).normal(0, 0.005, (2, 82))
    l64, _, tape = prob.forward(D)
    G64 = prob.backward(tape, 2)
    prob.precision = "mixed"
    lmx, _, tape = prob.forward(D)
    Gmx = prob.backward(tape, 2)
    for r in range(2):
        el = abs(lmx[r] - l64[r]) / l64[r]
        eg = np.max(np.abs(Gmx[r] - G64[r])) / np.max(np.abs(G64[r]))
        cs = np.dot(Gmx[r], G64[r]) / (np.linalg.norm(Gmx[r]) * np.linalg.norm(G64[r]))
        say("  realisation %d: loss rel %.3e  gradient normwise %.3e  cosine %.12f" % (r, el, eg, cs))
        assert el <= TOL_MIXED_GRAD and eg <= TOL_MIXED_GRAD


def test_mixed_pole_gradient_tilted_screen_general_phase(grad, oracle):
    """On a screen that is not one plane z = Z the mixed adjoint takes the general phase
    (root, quotient and reduction per term in FP64) instead of the paraxial one."""
    say("mixed pole-error gradient on a tilted screen (general phase) vs the FP64 one")
    p = oracle.ideal_electron(3.0, (0.0, 0.0, -1.5), (0.0, 0.0, 1.0), 0.5)
    xyz, _, _ = oracle.rect_surface("XY", 9, 9, 0.004, 0.004, (0.3, 0.2, 0), (0, 0, 30), 0)
    assert np.ptp(xyz[:, 2]) > 1e-4
    row = dict(x0=list(p.x0), beta0=list(p.beta0), qmg=p.qmg, t0=0.0, ct_start=0.0, ct_stop=3.0, npoints=4001)
    und = [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)]
    prob = grad.PoleProblem(und, 0, row, xyz, 143.52176088044024, p.q, p.current)
    prob.target = prob.flux(np.zeros((1, 82)))[0]
    D = np.random.default_rng(9).normal(0, 0.005, (2, 82))
    l64, _, tape = prob.forward(D)
    G64 = prob.backward(tape, 2)
    prob.precision = "mixed"
    lmx, _, tape = prob.forward(D)
    Gmx = prob.backward(tape, 2)
    for r in range(2):
        el = abs(lmx[r] - l64[r]) / l64[r]
        eg = np.max(np.abs(Gmx[r] - G64[r])) / np.max(np.abs(G64[r]))
        say("  realisation %d: loss rel %.3e  gradient normwise %.3e" % (r, el, eg))
        assert el <= TOL_MIXED_GRAD and eg <= TOL_MIXED_GRAD


def test_mixed_pole_gradient_screen_across_the_paraxial_bound(grad, oracle):
    """A 42 mm screen at 30 m: adjoint blocks of trajectory points far from the screen
    pass the paraxial bounds for the screen's whole extent and take the paraxial phase,
    near ones fail them and take the general phase; against the FP64 gradient."""
    say("mixed pole-error gradient, 42 mm screen at 30 m (paraxial and general adjoint blocks) vs FP64")
    p = oracle.ideal_electron(3.0, (0.0, 0.0, -1.5), (0.0, 0.0, 1.0), 0.5)
    xyz, _, _ = oracle.rect_surface("XY", 9, 9, 0.042, 0.042, (0, 0, 0), (0, 0, 30), 0)
    row = dict(x0=list(p.x0), beta0=list(p.beta0), qmg=p.qmg, t0=0.0, ct_start=0.0, ct_stop=3.0, npoints=4001)
    und = [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)]
    prob = grad.PoleProblem(und, 0, row, xyz, 143.52176088044024, p.q, p.current)
    prob.target = prob.flux(np.zeros((1, 82)))[0]
    D = np.random.default_rng(13).normal(0, 0.005, (2, 82))
    l64, _, tape = prob.forward(D)
    G64 = prob.backward(tape, 2)
    prob.precision = "mixed"
    lmx, _, tape = prob.forward(D)
    Gmx = prob.backward(tape, 2)
    for r in range(2):
        el = abs(lmx[r] - l64[r]) / l64[r]
        eg = np.max(np.abs(Gmx[r] - G64[r])) / np.max(np.abs(G64[r]))
        say("  realisation %d: loss rel %.3e  gradient normwise %.3e" % (r, el, eg))
        assert el <= TOL_MIXED_GRAD and eg <= TOL_MIXED_GRAD


def test_mixed_pole_gradient_beam_along_x_takes_the_fp64_terms(grad, oracle):
    """A realisation whose beam does not travel along z is flagged by the device RK4; the
    mixed forward then sums its FP64 terms and the mixed adjoint takes its FP64 terms, so
    the mixed gradient equals the FP64 one (undulator and beam along x, screen at x = 30 m)."""
    say("mixed pole-error gradient, beam along x: the FP64 terms on both passes")
    p = oracle.ideal_electron(3.0, (-1.5, 0.0, 0.0), (1.0, 0.0, 0.0), 0.5)
    xyz, _, _ = oracle.rect_surface("YZ", 7, 7, 0.004, 0.004, (0, 0, 0), (30, 0, 0), 0)
    row = dict(x0=list(p.x0), beta0=list(p.beta0), qmg=p.qmg, t0=0.0, ct_start=0.0, ct_stop=3.0, npoints=2001)
    und = [dict(field=(0, 1, 0), period=(0.05, 0, 0), nperiods=41)]
    prob = grad.PoleProblem(und, 0, row, xyz, 143.52176088044024, p.q, p.current)
    prob.target = prob.flux(np.zeros((1, 82)))[0]
    D = np.random.default_rng(5).normal(0, 0.005, (2, 82))
    l64, f64, tape = prob.forward(D, want_flux=True)
    G64 = prob.backward(tape, 2)
    prob.precision = "mixed"
    lmx, fmx, tape = prob.forward(D, want_flux=True)
    Gmx = prob.backward(tape, 2)
    ef = np.max(np.abs(fmx - f64)) / np.max(np.abs(f64))
    eg = np.max(np.abs(Gmx - G64)) / np.max(np.abs(G64))
    say("  flux %.3e  loss %.3e  gradient %.3e" % (ef, np.max(np.abs(lmx - l64) / l64), eg))
    assert np.max(np.abs(G64)) > 0
    assert ef <= 1e-13 and eg <= 1e-12


def test_autograd_function_and_tape(grad, oracle):
    import torch
    from srsoriginal_amd import abi
    say("torch.autograd.Function over the adjoint; tape replaced by a later forward")
    prob, p, pole, xyz, (c0, c1, n), g = _fixture_problem(grad, oracle)
    D = _deltas(g, 3)
    d = torch.tensor(D, dtype=torch.float64, requires_grad=True)
    w = torch.tensor([0.5, -2.0, 1.5], dtype=torch.float64)
    loss = grad.pole_loss(d, prob)
    (loss * w).sum().backward()
    ref_loss, _, tape = prob.forward(D)
    ref = prob.backward(tape, 3, w.numpy())
    assert np.array_equal(loss.detach().numpy(), ref_loss)
    assert np.array_equal(d.grad.numpy(), ref)
    stale = tape
    prob.forward(D[:1])
    with pytest.raises(abi.SrsError, match="tape"):
        prob.backward(stale, 3)


def _pd_problem(G, oracle, ct_start=0.0, npoints=None, screen=None, directional=False):
    name = "poles_pd_and_fd_gradient"
    a = cases.args(name)
    g = golden(name)
    pole = cases.pole(a)
    b = cases.beam(a)
    p = oracle.ideal_electron(b["energy"], b["x0"], b["d0"], b["current"])
    r = cases.rect(a)
    if screen is None:
        xyz, nrm, _ = oracle.rect_surface(r["plane"], r["nx1"], r["nx2"], r["w1"], r["w2"], r["rot"], r["trans"],
                                          r["normal"])
        target = g["pd_ideal"]
    else:
        xyz, nrm, _ = oracle.rect_surface(*screen)
        target = np.zeros(len(xyz))
    c0, c1, n = cases.ct(a)
    n = npoints or n
    row = dict(x0=list(p.x0), beta0=list(p.beta0), qmg=p.qmg, t0=0.0, ct_start=ct_start, ct_stop=c1, npoints=n)
    und = [dict(field=pole["field"], period=pole["period"], nperiods=pole["nperiods"])]
    prob = G.PoleProblem(und, 0, row, xyz, None, p.q, p.current, target=target, quantity="power_density",
                         normals=nrm, directional=directional)
    if screen is not None:
        prob.target = 0.9 * prob.flux(np.zeros((1, 82)))[0]
    return prob, p, pole, xyz, nrm, (ct_start, c1, n), g


# The power-density gradient is conditioned by 1 - beta.N ~ 1.5e-8 (it enters as
# 1/b^6): the oracle's own dual gradient moves by 1e-8 normwise when Num is written in
# the algebraically equal form u (N.A) - b A instead of N x (u x A) (2e-9 at 0.3 GeV),
# so two correct FP64 gradients agree to a few 1e-8, not to the flux's 1e-13.
PD_GRAD_TOL = 5e-8


@pytest.mark.parametrize("npoints", [2001, 20001])
def test_pd_pole_gradient_matches_oracle_dual_numbers(grad, oracle, npoints):
    # 20001 points run the forward in RK4 time segments overlapped with the power density
    say("power-density pole-error gradient: GPU adjoint vs oracle dual numbers (5x5 over 20 mm, %d points)" % npoints)
    prob, p, pole, xyz, nrm, (c0, c1, n), g = _pd_problem(grad, oracle, npoints=npoints)
    if npoints != 2001:
        prob.target = prob.flux(np.zeros((1, 82)))[0]
    D = np.array([g["deltas"]] + [np.random.default_rng(200 + k).normal(0, 0.005, 82) for k in range(2)])
    loss, pd, tape = prob.forward(D, want_flux=True)
    G = prob.backward(tape, len(D))
    for r in range(len(D)):
        F = oracle.FieldSet([], pole, D[r])
        lo, go, po = oracle.pole_pd_loss_gradient(F, p, c0, c1, n, xyz, nrm, prob.target, nthreads=8)
        ep = np.max(np.abs(pd[r] - po)) / np.max(np.abs(po))
        el = abs(loss[r] - lo) / lo
        eg = np.max(np.abs(G[r] - go)) / np.max(np.abs(go))
        say("  realisation %d: map %.3e  loss rel %.3e  gradient normwise %.3e" % (r, ep, el, eg))
        assert ep <= 1e-10 and el <= 1e-9 and eg <= PD_GRAD_TOL
    if npoints == 2001:
        assert np.max(np.abs(pd[0] - g["pd_scaled"])) <= 1e-10 * np.max(g["pd_scaled"])


def test_pd_pole_gradient_directional_backward_leg_and_checkpoints(grad, oracle):
    say("power-density gradient, directional tilted screen inside the undulator, ct from -0.3 m; "
        "checkpoints every 1000 / 7 / whole")
    screen = ("XY", 5, 4, 0.01, 0.008, (0.3, 0.1, 0.0), (0.001, 0.0, 0.4), 1)
    prob, p, pole, xyz, nrm, (c0, c1, n), g = _pd_problem(grad, oracle, ct_start=-0.3, npoints=2201, screen=screen,
                                                          directional=True)
    D = np.array([g["deltas"], np.random.default_rng(300).normal(0, 0.005, 82)])
    grads = []
    for every in (1000, 7, 100000):
        prob.checkpoint_every = every
        loss, _, tape = prob.forward(D)
        grads.append(prob.backward(tape, len(D)))
    for gk in grads[1:]:
        assert np.max(np.abs(gk - grads[0])) <= 1e-12 * np.max(np.abs(grads[0]))
    for r in range(len(D)):
        F = oracle.FieldSet([], pole, D[r])
        lo, go, _ = oracle.pole_pd_loss_gradient(F, p, c0, c1, n, xyz, nrm, prob.target, directional=True, nthreads=8)
        eg = np.max(np.abs(grads[0][r] - go)) / np.max(np.abs(go))
        say("  realisation %d: loss rel %.3e  gradient normwise %.3e" % (r, abs(loss[r] - lo) / lo, eg))
        assert abs(loss[r] - lo) <= 1e-9 * lo and eg <= PD_GRAD_TOL


def test_pd_pole_gradient_central_differences(grad, oracle):
    say("power-density gradient vs GPU central differences, 21x21 over 20 mm at 30 m, 20001 points, 2 realisations")
    b = dict(energy=3.0, x0=(0, 0, -1.5), d0=(0, 0, 1), current=0.5)
    p = oracle.ideal_electron(b["energy"], b["x0"], b["d0"], b["current"])
    xyz, nrm, _ = oracle.rect_surface("XY", 21, 21, 0.02, 0.02, (0, 0, 0), (0, 0, 30), 0)
    row = dict(x0=list(p.x0), beta0=list(p.beta0), qmg=p.qmg, t0=0.0, ct_start=0.0, ct_stop=3.0, npoints=20001)
    und = [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)]
    prob = grad.PoleProblem(und, 0, row, xyz, None, p.q, p.current, quantity="power_density", normals=nrm,
                            target=np.zeros(len(xyz)))
    prob.target = prob.flux(np.zeros((1, 82)))[0]
    rng = np.random.default_rng(11)
    D = rng.normal(0, 0.005, (2, 82))
    loss, _, tape = prob.forward(D)
    G = prob.backward(tape, 2)
    h = 1e-4
    for k in range(3):
        v = rng.normal(0, 1, (2, 82))
        fd = (prob.forward(D + h * v)[0] - prob.forward(D - h * v)[0]) / (2 * h)
        ad = np.sum(G * v, axis=1)
        err = np.abs(fd - ad) / (np.linalg.norm(G, axis=1) * np.linalg.norm(v, axis=1))
        say("  direction %d: adjoint %s  fd %s  err %s" % (k, ad, fd, err))
        assert np.all(err <= 1e-3)


def test_pd_autograd_function(grad, oracle):
    import torch
    say("torch.autograd.Function over the power-density adjoint")
    prob, p, pole, xyz, nrm, _, g = _pd_problem(grad, oracle)
    D = np.array([g["deltas"], -g["deltas"]])
    d = torch.tensor(D, dtype=torch.float64, requires_grad=True)
    loss = grad.pole_loss(d, prob)
    (loss * torch.tensor([1.0, 3.0], dtype=torch.float64)).sum().backward()
    ref_loss, _, tape = prob.forward(D)
    ref = prob.backward(tape, 2, np.array([1.0, 3.0]))
    assert np.array_equal(loss.detach().numpy(), ref_loss)
    assert np.array_equal(d.grad.numpy(), ref)
