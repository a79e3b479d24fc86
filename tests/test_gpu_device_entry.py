"""GPU tests of the device-resident entry points (srs_hip_flux_device,
srs_hip_power_density_device, srs_hip_spectrum_device) through the C ABI.

These are the entry points the benchmark times.  Each is called with the
trajectory as 6 device arrays (no accelerations) and as 9 (with A/c), on a
user stream that is not torch's, in both precisions, and compared with the
host-staged entry point on the same inputs: bit for bit in FP64 (same kernels,
same chunking), and bit for bit in mixed precision too (same staging kernel).
A canary region after each output buffer must stay untouched, which catches
a kernel that writes past the end of what the caller handed over."""
import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu

CANARY = 4096  # doubles of sentinel after each device output


@pytest.fixture(scope="module")
def env(oracle):
    import torch
    from srsoriginal_amd import abi
    assert abi.device_count() >= 1
    F = oracle.FieldSet([oracle.undulator((0, 1, 0), (0, 0, 0.05), 41)])
    p = oracle.ideal_electron(3, (0, 0, -1.5), (0, 0, 1), 0.5)
    t, dt = oracle.trajectory(F, p, 0, 3, 4001)
    st = torch.cuda.Stream()  # a stream of its own (default priority), not torch's current one
    yield dict(torch=torch, abi=abi, t=t, dt=dt, p=p, st=st, stream=st.cuda_stream)
    torch.cuda.synchronize()


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")


def _traj_ptrs(torch, t, rows):
    d = _dev(torch, t[:rows])
    return d, [d[k].data_ptr() if k < rows else None for k in range(9)]


def _out(torch, n):
    o = torch.zeros(n + CANARY, dtype=torch.float64, device="cuda:0")
    o[n:] = 12345.678
    torch.cuda.synchronize()  # inputs and canary were written on torch's stream; the call runs on another
    return o


def _finish(e, o, n):
    e["st"].synchronize()
    e["torch"].cuda.synchronize()
    h = o.cpu().numpy()
    assert np.all(h[n:] == 12345.678), "device entry point wrote past its output"
    return h[:n]


def _grid(nx, ny, w, z):
    xs = np.linspace(-w / 2, w / 2, nx)
    ys = np.linspace(-w / 2, w / 2, ny)
    return np.array([[x, y, z] for x in xs for y in ys])


@pytest.mark.parametrize("rows", [6, 9])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_flux_device_equals_host_entry(env, rows, prec):
    torch, abi, t, dt, p = env["torch"], env["abi"], env["t"], env["dt"], env["p"]
    P = abi.FP64 if prec == "fp64" else abi.MIXED
    xyz = _grid(23, 21, 0.01, 30.0)
    es = [143.5, 287.0]
    say("flux_device %s, %d-array trajectory, %d points x %d energies, user stream" % (prec, rows, len(xyz), len(es)))
    host = abi.flux(t[:6], dt, xyz, es, p.q, p.current, P)
    d_t, ptrs = _traj_ptrs(torch, t, rows)
    d_obs = _dev(torch, xyz.T)
    out = _out(torch, len(es) * len(xyz))
    abi.flux_device(ptrs, t.shape[1], dt, d_obs.data_ptr(), len(xyz), es, p.q, p.current, P, out.data_ptr(),
                    env["stream"])
    got = _finish(env, out, len(es) * len(xyz)).reshape(len(es), len(xyz))
    assert np.array_equal(got, host)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_power_density_device_equals_host_entry(env, prec):
    torch, abi, t, dt, p = env["torch"], env["abi"], env["t"], env["dt"], env["p"]
    P = abi.FP64 if prec == "fp64" else abi.MIXED
    xyz = _grid(17, 19, 0.03, 30.0)
    nrm = np.tile([0.0, 0.0, 1.0], (len(xyz), 1))
    say("power_density_device %s, 9-array trajectory, %d points, user stream" % (prec, len(xyz)))
    for directional in (False, True):
        host = abi.power_density(t, dt, xyz, nrm, directional, p.q, p.current, P)
        d_t, ptrs = _traj_ptrs(torch, t, 9)
        d_obs, d_nrm = _dev(torch, xyz.T), _dev(torch, nrm.T)
        out = _out(torch, len(xyz))
        abi.power_density_device(ptrs, t.shape[1], dt, d_obs.data_ptr(), d_nrm.data_ptr(), len(xyz), directional,
                                 p.q, p.current, P, out.data_ptr(), env["stream"])
        got = _finish(env, out, len(xyz))
        assert np.array_equal(got, host)


def test_power_density_device_needs_accelerations(env):
    """A 6-array trajectory has no A/c: the power density must refuse it, not read NULL."""
    torch, abi, t, dt, p = env["torch"], env["abi"], env["t"], env["dt"], env["p"]
    xyz = _grid(3, 3, 0.01, 30.0)
    nrm = np.tile([0.0, 0.0, 1.0], (len(xyz), 1))
    d_t, ptrs = _traj_ptrs(torch, t, 6)
    d_obs, d_nrm = _dev(torch, xyz.T), _dev(torch, nrm.T)
    out = _out(torch, len(xyz))
    with pytest.raises(abi.SrsError, match="acceleration"):
        abi.power_density_device(ptrs, t.shape[1], dt, d_obs.data_ptr(), d_nrm.data_ptr(), len(xyz), False, p.q,
                                 p.current, abi.FP64, out.data_ptr(), env["stream"])


@pytest.mark.parametrize("rows", [6, 9])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_spectrum_device_equals_host_entry(env, rows, prec):
    torch, abi, t, dt, p = env["torch"], env["abi"], env["t"], env["dt"], env["p"]
    P = abi.FP64 if prec == "fp64" else abi.MIXED
    es = 100.0 + (3000.0 - 100.0) / 399 * (np.arange(400) + 1)
    obs = (0.0004, -0.0002, 30.0)
    say("spectrum_device %s, %d-array trajectory, %d energies, user stream" % (prec, rows, len(es)))
    host = abi.spectrum(t[:6], dt, obs, es, p.q, p.current, P)
    d_t, ptrs = _traj_ptrs(torch, t, rows)
    out = _out(torch, len(es))
    abi.spectrum_device(ptrs, t.shape[1], dt, obs, es, p.q, p.current, P, out.data_ptr(), env["stream"])
    got = _finish(env, out, len(es))
    assert np.array_equal(got, host)
