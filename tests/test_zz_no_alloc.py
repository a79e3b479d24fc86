"""GPU test of srs_hip_flux_device_noalloc, the entry point that allocates nothing (its
scratch is the library's static device array): through the C ABI it must give what
srs_hip_flux_device gives on the same inputs -- bit for bit, both precisions -- on
three calls in a row, each with fewer observation points than the one before (the
201 x 201 map at one energy, then more energies, then a shorter trajectory), and
refuse a call whose scratch would not fit, with a clear error."""
import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu


def _case(oracle, npoints, grid, width, energies):
    F = oracle.FieldSet([oracle.undulator((0, 1, 0), (0, 0, 0.05), 41)])
    p = oracle.ideal_electron(3, (0, 0, -1.5), (0, 0, 1), 0.5)
    t, dt = oracle.trajectory(F, p, 0, 3, npoints)
    xs = np.linspace(-width / 2, width / 2, grid)
    xyz = np.array([[x, y, 30.0] for x in xs for y in xs])
    return p, t, dt, xyz, energies


def test_noalloc_flux_equals_flux_device_twice(oracle):
    import torch
    from srsoriginal_amd import abi
    assert abi.device_count() >= 1
    dev = "cuda:0"
    st = torch.cuda.Stream()
    cases = [_case(oracle, 20001, 201, 0.02, [143.52]),                            # the 201 x 201 flux map
             _case(oracle, 20001, 101, 0.02, [143.5, 287.0, 431.0, 574.5, 718.0]),  # two energy launches
             _case(oracle, 4001, 37, 0.01, [143.5])]                                  # then a smaller one
    for prec in (abi.FP64, abi.MIXED):
        for k, (p, t, dt, xyz, es) in enumerate(cases):
            say("noalloc flux call %d, %s: %d points x %d energies, %d trajectory points" %
                (k + 1, "fp64" if prec == abi.FP64 else "mixed", len(xyz), len(es), t.shape[1]))
            d_t = torch.from_numpy(np.ascontiguousarray(t[:6])).to(dev)
            d_obs = torch.from_numpy(np.ascontiguousarray(xyz.T)).to(dev)
            n_out = len(es) * len(xyz)
            ref = torch.zeros(n_out, dtype=torch.float64, device=dev)
            got = torch.full((n_out + 1024,), 7.5, dtype=torch.float64, device=dev)
            ptrs = [d_t[c].data_ptr() for c in range(6)] + [None] * 3
            torch.cuda.synchronize()
            abi.flux_device(ptrs, t.shape[1], dt, d_obs.data_ptr(), len(xyz), es, p.q, p.current, prec,
                            ref.data_ptr(), st.cuda_stream)
            abi.flux_device_noalloc(ptrs, t.shape[1], dt, d_obs.data_ptr(), len(xyz), es, p.q, p.current, prec,
                                    got.data_ptr(), st.cuda_stream)
            st.synchronize()
            g = got.cpu().numpy()
            assert np.all(g[n_out:] == 7.5), "wrote past its output"
            r = ref.cpu().numpy()
            say("  max |noalloc - flux_device| = %.3e (max flux %.4e)" % (np.max(np.abs(g[:n_out] - r)), np.max(r)))
            assert np.array_equal(g[:n_out], r)


def test_noalloc_flux_rejects_scratch_beyond_its_maximum(oracle):
    import torch
    from srsoriginal_amd import abi
    p, t, dt, xyz, es = _case(oracle, 20001, 3, 0.01, [143.5])
    ns = 400000  # ~20 chunks x 48 B x 4e5 points: well beyond 128 MiB; rejected before any launch
    d_t = torch.from_numpy(np.ascontiguousarray(t[:6])).to("cuda:0")
    d_obs = torch.zeros(3 * ns, dtype=torch.float64, device="cuda:0")
    d_obs[2 * ns:] = 30.0
    out = torch.zeros(ns, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    ptrs = [d_t[c].data_ptr() for c in range(6)] + [None] * 3
    with pytest.raises(abi.SrsError, match="bytes of scratch"):
        abi.flux_device_noalloc(ptrs, t.shape[1], dt, d_obs.data_ptr(), ns, es, p.q, p.current, abi.FP64,
                                out.data_ptr(), None)
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0  # nothing ran
