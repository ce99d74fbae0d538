"""pdx_get_u_device / pdx_set_u_device: the displacement between the context
and caller-order device memory, against the host path (pdx_get / pdx_set) on
a batch whose cases end in padding particles: the same values bit for bit,
and a state set on the device solves like the same state set from the host."""
import numpy as np
import pytest

from helpers import rel_l2
from simul_amd import workloads as W
from simul_amd.body import from_specimen

pytestmark = pytest.mark.gpu


def _batch():
    specs = [W.random_vertebra(77, dx=1.6), W.random_vertebra(78, dx=1.6)]
    packed, off = W.pack(specs)
    return packed, from_specimen(packed, case_offsets=off)


def test_device_displacement_matches_host_path(gpu):
    import torch
    packed, body = _batch()
    rng = np.random.default_rng(5)
    u = rng.standard_normal((packed.n, 3)) * 1e-3
    body.u = u
    d = body.get_u_device()
    assert d.shape == (packed.n, 3) and d.dtype == torch.float64 and d.is_cuda
    assert np.array_equal(d.cpu().numpy(), u)
    v = torch.from_numpy(rng.standard_normal((packed.n, 3)) * 2e-3).cuda()
    body.set_u_device(v)
    assert np.array_equal(body.u, v.cpu().numpy())
    with pytest.raises(ValueError):
        body.set_u_device(v.float())
    with pytest.raises(ValueError):
        body.set_u_device(v[:-1])
    # host memory given to the device entry points is refused before any launch
    import ctypes as C
    h = np.zeros((packed.n, 3))
    for fn in ("pdx_set_u_device", "pdx_get_u_device"):
        with pytest.raises(Exception, match="not device memory"):
            body.ctx.call(fn, C.c_void_p(h.ctypes.data))
    assert np.array_equal(body.u, v.cpu().numpy())


def test_device_state_solves_like_host_state(gpu):
    import torch
    packed, body = _batch()
    cold = body.cold_residual()
    u0 = np.zeros((packed.n, 3))
    body.u = u0
    r1 = body.solve("adaptive", tol=1e-6, refs=cold)
    u1 = body.u
    body.set_u_device(torch.zeros((packed.n, 3), dtype=torch.float64, device="cuda"))
    r2 = body.solve("adaptive", tol=1e-6, refs=cold)
    assert [x.iterations for x in r1] == [x.iterations for x in r2]
    assert np.array_equal(body.u, u1)
    assert rel_l2(body.get_u_device().cpu().numpy(), u1) == 0.0
