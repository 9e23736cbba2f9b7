"""Device-resident interop with an OpenMM GPU platform through the C ABI (needs an MI355X):
positions gathered from a posq buffer (float4 or double4 per slot) and forces added to an
int64 fixed-point force buffer (axis-major, padded), both in a permuted atom order, as the
CUDA platform's kernels read posq and add into its force buffer (reduceIntegralForce.cu:61-77).
Checked bitwise against the host entry points on the same positions.
"""
import ctypes

import numpy as np
import pytest
import torch

from openmmcustomgbintegralforce_amd import _native as N
from openmmcustomgbintegralforce_amd import synthetic as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("posq_double", [1, 0])
def test_posq_in_and_fixed_point_forces_out(posq_double):
    s = S.globule(700, 81)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p.n_lebedev = 50
    n = s.n
    padded = (n + 31) // 32 * 32
    slot = np.random.default_rng(82).permutation(n).astype(np.int32)
    dt = torch.float64 if posq_double else torch.float32
    posq = torch.zeros((padded, 4), dtype=dt)
    posq[torch.from_numpy(slot.astype(np.int64)), :3] = torch.from_numpy(s.pos).to(dt)
    posq[torch.from_numpy(slot.astype(np.int64)), 3] = torch.from_numpy(s.charge).to(dt)
    pos_used = posq[torch.from_numpy(slot.astype(np.int64)), :3].to(torch.float64).numpy().copy()
    d_posq = posq.cuda()
    d_slot = torch.from_numpy(slot).cuda()
    d_force = torch.full((3 * padded,), 7, dtype=torch.int64, device="cuda")  # other forces already there
    torch.cuda.synchronize()
    print(f"[interop] posq {'double4' if posq_double else 'float4'} N={n}", flush=True)

    eng = N.Engine(p, s.charge, s.radius)
    lib = eng._lib
    N._check(lib.gbi_set_positions_posq(eng._h, ctypes.c_void_p(d_posq.data_ptr()), posq_double,
                                        ctypes.c_void_p(d_slot.data_ptr())))
    eng.compute()
    eng.synchronize()
    N._check(lib.gbi_add_forces_fixed(eng._h, ctypes.c_void_p(d_force.data_ptr()), padded,
                                      ctypes.c_void_p(d_slot.data_ptr())))
    eng.synchronize()
    e_dev = eng.energy()
    fx = d_force.cpu().numpy().reshape(3, padded)
    eng.close()

    ref = N.Engine(p, s.charge, s.radius)
    e_host, f_host = ref.execute(pos_used)
    ref.close()
    f_dev = np.stack([(fx[d, slot] - 7) / 2.0 ** 32 for d in range(3)], axis=1)
    print(f"[interop] energy {e_dev:.6f} vs host {e_host:.6f}; forces bitwise {np.array_equal(f_dev, f_host)}",
          flush=True)
    assert e_dev == e_host
    assert np.array_equal(f_dev, f_host)
    untouched = np.setdiff1d(np.arange(padded), slot)
    assert np.all(fx[:, untouched] == 7)
