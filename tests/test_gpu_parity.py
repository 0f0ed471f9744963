"""Parity of the HIP path with the CPU oracle (restatement of pawan's
reference method), through the C ABI.  Tolerances: binary64 rates 1e-10 per
particle relative (floor 1e-3 x RMS); compensated binary32 1e-6 relative RMS."""
import numpy as np
import pytest

import oracle as orc
import pawan_amd as pa
from pawan_amd import generators as gen
from helpers import per_particle_rel, rel_rms

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 1e-6


def _gpu_rates(w, nu, wake_n=None, precision=pa.FP64, kernel=pa.KERNEL_HIGHORDER, uinf=(0, 0, 0),
               block=0, chunk=0):
    wake_n = wake_n or [w["pos"].shape[0]]
    offs = np.cumsum([0] + list(wake_n))
    wakes = [pa.Wake(w["pos"][a:b], w["vort"][a:b], w["radius"][a:b], w["volume"][a:b])
             for a, b in zip(offs[:-1], offs[1:])]
    S = pa.Interaction(*wakes, nu=nu, precision=precision, kernel=kernel, uinf=uinf, block=block,
                       chunk=chunk)
    S.solve()
    r = S.getRates()
    S.close()
    pos, _ = orc.unpack_state(np.zeros(6 * offs[-1]), wake_n)  # shape helper
    vel, dv = orc.unpack_state(r, wake_n)
    return vel, dv


@pytest.mark.parametrize("block", [64, 128, 256])
def test_single_ring_rates_fp64(block):
    w = gen.config_single_ring()
    print(f"case single ring N={w['n']} block={block}", flush=True)
    vel, dv = _gpu_rates(w, w["nu"], block=block)
    rv, rd = orc.rates(w["pos"], w["vort"], w["radius"], w["volume"], w["nu"], mode=orc.OMP_CLEAN)
    ev, ed = per_particle_rel(vel, rv), per_particle_rel(dv, rd)
    print(f"  vel {ev:.3e} dvort {ed:.3e}", flush=True)
    assert ev < TOL64 and ed < TOL64
