"""bh_set_potential_noalloc (the import of a host potential through the
library's fixed staging array, no device allocation) against
bh_set_potential (which stages through a buffer it allocates): the plan's
warm-start state and the warm solve that follows are bitwise equal, on two
calls in a row, the second on a smaller grid."""
import ctypes

import numpy as np
import pytest

from bipoly_hydro_amd import density
from bipoly_hydro_amd import potential as P

pytestmark = pytest.mark.gpu


def warm_state(plan):
    """The plan's warm-start potential in Fourier space (XJ)."""
    L = P.load_library()
    L.bhx_snapshot.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]
    out = np.empty(plan.shape)
    assert L.bhx_snapshot(plan.handle, 2, out.ctypes.data, out.size) == 0
    return out


@pytest.mark.parametrize("shape,precision", [((258, 130, 512), "fp64"), ((66, 66, 64), "mixed")])
def test_noalloc_import_matches_set_potential(shape, precision):
    numr, numz, numphi = shape
    print(f"no-alloc import {shape} {precision}", flush=True)
    rho1 = density.scaled_binary(numr, numz, numphi, seed=21)
    rho2 = density.scaled_binary(numr, numz, numphi, seed=22, mass_ratio=0.6, separation=0.5)
    with P.PotentialPlan(numr, numz, numphi, precision=precision) as a, \
            P.PotentialPlan(numr, numz, numphi, precision=precision) as b:
        pot1 = a.potential_solver(rho1, tstep=0)
        a.set_potential(pot1)
        b.set_potential_noalloc(pot1)
        np.testing.assert_array_equal(warm_state(b), warm_state(a))
        wa = a.potential_solver(rho2, tstep=2)
        wb = b.potential_solver(rho2, tstep=2)
        np.testing.assert_array_equal(wb, wa)
