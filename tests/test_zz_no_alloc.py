"""The allocation-free entry point sphnj_energy_noalloc against sphnj_energy, through the C
ABI: two calls in a row on one GPU (the second on a smaller case, so the shared per-GPU
scratch holds the first call's contents when it starts), each bit for bit equal to the
allocating call on the same state."""
import ctypes as C

import numpy as np
import pytest

from sphnj_amd import Case, Simulation, library

pytestmark = pytest.mark.gpu


def _both(sim):
    lib = library()
    a = np.zeros(6)
    b = np.zeros(6)
    p = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    assert lib.sphnj_energy_noalloc(sim._h, 0, p(b)) == 0
    assert lib.sphnj_energy(sim._h, 0, p(a)) == 0
    return a, b


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_energy_noalloc_equals_energy(gpu, precision):
    big = Simulation(Case.generate("gate_fsi", dx=2.0e-3, tank_w=0.3, tank_h=0.2, S=4), precision)
    big.step(5)
    a, b = _both(big)
    assert np.array_equal(a, b), (a, b)
    assert a[0] > 0 and a[3] > 0 and a[5] == big.cases[0].counts[0]
    small = Simulation(Case.generate("dam_break", tank_w=0.6, tank_h=0.5, col_w=0.2, col_h=0.3, dx=0.02), precision)
    small.step(3)
    a2, b2 = _both(small)
    assert np.array_equal(a2, b2), (a2, b2)
    assert a2[5] == small.cases[0].counts[0] < a[5] and a2[3] == 0.0
    # a third call on the first context still reads its own state only
    a3, b3 = _both(big)
    assert np.array_equal(a3, a) and np.array_equal(b3, a)
