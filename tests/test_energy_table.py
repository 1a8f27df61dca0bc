"""The library's Kulla-Conty table (mrg_conductor_energy_table, host code) equals the oracle's table
bit for bit: both bake the reference's bake_ibl_brdf() defaults in FP64 with the same operation order."""
import ctypes as C

import numpy as np

from minirender_amd import _native as N


def test_library_table_equals_oracle(oracle):
    out = np.zeros((64, 64, 3), np.float32)
    N.check(N.lib().mrg_conductor_energy_table(out.ctypes.data_as(C.c_void_p)))
    np.testing.assert_array_equal(out, oracle.energy_table())
