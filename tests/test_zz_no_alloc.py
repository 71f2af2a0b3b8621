"""misa_lat_hook_noalloc, the hook calls without any device allocation, against
misa_lat_rho / misa_lat_df / misa_lat_force through the C-ABI (needs an
MI355X).

Both run the same kernels on the reference's AoS array, so every output must
be bit-identical.  Two calls in a row share the library's fixed staging
array: the first on an 8x8x9 box, the second on a smaller 6x7x6 box, so the
second call runs over what the first left in the array.  A box whose array
exceeds the staging array is refused with MISA_E_ARG.
"""
import ctypes

import numpy as np
import pytest

import misa_eam as M
import oracle_py as O

pytestmark = pytest.mark.gpu

A = 2.855
F = 5.6 / 2.855


def _state(setfl, cells, seed):
    sim = O.Sim(cells, A, F, setfl)
    sim.build_world(seed, 600.0, (900, 70, 30), seed)
    sb = sim.sub_box()
    rng = np.random.default_rng(seed)
    sb["x"] = sb["x"] + rng.uniform(-0.1, 0.1, sb["x"].shape)
    sim.phase("exchange_atom_first")
    sim.phase("clear_force")
    return sim


def _hooks(dev, aos, rc, noalloc):
    """rho, df, force on `aos` in place, as atom::computeEam calls the hooks."""
    lib = dev.lib
    ptr = ctypes.c_void_p(aos.ctypes.data)
    for which, fn in ((0, lib.misa_lat_rho), (1, lib.misa_lat_df), (2, lib.misa_lat_force)):
        rc_ = lib.misa_lat_hook_noalloc(dev.h, which, ptr, rc) if noalloc else fn(dev.h, ptr, rc)
        assert rc_ == 0, lib.misa_last_error(dev.h).decode()


def test_noalloc_hooks_match_the_allocating_hooks(setfl_fecuni):
    for cells, seed in (((8, 8, 9), 11), ((6, 7, 6), 12)):
        print("no-alloc hooks", cells, flush=True)
        sim = _state(setfl_fecuni, cells, seed)
        dev = M.EamDevice(cells)
        dev.make_offsets()
        dev.set_potential(setfl_fecuni)
        want = sim.atoms().copy()
        got = sim.atoms().copy()
        _hooks(dev, want, sim.rc, noalloc=False)
        _hooks(dev, got, sim.rc, noalloc=True)
        for k in ("rho", "df", "f", "x", "type", "id"):
            np.testing.assert_array_equal(got[k], want[k], err_msg="%s %s" % (cells, k))
        sub = M.sub_box_view(got, dev)
        assert np.all(sub["rho"][sub["type"] >= 0] > 0)
        dev.close()


def test_noalloc_refuses_a_box_beyond_the_staging_array(setfl_fecuni):
    cells = (50, 50, 50)  # 2 * 56^3 slots * 104 B = 36.5 MB > 32 MiB
    dev = M.EamDevice(cells)
    dev.make_offsets()
    dev.set_potential(setfl_fecuni)
    aos = dev.empty_atoms()
    rc = dev.lib.misa_lat_hook_noalloc(dev.h, 0, ctypes.c_void_p(aos.ctypes.data), 5.6)
    assert rc == M.E_ARG
    assert "staging array" in dev.lib.misa_last_error(dev.h).decode()
    dev.close()
