"""pfgpu_value_noalloc: the energy with a fixed device scratch array (for
callers that cannot allocate) equals pfgpu_value_dev's, bitwise, on two calls
in a row on one GPU, the second with a smaller mesh (runs on an MI355X)."""
import ctypes as C

import numpy as np
import pytest

from polyfem_amd import NLAssembler
from polyfem_amd._ffi import check, lib
from polyfem_amd.mesh import BASE_SEED, kuhn_box, random_lame, smooth_displacement

pytestmark = pytest.mark.gpu


def _energy_both(a, x, torch):
    dev = torch.device("cuda:0")
    dx = torch.tensor(x, device=dev)
    e_dev = torch.full((1,), np.nan, dtype=torch.float64, device=dev)
    e_na = torch.full((1,), np.nan, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    check(lib().pfgpu_value_dev(a._h, C.c_void_p(dx.data_ptr()), C.c_void_p(e_dev.data_ptr())))
    check(lib().pfgpu_synchronize(a._h))
    check(lib().pfgpu_value_noalloc(a._h, C.c_void_p(dx.data_ptr()), C.c_void_p(e_na.data_ptr())))
    check(lib().pfgpu_synchronize(a._h))
    return float(e_dev.cpu()[0]), float(e_na.cpu()[0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_value_noalloc_matches_value_dev(precision):
    torch = pytest.importorskip("torch")
    # a large call (2,239,488 tets: more than 32 x 256 energy partials, so
    # both reduction levels use the scratch) then a smaller one on the same
    # GPU's scratch
    for grid, order in [((72, 72, 72), 1), ((3, 2, 2), 2)]:
        m = kuhn_box(*grid, order=order)
        lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 900)
        x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 901)
        a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=precision)
        print(f"noalloc {precision} P{order} E={m.n_elements}", flush=True)
        e_dev, e_na = _energy_both(a, x, torch)
        assert np.isfinite(e_dev)
        assert e_na == e_dev
        assert e_na == a.assemble_energy(x)
