"""shipfem_stable_timestep_noalloc, which stages the mesh in the library's static device
array instead of allocating, against shipfem_stable_timestep: the same bits on two calls
in a row, the second on a smaller mesh (the array keeps the first call's contents), in
either order; and a mesh larger than the array is refused with a message."""
import numpy as np
import pytest

import shipfem as sf
from shipfem import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if sf.device_info() is None:
        pytest.skip("no usable gfx950 device")


def test_noalloc_stable_step_equals_the_allocating_one():
    big = configs.plate(lx=2.0, ly=0.8, nx=70, ny=28)              # 47k tets
    small = configs.stiffened_panel(nx=24, ny_bay=4, nz_web=3, radius=0.05,
                                    geometry=dict(lx=0.36, ly=0.2, n_stiff=1, web_h=0.04))
    print("meshes:", big.T, "and", small.T, "tets", flush=True)
    got = []
    for case in (big, small, big):
        a = sf.stable_timestep(case.X, case.index, noalloc=True)
        b = sf.stable_timestep(case.X, case.index)
        c = sf.stable_timestep(case.X, case.index, noalloc=True)
        assert a == b == c, (case.T, a, b, c)
        got.append(a)
    assert got[0] == got[2] and got[0] != got[1]
    # and it is the FP64 bound of the oracle
    rs = O.rest_state(small.X, small.index, 7850.0)
    assert abs(got[1] * O.max_element_omega(rs["rest_inv"], 210e9, 0.3, 7850.0) - 2.0) <= 1e-11


def test_noalloc_refuses_a_mesh_beyond_its_array():
    # 24 B per node + 16 B per tet + 8 B per 64 tets against 32 MiB: refused on the host
    N, T = 800_000, 1_000_000
    X = np.zeros((N, 3))
    idx = np.zeros((T, 4), np.uint32)
    with pytest.raises(sf.ShipfemError) as e:
        sf.stable_timestep(X, idx, noalloc=True)
    assert e.value.status == sf.INVALID_ARG and "bytes of device scratch" in str(e.value)
