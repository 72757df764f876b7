"""CPU checks of the C-ABI library: it loads, exports every symbol include/shipfem.h
declares, its host-side case building matches the oracle exactly, and device entry
points report NO_DEVICE (skip, not fail) where there is no GPU."""
import os
import re

import numpy as np
import pytest

import shipfem as sf
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "shipfem.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shipfem_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = sf.lib()
    names = header_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the binding declares each one
    assert set(names) == set(sf.SIGNATURES), set(names) ^ set(sf.SIGNATURES)


def test_box_mesh_matches_oracle_and_reference(golden):
    m = sf.Mesh.box(0.4, 0.1, 0.02, 16, 4, 2)
    X, idx, flags, load = m.arrays()
    assert np.array_equal(idx, golden["box_index"])
    Xo, idxo = O.box_mesh(0.4, 0.1, 0.02, 16, 4, 2)
    assert np.array_equal(X, Xo)
    assert not flags.any() and not load.any()


def test_jitter_matches_oracle_and_keeps_boundaries():
    m = sf.Mesh.box(1.0, 0.4, 0.02, 35, 14, 4).jitter(0.05, 1234)
    X, idx, _, _ = m.arrays()
    Xo, _ = O.box_mesh(1.0, 0.4, 0.02, 35, 14, 4)
    Xj = O.box_jitter(Xo, 35, 14, 4, 1.0, 0.4, 0.02, 0.05, 1234)
    assert np.array_equal(X, Xj)
    # boundary faces stay planar, interior nodes moved by at most 5% of a spacing
    for axis, lo, hi in ((0, 0.0, 1.0), (1, 0.0, 0.4), (2, 0.0, 0.02)):
        on = np.isclose(Xo[:, axis], lo) | np.isclose(Xo[:, axis], hi)
        assert np.array_equal(X[on, axis], Xo[on, axis])
    h = np.array([1.0 / 35, 0.4 / 14, 0.02 / 4])
    assert np.all(np.abs(X - Xo) <= 0.05 * h + 1e-15)
    assert np.any(np.abs(X - Xo) > 0.01 * h)


def test_jitter_refuses_inversion():
    m = sf.Mesh.box(1.0, 1.0, 1.0, 4, 4, 4)
    with pytest.raises(sf.ShipfemError):
        m.jitter(0.5, 1)


def test_panel_mesh_matches_oracle():
    args = (6.0, 3.0, 0.012, 4, 0.2, 0.012, 24, 3, 2)
    X, idx, _, _ = sf.Mesh.panel(*args).arrays()
    Xo, idxo = O.panel_mesh(*args)
    assert np.array_equal(idx, idxo) and np.array_equal(X, Xo)
    # counts in closed form: 6 tets per plate cell and per web cell
    ny = 5 * 3 + 4
    assert len(idx) == 6 * (24 * ny + 24 * 4 * 2)
    rs = O.rest_state(X, idx, 7850.0)
    assert rs["volume"].min() > 0
    # total volume: plate + four webs
    assert abs(rs["volume"].sum() - (6 * 3 * 0.012 + 4 * 6 * 0.2 * 0.012)) < 1e-12


def test_traction_matches_oracle_and_closed_form():
    m = sf.Mesh.box(1.0, 0.1, 0.1, 40, 4, 4).traction(0, 1.0, (0.0, 0.0, -1e5))
    X, idx, flags, load = m.arrays()
    lo = O.face_traction(X, idx, 0, 1.0, 1e-9, (0.0, 0.0, -1e5))
    assert np.allclose(load, lo, rtol=0, atol=1e-9)
    assert abs(load[:, 2].sum() + 1000.0) < 1e-9  # 1 kN over the 0.01 m^2 face
    assert np.all((flags & sf.LOADED) == np.where(np.abs(load).sum(1) > 0, sf.LOADED, 0))


def test_edge_timestep_matches_reference(golden):
    X, idx = O.box_mesh(0.4, 0.1, 0.02, 16, 4, 2)
    dt = sf.lib().shipfem_edge_timestep(sf._p(X, sf.C.c_double), sf._p(idx, sf.C.c_uint32),
                                        len(idx), 210e9, 7850.0, 0.5)
    assert abs(dt - float(golden["rest_dt"])) <= 1e-6 * dt


def test_configs_have_the_stated_sizes():
    from shipfem import configs

    c1 = configs.cantilever()
    assert (c1.N, c1.T) == (1025, 3840)
    assert np.allclose(c1.X[configs.TIP_NODE], [1.0, 0.05, 0.05])
    assert abs(c1.load[:, 2].sum() + 1000.0) < 1e-9
    assert np.count_nonzero(c1.flags & sf.PIN_ALL) == 25
    assert configs.panel_tets(462, 45, 16) == 812196
    nx, nyb, nz = configs.panel_cells_for(1e5)
    assert abs(configs.panel_tets(nx, nyb, nz) - 1e5) / 1e5 < 0.15


@pytest.mark.skipif(sf.device_info() is not None, reason="a device is present")
def test_no_device_is_reported_as_skip_not_failure():
    case = sf.Case(*O.box_mesh(0.1, 0.1, 0.1, 1, 1, 1))
    with pytest.raises(sf.ShipfemError) as e:
        sf.Solver(case)
    assert e.value.status == sf.NO_DEVICE


def test_failure_strain_and_torn_groups_on_the_host():
    # regularisation against the oracle (pinned to the compiled reference)
    for length in (0.005, 0.02, 0.05, 1.0):
        assert sf.regularised_failure_strain(0.12, 0.8, length, 0.02) == \
            O.regularised_failure_strain(0.12, 0.8, length, 0.02)
    # torn groups: two groups of the cantilever's halves; tear all of the first half's
    # tets and a third of the second's, by volume
    X, idx = O.box_mesh(1.0, 0.1, 0.1, 40, 4, 4)
    cx = X[idx].mean(axis=1)[:, 0]
    group = np.where(cx < 0.5, 7, 3).astype(np.uint32)
    damage = np.zeros(len(idx), np.float32)
    damage[cx < 0.5] = 1.0
    third = np.nonzero(cx >= 0.5)[0][: (cx >= 0.5).sum() // 3]
    damage[third] = 1.0
    assert sf.torn_groups(X, idx, group, damage, 0.5) == [7]
    assert sf.torn_groups(X, idx, group, damage, 0.3) == [3, 7]
    damage[:] = 0.999
    assert sf.torn_groups(X, idx, group, damage, 0.01) == []
