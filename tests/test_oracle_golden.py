"""The oracle against golden vectors recorded from the compiled reference.

Fixtures: tests/golden/reference_fem.npz, written by tests/golden/make_golden.py from
engine/sim/fem.cpp and engine/sim/plasticity.cpp. CPU only.
"""
import numpy as np
import pytest

from oracle import oracle as O

E, NU, RHO = 210e9, 0.3, 7850.0


def test_box_mesh_matches_reference_exactly(golden):
    # makeBoxTetMesh (fem.cpp:217-263): float positions and Kuhn orientation swaps
    X32, idx32 = O.box_mesh_f32(0.4, 0.1, 0.02, 16, 4, 2)
    assert np.array_equal(idx32, golden["box_index"])
    assert np.array_equal(X32.view(np.uint32), golden["box_position"].view(np.uint32))
    X, idx = O.box_mesh(0.4, 0.1, 0.02, 16, 4, 2)
    assert np.array_equal(idx, golden["box_index"])
    assert np.max(np.abs(X - golden["box_position"])) < 1e-7


def test_cantilever_mesh_matches_reference(golden):
    X, idx = O.box_mesh(1.0, 0.1, 0.1, 40, 4, 4)
    assert idx.shape == (3840, 4) and X.shape == (1025, 3)
    assert np.array_equal(idx, golden["c1_index"])
    assert np.max(np.abs(X - golden["c1_position"])) < 1e-7
    # the tip-centre node the cantilever history follows: (i,j,k) = (40,2,2)
    assert np.allclose(X[(2 * 5 + 2) * 41 + 40], [1.0, 0.05, 0.05])


def test_rest_state_fp32_restatement_is_bit_identical(golden):
    rs = O.ref32_rest_state(golden["box_position"], golden["box_index"], RHO)
    for k, g in (("rest_inv", "rest_inv"), ("volume", "rest_volume"), ("mass", "rest_mass")):
        assert np.array_equal(rs[k].view(np.uint32), golden[g].view(np.uint32)), k
    assert np.array_equal(rs["adj_offset"], golden["rest_adj_offset"])
    assert np.array_equal(rs["adj_entry"], golden["rest_adj_entry"])


def test_rest_state_fp64_oracle_agrees(golden):
    X, idx = O.box_mesh(0.4, 0.1, 0.02, 16, 4, 2)
    rs = O.rest_state(X, idx, RHO)
    assert np.array_equal(rs["adj_offset"], golden["rest_adj_offset"])
    assert np.array_equal(rs["adj_entry"], golden["rest_adj_entry"])
    g = golden["rest_inv"].astype(np.float64)
    assert np.max(np.abs(rs["rest_inv"] - g)) <= 2e-6 * np.max(np.abs(g))
    assert np.allclose(rs["volume"], golden["rest_volume"], rtol=1e-6)
    assert np.allclose(rs["mass"], golden["rest_mass"], rtol=1e-6)
    # total volume is the box volume exactly (closed form)
    assert abs(rs["volume"].sum() - 0.4 * 0.1 * 0.02) < 1e-15
    dt = O.edge_timestep(X, idx, E, RHO, 0.5)
    assert abs(dt - float(golden["rest_dt"])) <= 1e-6 * dt


def test_tet_forces_fp32_restatement_is_bit_identical(golden):
    rs = O.ref32_rest_state(golden["box_position"], golden["box_index"], RHO)
    f = O.ref32_tet_forces(golden["twist_position"], golden["box_index"], rs, E, NU)
    assert np.array_equal(f.view(np.uint32), golden["twist_forces"].view(np.uint32))


def test_tet_forces_fp64_oracle_agrees_with_reference(golden):
    # same FP32 inputs, evaluated in double: differences are the reference's own
    # float rounding (the fem_spike one-step bound is 1e-4 relative)
    rs = O.ref32_rest_state(golden["box_position"], golden["box_index"], RHO)
    lam, mu = O.lame(E, NU)
    cur = golden["twist_position"].astype(np.float64)
    idx = golden["box_index"]
    ref = golden["twist_forces"].astype(np.float64)
    out = np.array([O.tet_force(cur[idx[t]].ravel(), rs["rest_inv"][t].astype(np.float64),
                                float(rs["volume"][t]), lam, mu) for t in range(len(idx))])
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert err < 1e-4, err
    # and a plausible wrong formula does not pass: a transposed Dm^-1 is far off
    bad = np.array([O.tet_force(cur[idx[t]].ravel(),
                                rs["rest_inv"][t].astype(np.float64).reshape(3, 3).T.ravel(),
                                float(rs["volume"][t]), lam, mu) for t in range(40)])
    assert np.linalg.norm(bad - ref[:40]) / np.linalg.norm(ref[:40]) > 1e-2


@pytest.mark.parametrize("threads", [1, 3])
def test_step_fp32_restatement_is_bit_identical(golden, threads):
    sim = O.Ref32(golden["box_position"], golden["box_index"], golden["twist_fixed"], E, NU, RHO)
    sim.position[:] = golden["twist_position"]
    dt = float(golden["rest_dt"])
    for n in range(10):
        sim.step(dt, 9.80665, 1.0, threads=threads)
        if n == 0:
            assert np.array_equal(sim.position.view(np.uint32),
                                  golden["step1_position"].view(np.uint32))
            assert np.array_equal(sim.velocity.view(np.uint32),
                                  golden["step1_velocity"].view(np.uint32))
    assert np.array_equal(sim.position.view(np.uint32), golden["step10_position"].view(np.uint32))
    assert np.array_equal(sim.velocity.view(np.uint32), golden["step10_velocity"].view(np.uint32))


def test_step_fp64_oracle_tracks_reference(golden):
    # fem_spike's own GPU-vs-CPU bounds (tools/fem_spike/main.cpp:210-211): one step
    # 1e-4, ten steps 1e-3 relative RMS velocity. What remains is the reference's
    # own FP32 rounding of absolute positions (it reads 2e-5 after one step).
    X = golden["box_position"].astype(np.float64)
    idx = golden["box_index"]
    dt = float(golden["rest_dt"])
    flags = np.where(golden["twist_fixed"] != 0, O.PIN_ALL, 0).astype(np.uint32)
    case = O.Case(X, idx, E=E, nu=NU, rho=RHO, dt=dt, flags=flags, gravity=9.80665)
    case.u[:] = golden["twist_position"].astype(np.float64) - X
    for steps, tol in ((1, 1e-4), (10, 1e-3)):
        case.run(steps - case.step)
        gv = golden[f"step{steps}_velocity"].astype(np.float64)
        rel = np.linalg.norm(case.v - gv) / np.linalg.norm(gv)
        assert rel < tol, (steps, rel)
        gx = golden[f"step{steps}_position"].astype(np.float64)
        assert np.max(np.abs(X + case.u - gx)) < 1e-7


@pytest.mark.parametrize("path", ["uniaxial", "shear", "mixed"])
def test_j2_matches_reference_point_law(golden, path):
    Ej, nuj, sy, H = golden["j2_material"]
    ep = np.zeros(6)
    q = 0.0
    for i, s in enumerate(golden[f"j2_{path}_strain"]):
        r = O.j2_update(Ej, nuj, sy, H, s, ep, q)
        ep, q = r["eps_p"], r["eqps"]
        assert r["yielded"] == bool(golden[f"j2_{path}_yielded"][i])
        assert np.max(np.abs(r["stress"] - golden[f"j2_{path}_stress"][i])) <= 1e-9 * sy
        assert np.max(np.abs(ep - golden[f"j2_{path}_eps_p"][i])) <= 1e-13
        assert abs(q - golden[f"j2_{path}_eqps"][i]) <= 1e-13
        d = golden[f"j2_{path}_dissipation"][i]
        assert abs(r["dissipation"] - d) <= 1e-9 * max(abs(d), sy * 1e-6)
    # linear hardening: the reference's Newton loop takes exactly two evaluations
    it = golden[f"j2_{path}_iterations"]
    assert set(it[golden[f"j2_{path}_yielded"] == 1].tolist()) <= {2}


def test_failure_strain_regularisation_and_triaxiality_match_reference(golden):
    eu, efr, thick = golden["fail_reg_input"]
    for length, ref in zip(golden["fail_reg_length"], golden["fail_reg_strain"]):
        assert O.regularised_failure_strain(eu, efr, length, thick) == pytest.approx(ref, rel=1e-15)
    for eta, ref in zip(golden["fail_eta"], golden["fail_factor"]):
        got = O.triaxiality_factor(eta)
        assert (np.isinf(got) and np.isinf(ref)) or got == pytest.approx(ref, rel=1e-14)


@pytest.mark.parametrize("path", ["pull", "rotated", "squeeze", "biaxial", "mixed"])
def test_j2_damage_matches_reference_point_law(golden, path):
    # plasticity::update with ductile failure (plasticity.cpp:457-477), state carried
    Ej, nuj, sy, H, ef, sens, eref, ecut = golden["fail_material"]
    ep, q, dmg = np.zeros(6), 0.0, 0.0
    for i, s in enumerate(golden[f"dmg_{path}_strain"]):
        r = O.j2_update_damage(Ej, nuj, sy, H, ef, s, ep, q, dmg, sens, eref, ecut)
        ep, q, dmg = r["eps_p"], r["eqps"], r["damage"]
        assert (dmg >= 1.0) == bool(golden[f"dmg_{path}_failed"][i])
        assert r["failed_now"] == bool(golden[f"dmg_{path}_failed_now"][i])
        assert dmg == pytest.approx(golden[f"dmg_{path}_damage"][i], rel=1e-12, abs=1e-15)
        assert np.max(np.abs(r["stress"] - golden[f"dmg_{path}_stress"][i])) <= 1e-9 * sy
        assert q == pytest.approx(golden[f"dmg_{path}_eqps"][i], rel=1e-12, abs=1e-15)
        if r["failed_now"]:
            # an eigenvector's sign is fixed by its largest component in both
            assert np.allclose(r["normal"], golden[f"dmg_{path}_normal"][i], atol=1e-9)
