"""The C++ mirror of sim::solidshell (csrc/solid_shell.hpp) against the compiled
reference's golden vectors (tests/golden/reference_hex.npz): a driver compiled here with
g++ against libshipfem.so calls elementStiffness, elementMass, criticalTimestep,
internalForce, elementRotation, makePlateMesh, uniformPressureLoad and the mesh's
criticalTimestep the way reference code would, and writes what they return.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import shipfem as sf

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "sinking-simulator_amd", "lib")
GOLD = np.load(os.path.join(HERE, "golden", "reference_hex.npz"))
E, RHO = GOLD["el_material"]
PLATES = ["plate_a", "plate_b"]


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def mirror_out(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    d = tmp_path_factory.mktemp("mirror")
    exe = str(d / "solid_shell_mirror")
    subprocess.run(["g++", "-std=c++20", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "solid_shell_mirror.cpp"), "-o", exe, "-L", LIBDIR,
                    "-lshipfem", "-Wl,-rpath," + LIBDIR], check=True)
    n = len(GOLD["el_nodes"])
    parts = [np.int32(n).tobytes()]
    for k in range(n):
        parts.append(np.ascontiguousarray(GOLD["el_nodes"][k], np.float64).tobytes())
        parts.append(np.array([GOLD["el_nu"][k], E, RHO], np.float64).tobytes())
        parts.append(np.ascontiguousarray(GOLD["el_current"][k], np.float64).tobytes())
    parts.append(np.int32(len(PLATES)).tobytes())
    for p in PLATES:
        parts.append(np.asarray(GOLD[f"{p}_spec"], np.float64).tobytes())
    (d / "in.bin").write_bytes(b"".join(parts))
    subprocess.run([exe, str(d / "in.bin"), str(d / "out.bin")], check=True, timeout=120)
    return (d / "out.bin").read_bytes()


def _elements_and_plates(buf):
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(buf, dtype, count, off)
        off += a.nbytes
        return a

    els = []
    for _ in range(len(GOLD["el_nodes"])):
        els.append(dict(K=take(np.float64, 576).reshape(24, 24), mass=take(np.float64, 8),
                        dt=take(np.float64, 1)[0], force=take(np.float64, 72).reshape(3, 24),
                        rotation=take(np.float64, 27).reshape(3, 9)))
    plates = []
    for _ in PLATES:
        N, En = take(np.int32, 2)
        plates.append(dict(position=take(np.float64, 3 * N).reshape(N, 3),
                           index=take(np.uint32, 8 * En).reshape(En, 8),
                           load=take(np.float64, 3 * N).reshape(N, 3), dt=take(np.float64, 1)[0]))
    refused = take(np.int32, 1)[0]
    assert off == len(buf)
    return els, plates, refused


def test_element_functions_match_compiled_reference(mirror_out):
    els, _, _ = _elements_and_plates(mirror_out)
    for k, el in enumerate(els):
        assert rel(el["K"], GOLD["el_K"][k]) <= 1e-13
        assert rel(el["mass"], GOLD["el_mass"][k]) <= 1e-14
        assert el["dt"] == pytest.approx(GOLD["el_dt"][k], rel=1e-9)
        # the mirror forms the force from its own K; the moved element works at |x| ~ 13 m
        assert rel(el["force"], GOLD["el_force"][k]) <= (1e-9 if k == 3 else 1e-12)
        assert np.abs(el["rotation"] - GOLD["el_rotation"][k]).max() <= 1e-12


def test_plate_mesh_load_and_step_match_compiled_reference(mirror_out):
    _, plates, _ = _elements_and_plates(mirror_out)
    for name, pl in zip(PLATES, plates):
        assert np.array_equal(pl["position"], GOLD[f"{name}_position"])
        assert np.array_equal(pl["index"], GOLD[f"{name}_index"])
        assert rel(pl["load"], GOLD[f"{name}_load"]) <= 1e-14
        # the mesh step is the smallest element step (StructuralMaterial defaults)
        X, idx = pl["position"], pl["index"]
        want = min(sf.hex_critical_timestep(X[idx[e]], 206e9, 0.3, 7850.0, 0.9)
                   for e in range(len(idx)))
        assert pl["dt"] == want


def test_control_formulations_are_refused(mirror_out):
    assert _elements_and_plates(mirror_out)[2] == 1
