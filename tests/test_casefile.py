"""The case wire format and one-run result rows (host only).

A case survives encode/decode and a file round trip with every array bit-equal and every
scalar equal; a damaged, truncated or inconsistent blob is refused; every result row
carries the run id, commit, source digest, time and step count.
"""
import csv

import numpy as np
import pytest

import shipfem as sf
from shipfem import configs, results


def full_case():
    # every optional array and scalar set to something distinctive
    c = configs.stiffened_panel(nx=12, ny_bay=3, nz_web=2, speed=7.5, radius=0.05,
                                geometry=dict(lx=0.3, ly=0.15, n_stiff=1, web_h=0.03),
                                failure=dict(failure_strain=0.12))
    rng = np.random.default_rng(5)
    c.prescribed = rng.normal(size=(c.N, 3)) * 1e-6
    c.u0 = rng.normal(size=(c.N, 3)) * 1e-7
    c.v0 = rng.normal(size=(c.N, 3)) * 1e-3
    c.load = rng.normal(size=(c.N, 3))
    c.load_ramp, c.drive_ramp, c.gravity, c.damping = 1e-3, 2e-3, 9.81, 0.9995
    c.dt, c.dt_safety, c.sample_every, c.max_steps = 3e-7, 0.85, 50, 777
    return c


def same_case(a, b):
    for k in ("X", "index", "flags", "load", "prescribed", "u0", "v0"):
        x, y = getattr(a, k), getattr(b, k)
        assert (x is None) == (y is None), k
        if x is not None:
            assert x.dtype == y.dtype and np.array_equal(x, y), k
    for k in ("E", "nu", "rho", "plastic", "yield_", "hardening", "load_ramp", "drive_ramp",
              "gravity", "damping", "dt", "dt_safety", "sample_every", "max_steps"):
        assert getattr(a, k) == getattr(b, k), k
    assert a.failure["failure_strain"] == b.failure["failure_strain"]
    for k in ("radius", "center", "velocity"):
        assert tuple(np.atleast_1d(a.indenter[k])) == tuple(np.atleast_1d(b.indenter[k])), k


def test_case_round_trips_through_the_wire_format_and_a_file(tmp_path):
    c = full_case()
    blob = c.encode()
    assert blob[:8] == b"SHIPFEMC"
    same_case(c, sf.Case.decode(blob))
    path = tmp_path / "panel.case"
    c.save(path)
    assert path.read_bytes() == blob
    same_case(c, sf.Case.load(path))
    # a minimal case: no optional arrays
    m = configs.cantilever(nx=4, ny=1, nz=1)
    m.flags = m.load = None
    d = sf.Case.decode(m.encode())
    assert d.flags is None and d.load is None and d.u0 is None
    assert np.array_equal(d.X, m.X) and np.array_equal(d.index, m.index)


def test_damaged_blobs_are_refused():
    blob = bytearray(full_case().encode())
    bad = bytearray(blob)
    bad[200] ^= 0x01  # one flipped bit in the payload: checksum
    with pytest.raises(sf.ShipfemError):
        sf.Case.decode(bytes(bad))
    with pytest.raises(sf.ShipfemError):
        sf.Case.decode(bytes(blob[:-1]))  # truncated
    bad = bytearray(blob)
    bad[0:8] = b"NOTACASE"
    with pytest.raises(sf.ShipfemError):
        sf.Case.decode(bytes(bad))
    # a connectivity entry out of range, with a matching checksum, is still refused
    c = configs.cantilever(nx=4, ny=1, nz=1)
    c.index = c.index.copy()
    c.index[3, 2] = c.N
    with pytest.raises(sf.ShipfemError):
        sf.Case.decode(c.encode())
    with pytest.raises(sf.ShipfemError) as e:
        sf.Case.load("/nonexistent/dir/none.case")
    assert e.value.status == sf.NOT_FOUND


def test_result_rows_carry_the_run_they_came_from(tmp_path):
    ident = results.run_identity("test device")
    assert len(ident["run_id"]) == 12 and len(ident["source_digest"]) == 16
    assert ident["source_digest"] == results.source_digest()  # a function of the sources
    path = tmp_path / "rows.csv"
    rows = results.Rows(str(path), ident)
    rows.add(2000, workload="a", tets=10, ms_per_step=0.5)
    rows.add(100, workload="b", tets=20, ms_per_step=0.25)
    rows.close()
    got = list(csv.DictReader(open(path)))
    assert len(got) == 2
    for r, steps in zip(got, ("2000", "100")):
        assert r["run_id"] == ident["run_id"] and r["commit"] == ident["commit"]
        assert r["source_digest"] == ident["source_digest"] and r["utc"]
        assert r["steps"] == steps
    busy = results.host_busy(0.05)
    assert busy is None or 0.0 <= busy <= 1.0
