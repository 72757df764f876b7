"""One-run result rows: every row of a table carries the run it came from.

Each row is stamped with the run id, the commit the library was built from, a digest of
the sources that built it, a UTC timestamp and the step count, so a table transcribed
from one file is a table from one run at one set of parameters (the discipline of the
reference's tools/zone_gpu_probe/sweep.sh:21-25). Timing rows also carry the host's busy
fraction in the second before the run, from /proc/stat.
"""
import csv
import datetime
import hashlib
import json
import os
import platform
import time
import uuid

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
STAMP = ["run_id", "commit", "source_digest", "utc", "steps"]


def source_digest():
    """sha256 over the library's sources and the ABI header, in a fixed file order."""
    h = hashlib.sha256()
    files = [os.path.join(ROOT, "include", "shipfem.h")]
    csrc = os.path.join(PKG, "csrc")
    files += [os.path.join(csrc, f) for f in sorted(os.listdir(csrc))]
    for path in files:
        h.update(os.path.basename(path).encode())
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def build_commit():
    try:
        with open(os.path.join(PKG, "lib", "build_info.json")) as f:
            info = json.load(f)
        return info["commit"][:12] + ("+dirty" if info.get("dirty") else "")
    except (OSError, ValueError, KeyError):
        return "unknown"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _cpu_times():
    with open("/proc/stat") as f:
        vals = [int(x) for x in f.readline().split()[1:]]
    idle = vals[3] + (vals[4] if len(vals) > 4 else 0)
    return sum(vals), idle


def host_busy(window=0.5):
    """Busy fraction of all host cores over `window` seconds (a /proc/stat delta, which
    sees only what runs during the window; None where /proc/stat is absent)."""
    try:
        t0, i0 = _cpu_times()
        time.sleep(window)
        t1, i1 = _cpu_times()
    except OSError:
        return None
    return 1.0 - (i1 - i0) / max(t1 - t0, 1)


def run_identity(device=None):
    return dict(run_id=uuid.uuid4().hex[:12], commit=build_commit(), source_digest=source_digest(),
                started_utc=datetime.datetime.now(datetime.timezone.utc).isoformat(timespec="seconds"),
                host_cpu=cpu_model(), host_cores=os.cpu_count(), device=device)


class Rows:
    """CSV rows stamped with the run's identity; columns are fixed by the first row."""

    def __init__(self, path, identity):
        self.path, self.ident = path, identity
        self.f = open(path, "w", newline="")
        self.w = None

    def add(self, steps, **fields):
        row = dict(run_id=self.ident["run_id"], commit=self.ident["commit"],
                   source_digest=self.ident["source_digest"],
                   utc=datetime.datetime.now(datetime.timezone.utc).isoformat(timespec="seconds"),
                   steps=steps)
        row.update(fields)
        if self.w is None:
            self.w = csv.DictWriter(self.f, fieldnames=list(row))
            self.w.writeheader()
        self.w.writerow(row)
        self.f.flush()
        return row

    def close(self):
        self.f.close()
