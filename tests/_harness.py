"""Helpers for the tests of the command-line harness (bin/emsa_register) and
the file tools: the reference's bunny scans (data/bun000.ply, data/bun045.ply,
committed gzip-compressed under tests/golden/bunny/), the runner, and the
prepared-cloud digest the golden fixture is keyed on."""
import gzip
import hashlib
import os
import shutil
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cuda_emicp_softassign_amd", "bin")
BUNNY = os.path.join(ROOT, "tests", "golden", "bunny")
FIXTURE = os.path.join(ROOT, "tests", "golden", "bunny_oracle.npz")
# the reference's demo (Makefile.demo:78-79): bun000 is the target X, bun045
# the source Y, both reduced to 5%
BUNNY_ARGS = ["-pointFileX=bun000.ply", "-pointFileY=bun045.ply", "-pointsReductionRate=5",
              "-noviewer"]


def tool(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        raise RuntimeError("%s is not built (run __graft_entry__.build())" % path)
    return path


def unpack_bunny(dest):
    for name in ("bun000.ply", "bun045.ply"):
        with gzip.open(os.path.join(BUNNY, name + ".gz"), "rb") as f, \
                open(os.path.join(dest, name), "wb") as g:
            shutil.copyfileobj(f, g)


def run(name, args, cwd, check=True, timeout=300):
    r = subprocess.run([tool(name)] + list(args), cwd=cwd, capture_output=True, text=True,
                       timeout=timeout)
    if check and r.returncode != 0:
        raise AssertionError("%s %s failed (%d):\n%s\n%s" % (name, " ".join(args), r.returncode,
                                                              r.stdout, r.stderr))
    return r


def load_xyz(path):
    return np.loadtxt(path, dtype=np.float64, ndmin=2)


def load_rt(path):
    v = np.loadtxt(path).ravel()
    assert v.size == 12
    return v[:9].reshape(3, 3), v[9:]


def digest(X, Y):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(X, np.float32).tobytes())
    h.update(np.ascontiguousarray(Y, np.float32).tobytes())
    return h.hexdigest()
