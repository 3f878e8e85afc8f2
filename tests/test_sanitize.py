"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only):

- libemsa built with the sanitizers on its host side (every source through
  hipcc with -Xarch_host before each flag; device code unchanged), driven by
  tests/cpp/sanitize_main.cpp: the drop-in header's host functions, the FP32
  host EM-ICP with trace and hook, and the argument checks of the host entry
  points;
- the command-line harness and the file loaders built with the sanitizers,
  on well-formed and malformed inputs (truncated binary PLY, list counts past
  the data, short PCD, junk text);
- the C oracle built with the sanitizers, loaded into a Python process that
  runs every oracle function on small inputs.

A sanitizer report makes the run fail (halt_on_error); leak checking is off
(the HIP runtime keeps allocations until exit)."""
import os
import struct
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from tests import _harness as H

ROOT = H.ROOT
CLANG = "/opt/rocm/llvm/bin/clang++"
SAN = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=all",
       "-g"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _ok(r, what):
    assert r.returncode == 0, "%s failed (%d):\n%s\n%s" % (what, r.returncode, r.stdout,
                                                             r.stderr[-4000:])
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr


@pytest.fixture(scope="module")
def san(tmp_path_factory):
    if not os.path.exists(CLANG):
        pytest.skip("no clang++ in /opt/rocm/llvm")
    d = str(tmp_path_factory.mktemp("san"))
    from cuda_emicp_softassign_amd import build as B
    host = []
    for f in SAN:
        host += ["-Xarch_host", f]
    lib = B.build(host_flags=host, out=os.path.join(d, "libemsa.so"))
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + B.TOOLS]
    link = ["-L" + d, "-lemsa", "-Wl,-rpath," + d, "-Wl,-rpath-link,/opt/rocm/lib",
            "-L/opt/rocm/lib", "-lamdhip64"]
    exe = {}
    for name, srcs, uses in (
            ("drv", [os.path.join(ROOT, "tests", "cpp", "sanitize_main.cpp")], True),
            ("emsa_register", [os.path.join(B.TOOLS, "emsa_register.cpp"),
                               os.path.join(B.TOOLS, "cloud_io.cpp")], True),
            ("pcd2ply", [os.path.join(B.TOOLS, "pcd2ply.cpp"),
                         os.path.join(B.TOOLS, "cloud_io.cpp")], False)):
        out = os.path.join(d, name)
        subprocess.run([CLANG, "-std=c++17", "-O1"] + SAN + inc + srcs + ["-o", out] +
                       (link if uses else []), check=True)
        exe[name] = out
    so = os.path.join(d, "liboracle.so")
    subprocess.run(["gcc", "-O1", "-fopenmp", "-fPIC", "-shared", "-std=c99"] + SAN +
                   [os.path.join(ROOT, "oracle", "emsa_oracle.c"), "-o", so, "-lm"], check=True)
    return d, lib, exe, so


def test_libemsa_host_code(san):
    d, _, exe, _ = san
    r = subprocess.run([exe["drv"]], capture_output=True, text=True, env=ENV, timeout=300)
    _ok(r, "sanitize_main")
    assert "sanitize OK" in r.stdout


def test_harness_and_loaders(san, tmp_path):
    d, _, exe, _ = san
    rng = np.random.default_rng(9)
    C = rng.random((400, 3)).astype(np.float32)
    np.savetxt(str(tmp_path / "c.txt"), C, fmt="%.9g")
    r = subprocess.run([exe["emsa_register"], "-pointFileX=c.txt", "-pointFileY=c.txt",
                        "-emicpcpu", "-noviewer", "-pointsReductionRate=50", "-alignScaleOnce",
                        "-sigma_inf=0.001", "-saveRTtoFile=rt.txt", "-saveRTtrace=tr.txt",
                        "-saveCloudsTo=p"], cwd=str(tmp_path), capture_output=True, text=True,
                       env=ENV, timeout=300)
    _ok(r, "emsa_register")
    r = subprocess.run([exe["emsa_register"], "-pointFileX=c.txt", "-pointFileY=c.txt",
                        "-loadRTfromFile=c.txt", "-emicpcpu"], cwd=str(tmp_path),
                       capture_output=True, text=True, env=ENV, timeout=300)
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr

    # malformed inputs: every one fails cleanly (status 1) or loads what is there
    head = ("ply\nformat binary_little_endian 1.0\nelement vertex 4\nproperty float x\n"
            "property float y\nproperty float z\nelement face 1\n"
            "property list uchar int vertex_indices\nend_header\n").encode()
    body = np.arange(12, dtype="<f4").tobytes() + struct.pack("<B", 200) + b"\0" * 8
    cases = {"trunc%d.ply" % k: head + body[:k] for k in (0, 5, 24, 47, 48, 49, 52)}
    cases["list.ply"] = head + body
    cases["ascii_list.ply"] = (b"ply\nformat ascii 1.0\nelement face 1\nproperty list uchar "
                               b"int v\nelement vertex 1\nproperty float x\nproperty float y\n"
                               b"property float z\nend_header\n9 1 2\n")
    cases["nohead.ply"] = b"ply\nformat ascii 1.0\nelement vertex 2\n"
    cases["badtype.ply"] = b"ply\nformat ascii 1.0\nelement vertex 1\nproperty half x\nend_header\n"
    cases["short.pcd"] = (b"VERSION 0.7\nFIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nCOUNT 1 1 1\n"
                          b"POINTS 9\nDATA binary\n" + b"\0" * 20)
    cases["mism.pcd"] = b"FIELDS x y z\nSIZE 4 4\nTYPE F F F\nPOINTS 1\nDATA ascii\n1 2 3\n"
    cases["junk.txt"] = bytes(rng.integers(0, 256, 4000, dtype=np.uint8))
    for name, data in cases.items():
        with open(str(tmp_path / name), "wb") as f:
            f.write(data)
        r = subprocess.run([exe["pcd2ply"], name, "o.ply"], cwd=str(tmp_path),
                           capture_output=True, text=True, env=ENV, timeout=60)
        assert r.returncode in (0, 1), (name, r.returncode, r.stderr[-2000:])
        assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, name
    ok = subprocess.run([exe["pcd2ply"], "trunc48.ply", "o.ply"], cwd=str(tmp_path),
                        capture_output=True, text=True, env=ENV)
    assert ok.returncode == 0  # all 4 vertices present; the face list is not needed


def test_oracle(san):
    d, _, _, so = san
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                             text=True).stdout.strip()
    code = textwrap.dedent("""
        import numpy as np
        from oracle import oracle as O
        from cuda_emicp_softassign_amd.clouds import uniform_cube
        X, Y, _, _ = uniform_cube(120, seed=3)
        for prec in ("f64", "f32"):
            dt = np.float64 if prec == "f64" else np.float32
            O.emicp(X.astype(dt), Y.astype(dt), precision=prec, sigma_inf=1e-3)
            O.emicp_step(X.astype(dt), Y.astype(dt), np.eye(3), np.zeros(3), 1e-3, precision=prec)
            O.softassign(X.astype(dt), Y[:100].astype(dt), JMAX=3, I0=2, precision=prec)
            o_r = np.ones(120, dt); o_c = np.ones(100, dt)
            O.softassign_inner(X.astype(dt), Y[:100].astype(dt), np.eye(3), np.zeros(3), 1.0,
                               3.0, 3, o_r, o_c, precision=prec)
            O.find_rt_from_s(np.zeros(3), np.zeros(3), np.eye(3).ravel(), precision=prec)
        O.icp(X, Y, maxIteration=3)
        O.icp_nearest(X, Y)
        O.jacobi4(np.eye(4) + 0.1)
        print("oracle OK")
    """)
    env = dict(ENV, LD_PRELOAD=libasan, EMSA_ORACLE_LIB=so, PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                       timeout=600, cwd=ROOT)
    _ok(r, "oracle under the sanitizers")
    assert "oracle OK" in r.stdout
