"""The command-line harness (bin/emsa_register, the counterpart of the
reference's src/main.cpp) and the file tools (add_rigid_transform, merge,
pcd2ply; src/add_rigid_transform.cpp, src/merge.cpp, src/pcd2ply.cpp), on the
CPU: file formats (ASCII / binary PLY with extra properties and elements,
PCD, text with CRLF), the zero-means-default flags, the seeded point
reduction, alignScaleOnce, (R, t) files, the per-iteration hook, and the
reference's bunny demo through emicp_cpu against the FP64 oracle
(tests/golden/bunny_oracle.npz)."""
import ctypes
import os
import struct

import numpy as np
import pytest

from tests import _harness as H
from cuda_emicp_softassign_amd.clouds import geodesic

PTS = np.array([[0.5, -1.25, 2.0], [1e-3, 2.5e-4, -7.0], [3.0, 4.0, 5.0], [-0.0625, 0.125, 9.5]],
               dtype=np.float32)


def _write(path, text, crlf=False):
    data = text.replace("\n", "\r\n") if crlf else text
    with open(path, "wb") as f:
        f.write(data.encode())


def _ply_ascii(pts, crlf=False, extra=True):
    head = ["ply", "format ascii 1.0", "comment made by the test",
            "element vertex %d" % len(pts), "property float x", "property float y",
            "property float z"]
    if extra:
        head += ["property float confidence", "property uchar intensity"]
    head += ["element face 2", "property list uchar int vertex_indices", "end_header"]
    rows = []
    for i, p in enumerate(pts):
        rows.append("%r %r %r" % tuple(float(v) for v in p) + (" 0.5 %d" % i if extra else ""))
    rows += ["3 0 1 2", "3 1 2 3"]
    return "\n".join(head + rows) + "\n"


def _read_ply_ascii(path):
    with open(path) as f:
        lines = f.read().splitlines()
    n = int([l for l in lines if l.startswith("element vertex")][0].split()[2])
    start = lines.index("end_header") + 1
    # %.9g rows: exact for float32 values read back as float32
    return np.array([[float(v) for v in l.split()[:3]] for l in lines[start:start + n]],
                    dtype=np.float32)


def _pcd2ply(tmp_path, src):
    out = str(tmp_path / "out.ply")
    H.run("pcd2ply", [src, out], str(tmp_path))
    return _read_ply_ascii(out)


@pytest.mark.parametrize("crlf", [False, True])
def test_ply_ascii_with_extra_properties_and_elements(tmp_path, crlf):
    src = str(tmp_path / "in.ply")
    _write(src, _ply_ascii(PTS, crlf=crlf), crlf=False)
    np.testing.assert_array_equal(_pcd2ply(tmp_path, src), PTS)


@pytest.mark.parametrize("endian", ["<", ">"])
def test_ply_binary(tmp_path, endian):
    fmt = "binary_little_endian" if endian == "<" else "binary_big_endian"
    head = ("ply\nformat %s 1.0\nelement vertex %d\nproperty double x\nproperty float y\n"
            "property short tag\nproperty float z\nelement edge 1\n"
            "property list uchar int vertex_indices\nend_header\n" % (fmt, len(PTS)))
    body = b"".join(struct.pack(endian + "dfhf", float(p[0]), float(p[1]), 7, float(p[2]))
                    for p in PTS)
    body += struct.pack(endian + "Bii", 2, 0, 1)
    src = str(tmp_path / "in.ply")
    with open(src, "wb") as f:
        f.write(head.encode() + body)
    np.testing.assert_array_equal(_pcd2ply(tmp_path, src), PTS)


def test_pcd_ascii_and_binary(tmp_path):
    head = ("# .PCD v0.7\nVERSION 0.7\nFIELDS x y z rgb\nSIZE 4 4 4 4\nTYPE F F F F\n"
            "COUNT 1 1 1 1\nWIDTH %d\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS %d\n"
            % (len(PTS), len(PTS)))
    src = str(tmp_path / "a.pcd")
    _write(src, head + "DATA ascii\n" + "".join("%r %r %r 1\n" % tuple(float(v) for v in p)
                                                  for p in PTS), crlf=True)
    np.testing.assert_array_equal(_pcd2ply(tmp_path, src), PTS)
    src = str(tmp_path / "b.pcd")
    with open(src, "wb") as f:
        f.write((head + "DATA binary\n").encode())
        f.write(np.hstack([PTS, np.ones((len(PTS), 1), np.float32)]).astype("<f4").tobytes())
    np.testing.assert_array_equal(_pcd2ply(tmp_path, src), PTS)


def test_text_crlf_comments_and_nan(tmp_path):
    src = str(tmp_path / "p.txt")
    rows = ["# comment", ""] + ["%r %r %r" % tuple(float(v) for v in p) for p in PTS[:2]] + \
        ["nan 1 2", "1 inf 2"] + ["%r %r %r 9 9" % tuple(float(v) for v in p) for p in PTS[2:]]
    _write(src, "\n".join(rows) + "\n", crlf=True)
    np.testing.assert_array_equal(_pcd2ply(tmp_path, src), PTS)


def test_reference_fixture_text_file(tmp_path):
    """The reference's own data file format (data/P101.txt: CRLF, 15 digits)."""
    src = os.path.join(H.ROOT, "tests", "golden", "P101.txt")
    got = _pcd2ply(tmp_path, src)
    ref = np.loadtxt(src)
    assert got.shape == ref.shape == (101, 3)
    np.testing.assert_array_equal(got, ref.astype(np.float32))


def test_bad_files_fail(tmp_path):
    bad = str(tmp_path / "bad.ply")
    _write(bad, "ply\nformat ascii 1.0\nelement vertex 5\nproperty float x\nproperty float y\n"
                "property float z\nend_header\n1 2 3\n")
    r = H.run("pcd2ply", [bad, str(tmp_path / "o.ply")], str(tmp_path), check=False)
    assert r.returncode == 1 and "short ascii data" in r.stderr
    r = H.run("pcd2ply", [str(tmp_path / "missing.ply"), str(tmp_path / "o.ply")],
              str(tmp_path), check=False)
    assert r.returncode == 1 and "cannot open" in r.stderr
    r = H.run("pcd2ply", [], str(tmp_path), check=False)
    assert r.returncode == 1 and "usage" in r.stderr


def test_add_rigid_transform_and_merge(tmp_path):
    src = str(tmp_path / "in.txt")
    np.savetxt(src, PTS)
    R = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]], np.float32)
    t = np.array([0.25, -0.5, 2.0], np.float32)
    rt = str(tmp_path / "rt.txt")
    with open(rt, "w") as f:
        f.write("".join("%f %f %f \n" % tuple(r) for r in R) + "%f %f %f \n" % tuple(t))
    for out in ("moved.pcd", "moved.ply", "moved.xyz"):
        H.run("add_rigid_transform", [src, out, rt], str(tmp_path))
        got = _pcd2ply(tmp_path, str(tmp_path / out))
        np.testing.assert_allclose(got, PTS @ R.T + t, rtol=0, atol=1e-6)
    H.run("merge", [src, "moved.pcd", "both.pcd"], str(tmp_path))
    both = _pcd2ply(tmp_path, str(tmp_path / "both.pcd"))
    np.testing.assert_allclose(both, np.vstack([PTS, PTS @ R.T + t]), rtol=0, atol=1e-6)
    with open(str(tmp_path / "both.pcd")) as f:
        head = f.read().split("DATA ascii")[0]
    assert "POINTS 8" in head and "FIELDS x y z" in head


def _params(stdout):
    out = {}
    for line in stdout.splitlines():
        w = line.split()
        if len(w) == 2:
            try:
                out[w[0]] = float(w[1])
            except ValueError:
                pass
    return out


def test_zero_means_default_flags(tmp_path):
    r = H.run("emsa_register", ["-pointFileX=none.ply", "-pointFileY=none.ply", "-softassign",
                                "-JMAX=0", "-I0=7", "--T_0=4", "-alpha=0"], str(tmp_path),
              check=False)
    assert r.returncode == 1 and "loadFile failed" in r.stderr  # after the parameters
    p = _params(r.stdout)
    assert p["JMAX"] == 100 and p["I0"] == 7 and p["I1"] == 3 and p["alpha"] == 3
    assert p["T_0"] == 4 and p["TFACTOR"] == pytest.approx(0.95)
    # the slack default follows the T_0 in effect (src/main.cpp:286)
    assert p["moutlier"] == pytest.approx(float(np.float32(1 / np.sqrt(4.0) * np.exp(-1.0))),
                                          rel=1e-5)
    r = H.run("emsa_register", ["-pointFileX=a", "-pointFileY=b", "-emicp", "-icp",
                                "-sigma_p2=0.02", "-d_02=0", "-maxIteration=0"], str(tmp_path),
              check=False)
    p = _params(r.stdout)
    assert p["sigma_p2"] == pytest.approx(0.02) and p["d_02"] == pytest.approx(0.01)
    assert p["sigma_inf"] == pytest.approx(1e-5) and p["maxIteration"] == 30
    assert "JMAX" not in p  # softassign only when selected or when nothing is
    r = H.run("emsa_register", ["-pointFileX=a"], str(tmp_path), check=False)
    assert r.returncode == 1 and "Wrong arguments" in r.stderr


def test_points_reduction_is_seeded(tmp_path):
    src = str(tmp_path / "c.txt")
    rng = np.random.default_rng(3)
    C = rng.random((20000, 3)).astype(np.float32)
    np.savetxt(src, C, fmt="%.9g")
    sizes = {}
    for seed, pre in ((0, "a"), (0, "b"), (5, "c")):
        args = ["-pointFileX=c.txt", "-pointFileY=c.txt", "-emicpcpu", "-noviewer",
                "-pointsReductionRate=10", "-sigma_p2=1", "-sigma_inf=0.5", "-saveCloudsTo=" + pre]
        if seed:
            args.append("-seed=%d" % seed)
        H.run("emsa_register", args, str(tmp_path))
        X = H.load_xyz(str(tmp_path / (pre + "_X.xyz")))
        Y = H.load_xyz(str(tmp_path / (pre + "_Y.xyz")))
        np.testing.assert_array_equal(X, Y)  # one seed for both clouds
        sizes[pre] = X
    np.testing.assert_array_equal(sizes["a"], sizes["b"])
    assert not np.array_equal(sizes["a"], sizes["c"])
    index = {C[i].tobytes(): i for i in range(len(C))}
    for X in sizes.values():
        # binomial(20000, 0.1): mean 2000, sd 42
        assert abs(len(X) - 2000) < 6 * 42
        # a subsequence of the input, order kept
        idx = [index[r.tobytes()] for r in X.astype(np.float32)]
        assert idx == sorted(idx) and len(set(idx)) == len(idx)


def test_align_scale_once(tmp_path):
    rng = np.random.default_rng(4)
    A = (rng.random((500, 3)) * 3 + 1).astype(np.float32)
    B = (rng.random((400, 3)) * 0.2 - 5).astype(np.float32)
    np.savetxt(str(tmp_path / "a.txt"), A, fmt="%.9g")
    np.savetxt(str(tmp_path / "b.txt"), B, fmt="%.9g")
    r = H.run("emsa_register", ["-pointFileX=a.txt", "-pointFileY=b.txt", "-emicpcpu",
                                "-noviewer", "-alignScaleOnce", "-sigma_p2=1", "-sigma_inf=0.5",
                                "-saveCloudsTo=s"], str(tmp_path))
    scales = [float(l.split("=")[1]) for l in r.stdout.splitlines() if l.startswith("scale")]

    def tr(P):
        P = P.astype(np.float64)
        return ((P - P.mean(0)) ** 2).sum(1).mean()

    assert scales[0] == pytest.approx(tr(A), rel=1e-5)
    assert scales[1] == pytest.approx(tr(B), rel=1e-5)
    assert scales[2] == pytest.approx(0.003, rel=1e-5) and scales[3] == pytest.approx(0.003,
                                                                                     rel=1e-5)
    for P, f in ((A, "s_X.xyz"), (B, "s_Y.xyz")):
        s = np.float32(np.sqrt(0.003) / np.sqrt(tr(P)))
        np.testing.assert_array_equal(H.load_xyz(str(tmp_path / f)).astype(np.float32), P * s)


def test_rt_files(tmp_path):
    C = np.random.default_rng(5).random((300, 3)).astype(np.float32)
    np.savetxt(str(tmp_path / "c.txt"), C, fmt="%.9g")
    with open(str(tmp_path / "rt0.txt"), "w") as f:
        f.write("0 -1 0\n1 0 0\n0 0 1\n0.5 0.25 -0.125\n")
    base = ["-pointFileX=c.txt", "-pointFileY=c.txt", "-emicpcpu", "-noviewer",
            "-sigma_p2=1", "-sigma_inf=0.5"]
    r = H.run("emsa_register", base + ["-loadRTfromFile=rt0.txt", "-saveRTtoFile=rt1.txt"],
              str(tmp_path))
    first = r.stdout.split("R\n", 1)[1].split("elapsed")[0]
    assert "0.000000 -1.000000 0.000000" in first and "0.500000 0.250000 -0.125000" in first
    with open(str(tmp_path / "rt1.txt")) as f:
        lines = f.read().split("\n")
    assert len(lines) == 5 and lines[4] == ""  # 3 rows of R, t, final newline
    assert all(len(l.split()) == 3 for l in lines[:4])
    assert all(len(v.split(".")[1]) == 6 for l in lines[:4] for v in l.split())  # "%f "
    last = r.stdout.rsplit("R\n", 1)[1]
    assert last.split() == [v for l in lines[:3] for v in l.split()] + ["t"] + lines[3].split()
    with open(str(tmp_path / "short.txt"), "w") as f:
        f.write("1 0 0 0 1 0\n")
    r = H.run("emsa_register", base + ["-loadRTfromFile=short.txt"], str(tmp_path), check=False)
    assert r.returncode == 1 and "Fail to read RT from file [short.txt]" in r.stderr


@pytest.fixture(scope="module")
def bunny(tmp_path_factory):
    d = tmp_path_factory.mktemp("bunny")
    H.unpack_bunny(str(d))
    return d, dict(np.load(H.FIXTURE, allow_pickle=False))


def test_bunny_emicp_cpu_matches_fp64_oracle(bunny):
    """The reference's demo pair through the harness with emicp_cpu (FP32, the
    reference's operation order) against the FP64 oracle of the identical
    iteration on the same prepared clouds; the per-iteration hook writes the
    66 sigma^2 steps, the last equal to the result."""
    d, g = bunny
    r = H.run("emsa_register", H.BUNNY_ARGS + ["-emicpcpu", "-saveRTtoFile=rt.txt",
                                               "-saveRTtrace=trace.txt", "-saveCloudsTo=prep"],
              str(d))
    X = H.load_xyz(str(d / "prep_X.xyz"))
    Y = H.load_xyz(str(d / "prep_Y.xyz"))
    assert H.digest(X, Y) == str(g["sha"]), "prepared clouds drifted from the fixture"
    assert "Xsize: %d" % int(g["N"]) in r.stdout and "Ysize: %d" % int(g["M"]) in r.stdout
    R, t = H.load_rt(str(d / "rt.txt"))
    ext = float((X.max(0) - X.min(0)).max())
    assert geodesic(R, g["emicp_R"]) < 1e-4
    assert np.abs(t - g["emicp_t"]).max() < 1e-5 * ext + 1e-6  # "%f" keeps 6 decimals
    tr = np.loadtxt(str(d / "trace.txt"))
    assert tr.shape == (66, 14) and (tr[:, 0] == 4).all()
    np.testing.assert_array_equal(tr[:, 1], np.arange(66))
    np.testing.assert_allclose(tr[-1, 2:11].reshape(3, 3), R, atol=1e-6)
    # the registration moves: the first step is not the identity, the last not the first
    assert geodesic(tr[0, 2:11].reshape(3, 3), np.eye(3)) > 1e-3
    assert geodesic(tr[0, 2:11].reshape(3, 3), R) > 1e-3


def test_iteration_hook_through_c_abi():
    """emsa_set_iteration_callback on the calling thread: one call per
    sigma^2 step of emsa_emicp_cpu, in order, the last with the final (R, t);
    removed again with NULL."""
    from cuda_emicp_softassign_amd import _lib
    from cuda_emicp_softassign_amd.clouds import uniform_cube
    L = _lib.load()
    X, Y, _, _ = uniform_cube(300, seed=2)
    Xs = np.ascontiguousarray(X.T, np.float32)
    Ys = np.ascontiguousarray(Y.T, np.float32)
    p = _lib.emsa_params()
    L.emsa_default_params(ctypes.byref(p))
    calls = []

    def hook(alg, it, R, t, user):
        calls.append((alg, it, np.ctypeslib.as_array(R, (9,)).copy(),
                      np.ctypeslib.as_array(t, (3,)).copy()))

    def run():
        R = np.eye(3, dtype=np.float32).ravel().copy()
        t = np.zeros(3, np.float32)
        fp = lambda a: a.ctypes.data_as(_lib._fp)  # noqa: E731
        n = L.emsa_emicp_cpu(fp(Xs), 300, fp(Ys), 300, fp(R), fp(t), ctypes.byref(p), None, 0)
        return n, R, t

    cb = _lib.emsa_iteration_fn(hook)
    L.emsa_set_iteration_callback(cb, None)
    try:
        n, R, t = run()
    finally:
        L.emsa_set_iteration_callback(_lib.emsa_iteration_fn(), None)
    assert n == 66 and len(calls) == 66
    assert [c[0] for c in calls] == [4] * 66 and [c[1] for c in calls] == list(range(66))
    np.testing.assert_array_equal(calls[-1][2], R)
    np.testing.assert_array_equal(calls[-1][3], t)
    calls.clear()
    _, R2, _ = run()
    assert calls == []
    np.testing.assert_array_equal(R2, R)
