"""The reference's bunny demo (Makefile.demo:78-79 and :92-95: bun045 onto
bun000, both reduced to 5%, reference defaults) through the command-line
harness bin/emsa_register on the GPU: EM-ICP, Softassign (both variants) and
ICP against the FP64 oracle of the identical iterations on the same prepared
clouds (tests/golden/bunny_oracle.npz, written by
tests/golden/make_bunny_oracle.py). Bar: R within 1e-4 rad, t within 1e-5 of
the target's extent (plus the 1e-6 of the "%f" (R, t) file)."""
import numpy as np
import pytest

from cuda_emicp_softassign_amd.clouds import geodesic
from tests import _harness as H

pytestmark = pytest.mark.gpu

CASES = {"emicp": ["-emicp"], "softassign": ["-softassign"],
         "softassign_logdomain": ["-softassign", "-logdomain"], "icp": ["-icp"]}
STEPS = {"emicp": 66, "softassign": 500, "softassign_logdomain": 500, "icp": 30}
ALG_ID = {"emicp": 1, "softassign": 2, "softassign_logdomain": 2, "icp": 3}


@pytest.fixture(scope="module")
def bunny(tmp_path_factory):
    d = tmp_path_factory.mktemp("bunny_gpu")
    H.unpack_bunny(str(d))
    return d, dict(np.load(H.FIXTURE, allow_pickle=False))


@pytest.mark.parametrize("case", list(CASES))
def test_bunny_through_harness_matches_fp64_oracle(bunny, case):
    d, g = bunny
    print("harness bunny case %s" % case, flush=True)
    r = H.run("emsa_register", H.BUNNY_ARGS + CASES[case] + [
        "-saveRTtoFile=%s.txt" % case, "-saveRTtrace=%s_trace.txt" % case,
        "-saveCloudsTo=%s_prep" % case], str(d))
    X = H.load_xyz(str(d / ("%s_prep_X.xyz" % case)))
    Y = H.load_xyz(str(d / ("%s_prep_Y.xyz" % case)))
    assert H.digest(X, Y) == str(g["sha"]), "prepared clouds drifted from the fixture"
    R, t = H.load_rt(str(d / ("%s.txt" % case)))
    key = "softassign" if case.startswith("softassign") else case
    ext = float((X.max(0) - X.min(0)).max())
    err = geodesic(R, g[key + "_R"])
    print("  %s: %s; rot err vs FP64 oracle %.3g rad, |dt| %.3g" % (
        case, r.stdout.split("elapsed")[1].split("\n")[0].strip(), err,
        np.abs(t - g[key + "_t"]).max()), flush=True)
    assert err < 1e-4
    assert np.abs(t - g[key + "_t"]).max() < 1e-5 * ext + 1e-6
    tr = np.loadtxt(str(d / ("%s_trace.txt" % case)), ndmin=2)
    assert tr.shape == (STEPS[case], 14) and (tr[:, 0] == ALG_ID[case]).all()
    np.testing.assert_allclose(tr[-1, 2:11].reshape(3, 3), R, atol=1e-6)


def test_algorithms_chain(bunny):
    """Several algorithms in one run continue from each other's (R, t), in the
    reference's order (icp, then emicp): the trace holds 30 ICP iterations,
    then 66 EM-ICP steps; the result equals EM-ICP started from ICP's result."""
    d, _ = bunny
    H.run("emsa_register", H.BUNNY_ARGS + ["-emicp", "-icp", "-saveRTtoFile=chain.txt",
                                           "-saveRTtrace=chain_trace.txt"], str(d))
    tr = np.loadtxt(str(d / "chain_trace.txt"))
    assert tr.shape == (96, 14)
    assert (tr[:30, 0] == 3).all() and (tr[30:, 0] == 1).all()
    with open(str(d / "icp_rt.txt"), "w") as f:
        R = tr[29, 2:11].reshape(3, 3)
        f.write("".join("%.9g %.9g %.9g\n" % tuple(r) for r in R) + "%.9g %.9g %.9g\n"
                % tuple(tr[29, 11:14]))
    H.run("emsa_register", H.BUNNY_ARGS + ["-emicp", "-loadRTfromFile=icp_rt.txt",
                                           "-saveRTtoFile=from_icp.txt"], str(d))
    R1, t1 = H.load_rt(str(d / "chain.txt"))
    R2, t2 = H.load_rt(str(d / "from_icp.txt"))
    assert geodesic(R1, R2) < 1e-5 and np.abs(t1 - t2).max() < 2e-6
