"""Writes tests/golden/bunny_oracle.npz: the FP64 oracle's EM-ICP, Softassign
and ICP registrations of the reference's bunny demo (Makefile.demo:78-79 and
:92-95: bun045 onto bun000, both reduced to 5%, reference default
parameters), on the clouds exactly as bin/emsa_register prepares them
(default seed), with their SHA-256 so that a test can check it prepared the
same input. Run after building: python tests/golden/make_bunny_oracle.py
(about a minute on 8 cores)."""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests import _harness as H  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    with tempfile.TemporaryDirectory() as d:
        H.unpack_bunny(d)
        H.run("emsa_register", H.BUNNY_ARGS + ["-emicpcpu", "-saveCloudsTo=prep"], d)
        X = H.load_xyz(os.path.join(d, "prep_X.xyz"))
        Y = H.load_xyz(os.path.join(d, "prep_Y.xyz"))
    X32, Y32 = X.astype(np.float32), Y.astype(np.float32)
    X64, Y64 = X32.astype(np.float64), Y32.astype(np.float64)
    out = {"sha": np.array(H.digest(X32, Y32)), "N": np.array(len(X)), "M": np.array(len(Y))}
    R, t, _ = O.emicp(X64, Y64)
    out["emicp_R"], out["emicp_t"] = R, t
    R, t, _ = O.softassign(X64, Y64)
    out["softassign_R"], out["softassign_t"] = R, t
    R, t, _ = O.icp(X64, Y64)
    out["icp_R"], out["icp_t"] = R, t
    np.savez_compressed(H.FIXTURE, **out)
    print("wrote", H.FIXTURE, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
