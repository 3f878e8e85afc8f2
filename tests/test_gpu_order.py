"""The working order EM-ICP gives each cloud (csrc/emicp.hip, emicp_order_kernel):
a stable sort of the points by a 30-bit 3D Hilbert key over the cloud's
bounding box, non-finite points last. Checked exactly against a numpy
restatement: the same FP32 quantisation to 10 bits per axis, Skilling's
transpose algorithm ("Programming the Hilbert curve", AIP Conf. Proc. 707,
2004) for the key, and numpy's stable argsort. The key itself is pinned by the
curve's defining property on a 16^3 grid (consecutive cells one step apart).
"""
import numpy as np
import pytest

BITS = 10


def hilbert_keys(q):
    """Keys of integer cells q (n x 3, values < 2^BITS), transposed-index
    algorithm, bits interleaved from the most significant level."""
    x = [q[:, 0].astype(np.int64), q[:, 1].astype(np.int64), q[:, 2].astype(np.int64)]
    m = 1 << (BITS - 1)
    b = m
    while b > 1:
        p = b - 1
        for i in range(3):
            hit = (x[i] & b) != 0
            x[0] = np.where(hit, x[0] ^ p, x[0])
            t = np.where(hit, 0, (x[0] ^ x[i]) & p)
            x[0] = x[0] ^ t
            x[i] = x[i] ^ t
        b >>= 1
    x[1] ^= x[0]
    x[2] ^= x[1]
    t = np.zeros_like(x[0])
    b = m
    while b > 1:
        t = np.where((x[2] & b) != 0, t ^ (b - 1), t)
        b >>= 1
    key = np.zeros(len(q), np.int64)
    for lev in range(BITS - 1, -1, -1):
        for i in range(3):
            key = (key << 1) | (((x[i] ^ t) >> lev) & 1)
    return key


def expected_order(P):
    """P: n x 3 float32. The permutation the kernel must produce."""
    P = P.astype(np.float32)
    fin = np.isfinite(P).all(axis=1)
    lo = np.array([P[np.isfinite(P[:, a]), a].min() if np.isfinite(P[:, a]).any() else np.inf
                   for a in range(3)], np.float32)
    hi = np.array([P[np.isfinite(P[:, a]), a].max() if np.isfinite(P[:, a]).any() else -np.inf
                   for a in range(3)], np.float32)
    q = np.zeros((len(P), 3), np.int64)
    with np.errstate(invalid="ignore", over="ignore"):
        for a in range(3):
            scale = (np.float32((1 << BITS) - 1) / (hi[a] - lo[a])) if hi[a] > lo[a] else np.float32(0)
            f = (P[:, a] - lo[a]) * np.float32(scale)
            f = np.where(np.isnan(f), np.float32(0), f)
            f = np.minimum(np.maximum(f, np.float32(0)), np.float32((1 << BITS) - 1))
            q[:, a] = f.astype(np.int64)
    keys = np.where(fin, hilbert_keys(q), 0xFFFFFFFF)
    return np.argsort(keys, kind="stable")


def test_hilbert_curve_steps_are_unit():
    """CPU: on a 16^3 corner of the 1024^3 grid (a complete level-4 cell)
    the curve visits every cell once, moving one cell at a time."""
    g = np.array([[x, y, z] for x in range(16) for y in range(16) for z in range(16)])
    k = hilbert_keys(g)
    assert len(set(k.tolist())) == len(k)
    s = g[np.argsort(k)]
    assert (np.abs(np.diff(s, axis=0)).sum(axis=1) == 1).all()


@pytest.mark.gpu
# up to 16384 points the sort keeps its keys in registers and LDS; larger
# clouds take the global-memory passes
@pytest.mark.parametrize("n", [1, 2, 63, 1000, 4096, 4097, 10000, 12345, 16383, 16384, 16385,
                               40000])
def test_order_matches_numpy(n):
    import torch
    from cuda_emicp_softassign_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(n)
    P = rng.uniform(-0.7, 1.3, size=(n, 3)).astype(np.float32)
    if n > 10:
        P[5:9] = P[2]            # duplicates: equal keys keep index order
        P[n // 3, 1] = np.nan    # non-finite points go last, in index order
        P[n // 2, 0] = np.inf
    dP = torch.from_numpy(np.ascontiguousarray(P.T)).cuda()
    dperm = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    _lib.check("hilbert_order", L.emsa_probe_hilbert_order(dP.data_ptr(), n, dperm.data_ptr()))
    got = dperm.cpu().numpy().astype(np.int64)
    print("case n=%d" % n, flush=True)
    np.testing.assert_array_equal(got, expected_order(P))


@pytest.mark.gpu
def test_order_flat_cloud():
    """A cloud with zero extent along an axis (all z equal) and all points
    equal: a zero scale keeps every key defined."""
    import torch
    from cuda_emicp_softassign_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(3)
    for P in (np.c_[rng.uniform(size=(500, 2)), np.full(500, 0.25)].astype(np.float32),
              np.tile(np.float32([[0.1, 0.2, 0.3]]), (300, 1))):
        n = len(P)
        dP = torch.from_numpy(np.ascontiguousarray(P.T)).cuda()
        dperm = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        _lib.check("hilbert_order", L.emsa_probe_hilbert_order(dP.data_ptr(), n, dperm.data_ptr()))
        np.testing.assert_array_equal(dperm.cpu().numpy().astype(np.int64), expected_order(P))
