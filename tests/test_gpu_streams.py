"""Unrolled (differentiable) registrations issued on two torch streams at once.

The unrolled forward and backward entry points keep one workspace per device
and return while their kernels are still queued on the caller's stream. A
call on another stream waits for the previous call's work before it rewrites
that workspace. Bar: gradients and results of two problems run on two streams
without synchronising in between are bit-identical to the same calls made one
after the other on one stream.
"""
import numpy as np
import pytest
import torch

from cuda_emicp_softassign_amd.clouds import range_scans, uniform_cube

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def UR():
    from cuda_emicp_softassign_amd import unrolled
    return unrolled


def _em_problem(n, m, seed, B):
    Xs, Ys = [], []
    for b in range(B):
        X, Y, _, _ = range_scans(n, m, seed=seed + b)
        Xs.append(X)
        Ys.append(Y)
    return np.stack(Xs).astype(np.float32), np.stack(Ys).astype(np.float32)


def _sa_problem(n, m, seed, B):
    Xs, Ys = [], []
    for b in range(B):
        X, Y, _, _ = uniform_cube(max(n, m), seed=seed + b)
        Xs.append(X[:n])
        Ys.append(Y[:m])
    return np.stack(Xs).astype(np.float32), np.stack(Ys).astype(np.float32)


def _fwd_bwd(run, X, Y):
    Xd = torch.tensor(X, device="cuda").requires_grad_(True)
    Yd = torch.tensor(Y, device="cuda").requires_grad_(True)
    Rt = run(Xd, Yd)
    (Rt.double() ** 2).sum().backward()
    return Rt, Xd, Yd


def _gather(res):
    Rt, Xd, Yd = res
    return [Rt.detach().cpu().numpy(), Xd.grad.cpu().numpy(), Yd.grad.cpu().numpy()]


def _two_streams(run, probs):
    torch.cuda.synchronize()
    seq = [_gather(_fwd_bwd(run, X, Y)) for X, Y in probs]
    torch.cuda.synchronize()
    main = torch.cuda.current_stream()
    streams = [torch.cuda.Stream() for _ in probs]
    out = []
    for s, (X, Y) in zip(streams, probs):
        s.wait_stream(main)
        with torch.cuda.stream(s):
            out.append(_fwd_bwd(run, X, Y))  # no synchronisation between the two
    torch.cuda.synchronize()
    for k, (a, b) in enumerate(zip(seq, [_gather(r) for r in out])):
        for name, x, y in zip(("Rt", "gX", "gY"), a, b):
            assert np.array_equal(x, y), (k, name, np.abs(x - y).max())


def test_emicp_unrolled_two_streams(UR):
    probs = [_em_problem(700, 650, 300, 2), _em_problem(1100, 900, 310, 3)]
    print("emicp unrolled on two streams: B=2 700x650, B=3 1100x900", flush=True)
    _two_streams(lambda X, Y: UR.emicp_unrolled(X, Y, sigma2_0=0.01, steps=12), probs)


def test_softassign_unrolled_two_streams(UR):
    probs = [_sa_problem(400, 380, 320, 2), _sa_problem(600, 520, 330, 1)]
    print("softassign unrolled on two streams: B=2 400x380, B=1 600x520", flush=True)
    _two_streams(lambda X, Y: UR.softassign_unrolled(X, Y, JMAX=6, I0=1, I1=3), probs)


def test_cached_sequence_replays_on_two_streams():
    """A recorded batched sequence (same buffers, sizes and parameters) owns
    one workspace: replayed on two streams back to back with no
    synchronisation in between, the second replay starts on the device after
    the first, and the outputs equal the eager run's bit for bit."""
    import cuda_emicp_softassign_amd as E
    pairs = []
    for k, (n, m) in enumerate([(900, 700), (1500, 1300), (400, 600)]):
        X, Y, _, _ = range_scans(n, m, seed=300 + k)
        pairs.append((X, Y))
    Rt0 = np.stack([E.multistart_rt0(X, Y, [np.eye(3)]) for X, Y in pairs]).astype(np.float32)
    p = E.registrationParameters(sigma_inf=2e-4)
    ref, ref_res = E.emicp_batched(pairs, Rt0, p)
    keep, dpairs = [], []
    for X, Y in pairs:
        dx = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
        dy = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()
        keep += [dx, dy]
        dpairs.append((dx.data_ptr(), len(X), dy.data_ptr(), len(Y)))
    d0 = torch.from_numpy(Rt0.reshape(-1)).cuda()
    dout = torch.zeros(12 * len(pairs), device="cuda")
    dres = torch.zeros(len(pairs), dtype=torch.float64, device="cuda")
    main = torch.cuda.current_stream()
    side = torch.cuda.Stream()
    for _ in range(2):  # eager, then recorded and replayed
        E.emicp_batched_dev(dpairs, d0.data_ptr(), 1, dout.data_ptr(), dres.data_ptr(), p,
                            main.cuda_stream)
        torch.cuda.synchronize()
    for rep in range(3):
        dout.fill_(float("nan"))
        dres.fill_(float("nan"))
        torch.cuda.synchronize()
        E.emicp_batched_dev(dpairs, d0.data_ptr(), 1, dout.data_ptr(), dres.data_ptr(), p,
                            main.cuda_stream)
        E.emicp_batched_dev(dpairs, d0.data_ptr(), 1, dout.data_ptr(), dres.data_ptr(), p,
                            side.cuda_stream)
        torch.cuda.synchronize()
        assert np.array_equal(dout.cpu().numpy().reshape(len(pairs), 1, 12), ref), rep
        assert np.array_equal(dres.cpu().numpy().reshape(len(pairs), 1), ref_res), rep
    E.graph_cache_clear()
