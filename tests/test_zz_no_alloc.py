"""scxr_vec_reduce_noalloc (static device workspace, no allocation) gives
bitwise the results of scxr_vec_reduce (caller workspace) for every
reduction kind, on two calls in a row with the second input smaller, so a
stale partial left in the shared workspace by the first call would show."""

import numpy as np
import pytest
import torch

from scico_amd import _lib

KINDS = {"dot": 0, "sumsq": 1, "sqdist": 2, "sum": 3, "l1": 4, "l21": 5}


def _reduce(lib, name, kind, n, K, x, y, post, stream, ws=None):
    out = torch.full((1,), np.nan, dtype=torch.float64, device=x.device)
    yp = y.data_ptr() if y is not None else None
    if ws is None:
        rc = lib.scxr_vec_reduce_noalloc(kind, n, K, x.data_ptr(), yp, 0.5, post,
                                         out.data_ptr(), None, stream)
    else:
        rc = lib.scxr_vec_reduce(kind, n, K, x.data_ptr(), yp, 0.5, post, out.data_ptr(),
                                 ws.data_ptr(), None, stream)
    _lib.check(name, rc)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("kind", list(KINDS))
def test_noalloc_reduce_matches_workspace_entry(gpu, kind):
    lib = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(7)
    K = 3 if kind == "l21" else 1
    # 2.5 M elements fill all 1024 first-pass workgroups; 5000 use 5, so the
    # second call's pass 2 must not pick up the first call's other partials.
    for n in (2_500_000, 5_000):
        xh = rng.standard_normal(K * n).astype(np.float32)
        yh = rng.standard_normal(n).astype(np.float32)
        x = torch.as_tensor(xh, device=gpu)
        y = torch.as_tensor(yh, device=gpu)
        yarg = y if kind in ("dot", "sqdist") else None
        ws = torch.empty(1024, dtype=torch.float64, device=gpu)
        print(f"noalloc reduce kind={kind} n={n}", flush=True)
        for post in (0, 1):
            if post == 1 and kind in ("dot", "sum"):
                continue  # the square root only of reductions that cannot be negative
            a = _reduce(lib, "scxr_vec_reduce_noalloc", KINDS[kind], n, K, x, yarg, post,
                        stream)
            b = _reduce(lib, "scxr_vec_reduce", KINDS[kind], n, K, x, yarg, post, stream, ws)
            torch.cuda.synchronize()
            av, bv = float(a.item()), float(b.item())
            assert av == bv, (kind, n, post, av, bv)
            # and the value itself, against a float64 host sum
            xd = xh.astype(np.float64)
            yd = yh.astype(np.float64)
            ref = {
                "dot": lambda: np.dot(xd, yd),
                "sumsq": lambda: np.dot(xd, xd),
                "sqdist": lambda: np.sum((xd - yd) ** 2),
                "sum": lambda: np.sum(xd),
                "l1": lambda: np.sum(np.abs(xd)),
                "l21": lambda: np.sum(np.sqrt(np.sum(xd.reshape(K, n) ** 2, axis=0))),
            }[kind]() * 0.5
            if post == 1:
                ref = np.sqrt(ref)
            assert abs(av - ref) <= 1e-9 * max(abs(ref), np.sqrt(n)), (kind, n, post, av, ref)


@pytest.mark.gpu
def test_noalloc_reduce_rejects_bad_arguments(gpu):
    lib = _lib.load()
    x = torch.zeros(16, dtype=torch.float32, device=gpu)
    out = torch.zeros(1, dtype=torch.float64, device=gpu)
    # a dot product without its second operand, an unknown kind, a null output
    assert lib.scxr_vec_reduce_noalloc(0, 16, 1, x.data_ptr(), None, 1.0, 0, out.data_ptr(),
                                       None, None) == -1
    assert lib.scxr_vec_reduce_noalloc(9, 16, 1, x.data_ptr(), None, 1.0, 0, out.data_ptr(),
                                       None, None) == -1
    assert lib.scxr_vec_reduce_noalloc(1, 16, 1, x.data_ptr(), None, 1.0, 0, None, None,
                                       None) == -1
    assert b"reduction" in lib.scxr_last_error()
