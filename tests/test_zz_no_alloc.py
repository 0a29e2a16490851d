"""fgs_sort_pairs_noalloc (no workspace argument; its scratch is a static
device array of the library) gives the same result as fgs_sort_pairs with a
caller workspace, on two calls in a row, the second smaller: the scratch
keeps the first call's contents, which the second must not see."""
import ctypes

import numpy as np
import pytest
import torch

from fastergs_amd._lib import check, lib

pytestmark = pytest.mark.gpu


def _keys(n, bits, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 2 ** bits, n, dtype=np.uint64, endpoint=False)
    keys[rng.random(n) < 0.3] = keys[0]  # many equal keys: stability matters
    return keys


def _sort(fn_noalloc, keys, bits):
    n = keys.size
    k = torch.from_numpy(keys.view(np.int64)).cuda()
    v = torch.arange(n, dtype=torch.int32, device="cuda")
    ka, va = torch.empty_like(k), torch.empty_like(v)
    L = lib()
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if fn_noalloc:
        check(L.fgs_sort_pairs_noalloc(p(k), p(v), p(ka), p(va), n, 0, bits, s),
              "fgs_sort_pairs_noalloc")
    else:
        ws = torch.empty(L.fgs_sort_ws_size(n), dtype=torch.uint8, device="cuda")
        check(L.fgs_sort_pairs(p(k), p(v), p(ka), p(va), n, 0, bits, p(ws), ws.numel(), s),
              "fgs_sort_pairs")
        check(L.fgs_sort_check(p(ws), n, s), "fgs_sort_check")
    torch.cuda.synchronize()
    return k.cpu().numpy().view(np.uint64), v.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("n,bits,seed", [(1 << 24, 45, 1), (1_000_003, 13, 2)],
                         ids=["max_2^24", "smaller"])
def test_noalloc_sort_matches_workspace_sort(n, bits, seed):
    print(f"\n[noalloc] n={n} bits={bits}", flush=True)
    keys = _keys(n, bits, seed)
    k_ref, v_ref = _sort(False, keys, bits)
    k_got, v_got = _sort(True, keys, bits)
    np.testing.assert_array_equal(k_got, k_ref)
    np.testing.assert_array_equal(v_got, v_ref)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(v_got, order.astype(np.uint32))
