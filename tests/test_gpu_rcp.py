"""The reverse walk's reciprocal of (1 - alpha) equals the IEEE quotient for
every float in [2^-7, 1] (the walk only sees [0.01, 1]), checked
exhaustively on the device: this is what lets the GPU backward recover the
same transmittance as the oracle's division."""
import ctypes

import pytest
import torch

from fastergs_amd._lib import check, lib

pytestmark = pytest.mark.gpu


def test_backward_reciprocal_is_correctly_rounded():
    bad = torch.full((2,), 12345, dtype=torch.int64, device="cuda")
    check(lib().fgs_rcp_selftest(ctypes.c_void_p(bad.data_ptr()),
                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
          "fgs_rcp_selftest")
    torch.cuda.synchronize()
    assert int(bad[0].item()) == 0
