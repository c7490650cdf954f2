"""mpc_quantize_panel_noalloc (scratch in a static per-device array, no
device allocation) against mpc_quantize_panel (scratch in the handle's
workspace): bitwise equal payload and scales on two calls in a row, the
second smaller, through the C ABI; sizes beyond the static scratch are
rejected before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from cholesky_gpu_amd import _lib as L
from cholesky_gpu_amd import solver as S

pytestmark = pytest.mark.gpu

FORMATS = [L.PREC_E4M3, L.PREC_MXE5M2, L.PREC_BF16, L.PREC_FP32, L.PREC_MXE4M3, L.PREC_FP16,
           L.PREC_E5M2]


def say(*a):
    print("[case]", *a, flush=True)


def run(handle, fn, P, precs, tile_rows):
    """payload and scales from entry point fn (same buffers' shapes for both)."""
    k, m = P.shape
    pt = torch.as_tensor(precs, dtype=torch.uint8).cuda()
    pay = torch.full((m + 128, 4 * k), 0xAB, dtype=torch.uint8, device="cuda")
    sc = torch.full((m, k // 32), 0xCD, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    rc = fn(handle.h, C.c_void_p(P.data_ptr()), m, k, m, tile_rows, C.c_void_p(pt.data_ptr()),
            C.c_void_p(pay.data_ptr()), 4 * k, C.c_void_p(sc.data_ptr()))
    return rc, pay.cpu().numpy(), sc.cpu().numpy()


def test_noalloc_matches_quantize_panel(handle):
    lib = L.load()
    rng = np.random.default_rng(4)
    for m, k, tr in ((5 * 256 + 77, 256, 256), (3 * 128 + 5, 128, 128)):
        say("noalloc quantize", m, k, tr)
        Pm = (rng.standard_normal((m, k)) * np.exp(rng.uniform(-6, 6, (m, 1)))).astype(np.float32)
        P = torch.from_numpy(np.ascontiguousarray(Pm.T)).cuda()       # column-major m x k
        nt = (m + tr - 1) // tr
        precs = [FORMATS[t % len(FORMATS)] for t in range(nt)]
        rc0, pay0, sc0 = run(handle, lib.mpc_quantize_panel, P, precs, tr)
        rc1, pay1, sc1 = run(handle, lib.mpc_quantize_panel_noalloc, P, precs, tr)
        assert rc0 == 0 and rc1 == 0, (rc0, rc1)
        assert np.array_equal(pay0, pay1)
        assert np.array_equal(sc0, sc1)


def test_noalloc_rejects_beyond_scratch(handle):
    lib = L.load()
    say("noalloc size limit")
    tr = 32
    m = (L.NOALLOC_MAX_TILES + 1) * tr
    P = torch.zeros(32, m, dtype=torch.float32, device="cuda")
    rc, _, _ = run(handle, lib.mpc_quantize_panel_noalloc, P, [L.PREC_FP32] * (m // tr), tr)
    assert rc == L.ERR_ARG
    assert "static scratch" in lib.mpc_last_error_string().decode()
