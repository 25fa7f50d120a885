"""pc_canonical_fused_noalloc, the entry point that allocates nothing (its
partial sums live in the library's fixed per-GPU scratch array): called
through the C ABI, its results equal pc_canonical_fused's with a caller
workspace bit for bit, on two calls in a row (the second smaller, so stale
scratch contents from the first would show), at both precisions; a problem
above the scratch size is refused before any launch."""
import ctypes

import numpy as np
import pytest

from pypercolate_amd import _engine as E
from pypercolate_amd import _native as nat

pytestmark = pytest.mark.gpu


def _inputs(torch, n, cols, num_p, seed):
    rng = np.random.RandomState(seed)
    x = rng.rand(n + 1, cols) * np.logspace(0, 12, cols)[None, :]
    ps = np.sort(rng.uniform(0.05, 0.95, num_p))
    start, length = E.binomial_windows(n, ps)
    dev = torch.device('cuda', 0)
    return {'x': torch.from_numpy(x).to(dev), 'ps': torch.from_numpy(ps).to(dev),
            'start': torch.from_numpy(start).to(dev),
            'len': torch.from_numpy(length).to(dev), 'n': n, 'cols': cols,
            'P': num_p}


def _with_workspace(torch, lib, a, precision):
    b = ctypes.c_uint64(0)
    nat.check(lib.pc_canonical_fused_workspace_size(a['P'], a['cols'], precision,
                                                    ctypes.byref(b)), 'size')
    ws = torch.empty((int(b.value),), dtype=torch.uint8, device=a['x'].device)
    out = torch.empty((a['P'], a['cols']), dtype=torch.float64, device=a['x'].device)
    nat.check(lib.pc_canonical_fused(
        a['n'], a['cols'], nat.ptr(a['x']), a['cols'], a['P'], nat.ptr(a['ps']),
        nat.ptr(a['start']), nat.ptr(a['len']), nat.ptr(out), precision,
        nat.ptr(ws), int(b.value), nat.stream_handle(torch)), 'pc_canonical_fused')
    return out


def _noalloc(torch, lib, a, precision):
    out = torch.empty((a['P'], a['cols']), dtype=torch.float64, device=a['x'].device)
    nat.check(lib.pc_canonical_fused_noalloc(
        a['n'], a['cols'], nat.ptr(a['x']), a['cols'], a['P'], nat.ptr(a['ps']),
        nat.ptr(a['start']), nat.ptr(a['len']), nat.ptr(out), precision,
        nat.stream_handle(torch)), 'pc_canonical_fused_noalloc')
    return out


@pytest.mark.parametrize('precision', [nat.PC_PREC_FP64_MFMA, nat.PC_PREC_SPLIT_BF16])
def test_noalloc_equals_workspace_form(gpu, precision):
    torch = gpu
    lib = nat.load_library()
    big = _inputs(torch, 32512, 21, 4096, 1)
    small = _inputs(torch, 480, 7, 101, 2)
    for a in (big, small):
        got = _noalloc(torch, lib, a, precision)
        ref = _with_workspace(torch, lib, a, precision)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (a['P'], a['cols'])
        assert torch.isfinite(got).all()


def test_noalloc_refuses_above_scratch(gpu):
    torch = gpu
    lib = nat.load_library()
    a = _inputs(torch, 480, 63, 4097, 3)
    b = ctypes.c_uint64(0)
    nat.check(lib.pc_canonical_fused_workspace_size(4097, 63, nat.PC_PREC_FP64_MFMA,
                                                    ctypes.byref(b)), 'size')
    assert b.value > nat.PC_NOALLOC_SCRATCH_BYTES
    with pytest.raises(MemoryError, match='scratch'):
        _noalloc(torch, lib, a, nat.PC_PREC_FP64_MFMA)
    # a problem at the limit still runs
    ok = _inputs(torch, 480, 63, 4096, 4)
    got = _noalloc(torch, lib, ok, nat.PC_PREC_FP64_MFMA)
    ref = _with_workspace(torch, lib, ok, nat.PC_PREC_FP64_MFMA)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
