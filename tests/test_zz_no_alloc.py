"""lp_label26_noalloc (device pointers, static working memory) against
lp_label26 (host arrays, allocating): the same labels and count, on two
calls in a row, the second on a smaller grid, and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _label_host(lib, mask):
    nx, ny, nz = mask.shape
    m = np.ascontiguousarray(mask, np.uint8)
    lab = np.empty(mask.shape, np.int32)
    cnt = C.c_int()
    assert lib.lp_label26(nx, ny, nz, m.ctypes.data, lab.ctypes.data, C.byref(cnt)) == 0
    return lab, cnt.value


def _label_noalloc(lib, mask):
    import torch

    nx, ny, nz = mask.shape
    dm = torch.from_numpy(np.ascontiguousarray(mask, np.uint8)).cuda()
    dl = torch.full(mask.shape, 12345, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    cnt = C.c_int()
    rc = lib.lp_label26_noalloc(nx, ny, nz, dm.data_ptr(), dl.data_ptr(), C.byref(cnt))
    assert rc == 0, rc
    return dl.cpu().numpy(), cnt.value


def test_noalloc_labels_equal_the_allocating_entry(gpu):
    from loadpath_amd import _lib

    lib = _lib.lib()
    rng = np.random.default_rng(11)
    # first call: a large random mask (many components); second: smaller
    for shape, p in (((40, 36, 28), 0.28), ((9, 7, 5), 0.35), ((1, 1, 1), 1.0), ((12, 1, 3), 0.0)):
        mask = rng.random(shape) < p
        want, nw = _label_host(lib, mask)
        got, ng = _label_noalloc(lib, mask)
        assert ng == nw, (shape, ng, nw)
        assert np.array_equal(got, want), shape
    # a spiral-like single component on a full slab and an all-ones grid
    full = np.ones((16, 16, 16), bool)
    got, ng = _label_noalloc(lib, full)
    assert ng == 1 and np.all(got == 1)


def test_noalloc_rejects_grids_over_the_limit(gpu):
    import torch

    from loadpath_amd import _lib

    lib = _lib.lib()
    dm = torch.zeros(8, dtype=torch.uint8, device="cuda")
    dl = torch.zeros(8, dtype=torch.int32, device="cuda")
    cnt = C.c_int(-7)
    limit = 1 << 26
    rc = lib.lp_label26_noalloc(limit // 64 + 1, 8, 8, dm.data_ptr(), dl.data_ptr(), C.byref(cnt))
    assert rc == _lib.LP_ERR_ARG
    assert "limit" in lib.lp_last_error().decode()
    assert cnt.value == -7
