"""mb_radix_sort_static (the entry point that allocates nothing) against
mb_radix_sort with a workspace, through the C ABI: two calls in a row, the
second smaller, with generated and with given values; and the size limit."""
import ctypes as C

import numpy as np
import pytest

from bevy_millions_ball_amd import _lib

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 20


def _u32(t):
    return t.cpu().numpy().view(np.uint32)


def _dev(a, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).to(device)


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def sort_both(device, keys, vals, passes):
    import torch

    L = _lib.lib()
    n = keys.shape[0]
    kd = _dev(keys, device)
    vd = None if vals is None else _dev(vals, device)
    ws = torch.zeros(L.mb_radix_sort_workspace_size(n), dtype=torch.uint8, device=device)
    ka, va = torch.empty_like(kd), torch.empty_like(kd)
    kb, vb = torch.empty_like(kd), torch.empty_like(kd)
    _lib.check(L.mb_radix_sort(_ptr(ws), _ptr(kd), _ptr(vd), _ptr(ka), _ptr(va), n, 0, passes, None))
    _lib.check(L.mb_radix_sort_static(_ptr(kd), _ptr(vd), _ptr(kb), _ptr(vb), n, 0, passes, None))
    torch.cuda.synchronize()
    return (_u32(ka), _u32(va)), (_u32(kb), _u32(vb))


@pytest.mark.parametrize("passes", [3, 4])
def test_static_sort_equals_workspace_sort(gpu, passes):
    rng = np.random.default_rng(passes)
    for n, with_vals in ((MAX_KEYS, False), (77_777, True), (MAX_KEYS - 5, True), (3000, False)):
        print(f"[no-alloc] n={n} passes={passes} vals={with_vals}", flush=True)
        keys = rng.integers(0, 1 << 24, n, dtype=np.uint64).astype(np.uint32)
        keys[rng.integers(0, n, n // 4)] = 7          # many equal keys: stability
        vals = rng.permutation(n).astype(np.uint32) if with_vals else None
        (ka, va), (kb, vb) = sort_both(gpu, keys, vals, passes)
        assert np.array_equal(ka, kb) and np.array_equal(va, vb)
        mask = np.uint32(0xFFFFFFFF if passes == 4 else (1 << (8 * passes)) - 1)
        order = np.argsort(keys & mask, kind="stable")
        assert np.array_equal(kb, keys[order])
        assert np.array_equal(vb, order.astype(np.uint32) if vals is None else vals[order])


def test_static_sort_rejects_larger_inputs(gpu):
    L = _lib.lib()
    rc = L.mb_radix_sort_static(C.c_void_p(16), None, C.c_void_p(32), C.c_void_p(48),
                                MAX_KEYS + 1, 0, 3, None)
    assert rc == _lib.MB_ERR_INVALID_ARGUMENT
    assert b"MB_RADIX_SORT_STATIC_MAX_KEYS" in L.mb_last_error()
