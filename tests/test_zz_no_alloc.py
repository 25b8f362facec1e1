"""vkgs_sort_kv_noalloc (static device scratch) against vkgs_sort_kv (caller storage), through the C ABI.

Two calls in a row on the same GPU, the second with a smaller input, must each give the keys and
values vkgs_sort_kv gives for the same input, bit for bit; a capacity above the static buffer's
maximum is refused.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from vkgs_amd import _lib  # noqa: E402

pytestmark = pytest.mark.gpu

MAX_N = 1 << 23  # VKGS_SORT_NOALLOC_MAX_N


def _inputs(n, cap, bits, distinct, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, min(distinct, 1 << bits), cap, dtype=np.uint64).astype(np.uint32)
    vals = rng.permutation(cap).astype(np.uint32)
    return keys, vals


def _dev(a):
    return torch.from_numpy(a.view(np.int32).copy()).cuda()


def _host(t):
    return t.cpu().numpy().view(np.uint32)


def _sort_with_storage(cap, n, keys, vals, bits):
    lib = _lib.lib()
    dk, dv = _dev(keys), _dev(vals)
    cnt = torch.tensor([n], dtype=torch.int32, device="cuda")
    nbytes = lib.vkgs_sort_kv_storage_bytes(cap)
    storage = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.vkgs_sort_kv(cap, cnt.data_ptr(), dk.data_ptr(), dv.data_ptr(), bits, storage.data_ptr(),
                                nbytes, stream), "vkgs_sort_kv")
    torch.cuda.synchronize()
    return _host(dk), _host(dv)


def _sort_noalloc(cap, n, keys, vals, bits):
    lib = _lib.lib()
    dk, dv = _dev(keys), _dev(vals)
    cnt = torch.tensor([n], dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.vkgs_sort_kv_noalloc(cap, cnt.data_ptr(), dk.data_ptr(), dv.data_ptr(), bits, stream),
               "vkgs_sort_kv_noalloc")
    torch.cuda.synchronize()
    return _host(dk), _host(dv)


def test_noalloc_sort_matches_storage_sort_on_two_calls():
    # first call: near the maximum, 32-bit keys with ties, live count below capacity
    cases = [(MAX_N, MAX_N - 1234, 32, 1 << 20, 1),
             # second call, smaller: 30-bit keys (the depth-key width), heavy ties
             (100_003, 99_000, 30, 97, 2)]
    for cap, n, bits, distinct, seed in cases:
        print("noalloc sort cap", cap, "n", n, "bits", bits, flush=True)
        keys, vals = _inputs(n, cap, bits, distinct, seed)
        want_k, want_v = _sort_with_storage(cap, n, keys, vals, bits)
        got_k, got_v = _sort_noalloc(cap, n, keys, vals, bits)
        np.testing.assert_array_equal(got_k, want_k)
        np.testing.assert_array_equal(got_v, want_v)
        order = np.argsort(keys[:n], kind="stable")  # and both are the stable sort
        np.testing.assert_array_equal(got_k[:n], keys[:n][order])
        np.testing.assert_array_equal(got_v[:n], vals[:n][order])
        np.testing.assert_array_equal(got_k[n:], keys[n:])  # untouched tail


def test_noalloc_sort_refuses_capacity_above_its_buffer():
    lib = _lib.lib()
    k = torch.zeros(4, dtype=torch.int32, device="cuda")
    cnt = torch.tensor([1], dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    st = lib.vkgs_sort_kv_noalloc(MAX_N + 1, cnt.data_ptr(), k.data_ptr(), k.data_ptr(), 32, stream)
    assert st == 2  # VKGS_ERR_CAPACITY, before any launch
    st = lib.vkgs_sort_kv_noalloc(4, cnt.data_ptr(), k.data_ptr(), k.data_ptr(), 0, stream)
    assert st == 1  # VKGS_ERR_INVALID_ARG
    assert lib.vkgs_sort_kv_noalloc(0, None, None, None, 32, stream) == 0  # empty sort: nothing to do
    assert C.sizeof(C.c_void_p) == 8
