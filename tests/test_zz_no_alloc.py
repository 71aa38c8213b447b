"""fdm_buildHashStartEndNoAlloc (the hash-table build whose sort workspace is
the library's fixed device array) against fdm_buildHashStartEnd (caller's
workspace) through the C-ABI: sorted keys, body order, cell ranges and the
reset bookkeeping are bit-identical on two calls in a row, the second with a
smaller input, and equal the stable order of the keys computed on the host.
Sizes beyond FDM_NOALLOC_MAX_BODIES are refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from fundem_amd import build

pytestmark = pytest.mark.gpu

FDM_ECAPACITY = -3
MAX_BODIES = 1 << 21


class Double3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class Int3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class Grid(C.Structure):
    _fields_ = [("min", Double3), ("max", Double3), ("invCell", Double3), ("size", Int3),
                ("num", C.c_longlong)]


class HashTable(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in ("hashValue", "sortKey", "sortIndex", "cellStart",
                                                "cellEnd", "prevKey", "prevCount",
                                                "validAtBuild")]


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "needs a GPU"
    lib = C.CDLL(build.CORE_LIB)
    lib.fdm_radixSortWorkspaceBytes.restype = C.c_size_t
    lib.fdm_radixSortWorkspaceBytes.argtypes = [C.c_size_t]
    lib.fdm_buildHashStartEnd.argtypes = [HashTable, C.c_int, Grid, C.c_void_p, C.c_size_t,
                                          C.c_void_p]
    lib.fdm_buildHashStartEndNoAlloc.argtypes = [HashTable, C.c_int, Grid, C.c_void_p]
    return lib


class Table:
    """Device buffers of one hash table, sized for `cap` bodies and `cells` cells."""

    def __init__(self, cap, cells):
        import torch
        dev = torch.device("cuda:0")
        self.key = torch.zeros(cap, dtype=torch.int32, device=dev)
        self.index = torch.zeros(cap, dtype=torch.int32, device=dev)
        self.start = torch.full((cells,), -7, dtype=torch.int32, device=dev)
        self.end = torch.full((cells,), -7, dtype=torch.int32, device=dev)
        self.prev_key = torch.zeros(cap, dtype=torch.int32, device=dev)
        self.prev_count = torch.full((1,), -1, dtype=torch.int32, device=dev)

    def load(self, keys):
        import torch
        n = keys.size
        self.key[:n] = torch.from_numpy(keys.astype(np.int32)).to(self.key.device)
        self.index[:n] = torch.arange(n, dtype=torch.int32, device=self.key.device)

    def view(self):
        return HashTable(None, self.key.data_ptr(), self.index.data_ptr(), self.start.data_ptr(),
                         self.end.data_ptr(), self.prev_key.data_ptr(), self.prev_count.data_ptr(),
                         None)

    def host(self, n):
        return {k: getattr(self, a)[:m].cpu().numpy().copy() for k, a, m in
                (("key", "key", n), ("index", "index", n), ("start", "start", None),
                 ("end", "end", None), ("prev_key", "prev_key", n),
                 ("prev_count", "prev_count", None))}


def grid_of(cells):
    g = Grid()
    g.size = Int3(cells, 1, 1)
    g.num = cells
    return g


def expected(keys, cells):
    order = np.argsort(keys, kind="stable")
    sk = keys[order]
    start = np.full(cells, -1, np.int32)
    end = np.full(cells, -1, np.int32)
    for i, k in enumerate(sk):
        h = int(k) - 1
        if 0 <= h < cells:
            if i == 0 or sk[i - 1] != k:
                start[h] = i
            if i == sk.size - 1 or sk[i + 1] != k:
                end[h] = i + 1
    return sk, order.astype(np.int32), start, end


def test_no_alloc_hash_build_equals_workspace_build(lib):
    import torch
    cells = 6000
    sizes = (150_000, 23_457)  # the second call is smaller
    rng = np.random.default_rng(20261203)
    print(f"no-alloc hash build: cells {cells}, sizes {sizes}", flush=True)
    ref, test = Table(sizes[0], cells), Table(sizes[0], cells)
    stream = torch.cuda.current_stream().cuda_stream
    g = grid_of(cells)
    for n in sizes:
        # keys hash + 1: 0 (outside the grid), inside, and rounded up to the grid size
        keys = rng.integers(0, cells + 1, size=n).astype(np.int64)
        keys[rng.random(n) < 0.05] = 0
        keys[rng.random(n) < 0.01] = cells + 1
        ref.load(keys)
        test.load(keys)
        ws_bytes = lib.fdm_radixSortWorkspaceBytes(n)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda:0")
        torch.cuda.synchronize()
        assert lib.fdm_buildHashStartEnd(ref.view(), n, g, ws.data_ptr(), ws_bytes, stream) == 0
        assert lib.fdm_buildHashStartEndNoAlloc(test.view(), n, g, stream) == 0
        torch.cuda.synchronize()
        a, b = ref.host(n), test.host(n)
        for k in a:
            assert np.array_equal(a[k], b[k]), (n, k)
        sk, order, start, end = expected(keys, cells)
        assert np.array_equal(b["key"], sk) and np.array_equal(b["index"], order), n
        assert np.array_equal(b["start"], start) and np.array_equal(b["end"], end), n
        assert b["prev_count"][0] == n and np.array_equal(b["prev_key"], sk), n
        del ws


def test_no_alloc_refuses_sizes_beyond_its_scratch(lib):
    t = Table(16, 8)
    assert lib.fdm_buildHashStartEndNoAlloc(t.view(), MAX_BODIES + 1, grid_of(8), None) == \
        FDM_ECAPACITY
