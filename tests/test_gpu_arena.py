"""hipMalloc VRAM arena through the C ABI: the reference's arena cases
(tests/unit/test_vram_arena.cu:25-152): 256-byte granules, first fit, coalescing of free
neighbours, occupancy in whole percent, bad_alloc on exhaustion, invalid_argument on a double
free, free(NULL) a no-op, and device writes / reads through the returned pointers."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_construction(gpu):
    a = gpu.Arena(1024 * 1024)
    assert a.occupancy_percentage() == 0
    a.close()


def test_allocate_write_read_free_reuse(gpu):
    a = gpu.Arena(16 * 1024)
    p = a.allocate(1024)
    assert p != 0 and a.occupancy_percentage() > 0
    L = gpu.lib()
    L.lkd_copy_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    src = gpu.DeviceBuffer(np.arange(128, dtype=np.float64) + 42.0)
    assert L.lkd_copy_f64(C.c_void_p(p), src.ptr, 128, None) == 0     # a kernel writes the block
    assert L.lkd_stream_sync(None) == 0
    out = np.zeros(128)
    assert L.lkd_memcpy_dtoh(out.ctypes.data, C.c_void_p(p), out.nbytes, None) == 0
    np.testing.assert_array_equal(out, np.arange(128) + 42.0)
    a.free(p)
    assert a.occupancy_percentage() == 0
    p2 = a.allocate(1024)
    assert p2 != 0
    a.free(p2)
    assert a.occupancy_percentage() == 0
    a.close()


def test_granules_and_coalescing(gpu):
    a = gpu.Arena(1024)                     # four 256-byte granules
    pa, pb, pc = a.allocate(200), a.allocate(200), a.allocate(200)
    assert len({pa, pb, pc}) == 3 and 0 not in (pa, pb, pc)
    assert a.occupancy_percentage() == 75
    a.free(pb)
    a.free(pc)
    assert a.occupancy_percentage() == 25
    big = a.allocate(500)                   # needs the two freed granules plus the tail
    assert big != 0 and a.occupancy_percentage() == 75
    a.free(pa)
    a.free(big)
    assert a.occupancy_percentage() == 0
    a.close()


def test_edge_cases(gpu):
    a = gpu.Arena(1024 * 1024)
    assert a.allocate(0) == 0 and a.occupancy_percentage() == 0
    with pytest.raises(MemoryError):
        a.allocate(2 * 1024 * 1024)
    assert a.occupancy_percentage() == 0
    p = a.allocate(100)
    a.free(p)
    with pytest.raises(ValueError):
        a.free(p)
    a.free(None)
    a.close()


def test_occupancy_tracking(gpu):
    a = gpu.Arena(1024)
    ps = []
    for want in (25, 50, 75, 100):
        ps.append(a.allocate(100))
        assert a.occupancy_percentage() == want
    a.free(ps[1])
    assert a.occupancy_percentage() == 75
    for p in (ps[0], ps[2], ps[3]):
        a.free(p)
    assert a.occupancy_percentage() == 0
    a.close()
