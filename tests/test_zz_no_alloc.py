"""pat_ir_csr_build_dev_noalloc: the CSR build with its work space in a
static device array instead of caller or library allocations.  Through the C
ABI, its rowptr and cols equal pat_ir_csr_build_dev's bit for bit, in both
row orders, on two calls in a row (the second on a smaller edge list, so
stale counters in the shared array would show), and a build beyond the
array's capacity is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def build_noalloc(pa, edges_dev, ne, n, order):
    rowptr = pa.DeviceArray((n + 1,), np.int32)
    cols = pa.DeviceArray((max(2 * ne, 1),), np.int32)
    pa.call("pat_ir_csr_build_dev_noalloc", pa.ptr(edges_dev), ne, n, order, rowptr.ptr, cols.ptr, None)
    pa.sync()
    r, c = rowptr.numpy(), cols.numpy()[:2 * ne]
    rowptr.free()
    cols.free()
    return r, c


@pytest.mark.parametrize("order", [0, 1])
def test_noalloc_csr_equals_workspace_build(pa, gpu, order):
    for cells, seed in ((12, 3), (7, 4)):  # the second call is smaller
        pos, box, edges = pa.moldyn_system(cells, seed=seed)
        n, ne = len(pos), len(edges)
        ed = pa.DeviceArray.from_numpy(edges)
        rp, cl = pa.csr_build_dev(ed, ne, n, order=order)
        want = (rp.numpy(), cl.numpy()[:2 * ne])
        rp.free()
        cl.free()
        got = build_noalloc(pa, ed, ne, n, order)
        ed.free()
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), (cells, order)
        assert got[0][-1] == 2 * ne


def test_noalloc_refuses_beyond_capacity(pa, gpu):
    n, ne = 1000, 9_000_000  # ~144 MB of work space: beyond the 128 MiB array
    rowptr = pa.DeviceArray((n + 1,), np.int32)
    with pytest.raises(pa.PatternError, match="static scratch holds"):
        pa.call("pat_ir_csr_build_dev_noalloc", rowptr.ptr, ne, n, 0, rowptr.ptr, rowptr.ptr, None)
    rowptr.free()
