"""Application irregular functors run on the device through their compiled
tables (apps/moldyn, apps/minimd gpu_table.hip -> apps/bin/lib*_table.so):
every node is a partition of its own, its row of the CSR by destination
lists its edges in increasing order, and the map functor inserts only that
node's values (part_index[node] == part_id) into an object that copies the
first value and folds the rest with the reduce functor.

Bar: bit-exact against the oracle applying the same arithmetic serially over
the edge list with the same object semantics -- the reference functors' own
FP64 arithmetic (map 1 of each table) and the workload's FP32 force (map 0,
through the table's kernel and through the library kernel it is declared
equal to).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def moldyn6(pa):
    return pa.moldyn_system(6, seed=21)  # 864 molecules, shuffled pairs


def run_ir(pa, edges, pos64, table, map_idx, param, variant=None, order=None):
    ir = pa.IrregularRuntime(edges, len(pos64), map_idx, param, table=table)
    if variant is not None:
        ir.set_variant(variant)
    if order is not None:
        ir.set_row_order(order)
    ir.reset_node_data(pos64)
    ir.start()
    f = ir.get_reduction_result()
    ir.close()
    return f


@pytest.mark.parametrize("cells", [6, 16])
def test_moldyn_reference_functor_vs_oracle(pa, orc, gpu, cells):
    """moldyn's map in its own arithmetic (FP64, r148 d, minimum image)."""
    pos, box, edges = pa.moldyn_system(cells, seed=30 + cells)
    pos64 = pos.astype(np.float64) + np.random.default_rng(cells).uniform(-1e-3, 1e-3, pos.shape)
    pos64 %= box
    param = np.array([box, 2.5])
    got = run_ir(pa, edges, pos64, pa.app_table("moldyn", "ir"), 1, param)
    assert np.array_equal(got, orc.ir_lj_ref(pos64, edges, box, 2.5 * 2.5, 0))


def test_moldyn_workload_functor_table_kernel_and_library(pa, orc, gpu, moldyn6):
    """map 0 (the FP32 workload force): the table's own kernel equals the
    oracle's serial scatter over the list as given, and so does the library
    kernel it is declared equal to, with rows in edge order."""
    pos, box, edges = moldyn6
    pos64 = pos.astype(np.float64)
    param = np.array([box, 2.5])
    tab = pa.app_table("moldyn", "ir")
    ref, _ = orc.lj_force(pos, edges, box, 2.5)
    own = run_ir(pa, edges, pos64, tab, 0, param, variant=pa.PAT_IR_TABLE_KERNEL)
    assert np.array_equal(own, ref)
    lib_ = run_ir(pa, edges, pos64, tab, 0, param, order=pa.PAT_IR_ORDER_EDGE)
    assert np.array_equal(lib_, ref)


def test_minimd_functors_vs_oracle(pa, orc, gpu):
    """miniMD's map1 (no images) on 8^3 cells with the half list within 2.8:
    the reference arithmetic (map 1) and the workload force (map 0, table
    kernel)."""
    pos, box = pa.gen_fcc(8, 8, 8, rho=0.8442, jitter=0.05, seed=41)
    edges = orc.pairs_brute(pos, box[0], 2.8, pbc=False)
    edges = edges[np.random.default_rng(42).permutation(len(edges))]
    pos64 = pos.astype(np.float64)
    tab = pa.app_table("minimd", "ir")
    param = np.array([2.5 * 2.5])
    got = run_ir(pa, edges, pos64, tab, 1, param)
    assert np.array_equal(got, orc.ir_lj_ref(pos64, edges, box[0], 6.25, 1))
    got0 = run_ir(pa, edges, pos64, tab, 0, param, variant=pa.PAT_IR_TABLE_KERNEL)
    ref0, _ = orc.lj_force(pos, edges, box[0], 2.5, pbc=False)
    assert np.array_equal(got0, ref0)


def test_table_index_errors(pa, gpu, moldyn6):
    pos, box, edges = moldyn6
    tab = pa.app_table("moldyn", "ir")
    ir = pa.IrregularRuntime(edges, len(pos), 0, np.array([box, 2.5]), table=tab)
    with pytest.raises(pa.PatternError, match="not in the application's table"):
        pa.call("pat_ir_set_map_idx", ir.h, 2)
    with pytest.raises(pa.PatternError, match="reduce index"):
        pa.call("pat_ir_set_reduce_idx", ir.h, 1)
    with pytest.raises(pa.PatternError, match="no node data"):
        ir.start()
    ir.close()
