"""GPU parity of the irregular engine (through libpattern's C ABI) vs the oracle.

Bar: the in-order CSR kernels (the default, on the plain or the sliced
layout) are bit-identical to the FP32 oracle's serial scatter -- forces, and
whole velocity-Verlet trajectories -- because each row lists a node's edges
in the row order and sums them in that order with the same per-pair
arithmetic.  Rows in edge order match the scatter over the list as given;
rows in node order (the handle default) match the scatter over the list
sorted by (min(i, j), max(i, j)).  The atomic and tree-order kernels sum in
another order and are held to 1e-5 relative L-infinity (max |a-b| / max |b|)
per force evaluation.  The CSR must equal a NumPy construction exactly.
Errors against the FP64 oracle are printed.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / max(np.max(np.abs(b)), 1e-30))


def node_sorted(edges):
    """The edge list in the order whose serial scatter adds each row's
    neighbours by increasing index: sorted by (min(i, j), max(i, j))."""
    lo, hi = np.minimum(edges[:, 0], edges[:, 1]), np.maximum(edges[:, 0], edges[:, 1])
    return edges[np.lexsort((hi, lo))]


def oracle_edges(pa, edges, order):
    return node_sorted(edges) if order == pa.PAT_IR_ORDER_NODE else edges


@pytest.fixture(scope="module")
def small(pa):
    return pa.moldyn_system(6, seed=21)  # 864 atoms


def test_csr_build_exact(pa, gpu, small):
    pos, box, edges = small
    n = len(pos)
    ed = pa.DeviceArray.from_numpy(edges)
    rowptr, cols = pa.csr_build_dev(ed, len(edges), n)
    rp, cl = rowptr.numpy(), cols.numpy()
    # row r: r's neighbours in the order their edges appear in the list
    both = np.concatenate([edges, edges[:, ::-1]])
    eid = np.concatenate([np.arange(len(edges)), np.arange(len(edges))])
    order = np.lexsort((eid, both[:, 0]))
    both = both[order]
    expect_rp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(expect_rp, both[:, 0] + 1, 1)
    assert np.array_equal(rp, np.cumsum(expect_rp))
    assert np.array_equal(cl, both[:, 1])


def test_csr_build_node_order(pa, gpu, small):
    pos, box, edges = small
    n = len(pos)
    rowptr, cols = pa.csr_build_dev(pa.DeviceArray.from_numpy(edges), len(edges), n, order=pa.PAT_IR_ORDER_NODE)
    both = np.concatenate([edges, edges[:, ::-1]])
    both = both[np.lexsort((both[:, 1], both[:, 0]))]
    assert np.array_equal(cols.numpy(), both[:, 1])
    # the sorted list's scatter visits each row in exactly this order
    s = node_sorted(edges)
    eid = np.concatenate([np.arange(len(s)), np.arange(len(s))])
    b2 = np.concatenate([s, s[:, ::-1]])
    b2 = b2[np.lexsort((eid, b2[:, 0]))]
    assert np.array_equal(cols.numpy(), b2[:, 1])


def test_csr_build_edge_cases(pa, gpu):
    # isolated nodes, a hub of high degree (> 64, multi-chunk rows), duplicates
    n = 300
    hub = [(0, j) for j in range(1, 200)]
    edges = np.array(hub + [(5, 7), (5, 7), (299, 3)], dtype=np.int32)
    rowptr, cols = pa.csr_build_dev(pa.DeviceArray.from_numpy(edges), len(edges), n)
    rp, cl = rowptr.numpy(), cols.numpy()
    assert rp[0] == 0 and rp[-1] == 2 * len(edges)
    assert np.array_equal(cl[rp[0]:rp[1]], np.arange(1, 200))
    assert np.array_equal(cl[rp[5]:rp[6]], [0, 7, 7])  # edge order: the hub's edge, then the duplicates
    assert rp[251] - rp[250] == 0


@pytest.mark.parametrize("lpr", [1, 4, 8, 16, 32, -4, -8, -16, -32, -64])
def test_force_csr_vs_oracle(pa, orc, gpu, small, lpr):
    pos, box, edges = small
    n = len(pos)
    rowptr, cols = pa.csr_build_dev(pa.DeviceArray.from_numpy(edges), len(edges), n)
    p4 = pa.DeviceArray.from_numpy(pa.to_float4(pos))
    f4 = pa.DeviceArray((n, 4), np.float32)
    pa.call("pat_lj_force_csr_dev", p4.ptr, rowptr.ptr, cols.ptr, n, box, 1, 2.5, lpr, f4.ptr, None)
    pa.sync()
    got = f4.numpy()[:, :3]
    ref, _ = orc.lj_force(pos, edges, box, 2.5)
    if lpr > 0:
        assert np.array_equal(got, ref)  # in order: the serial scatter's sums
    else:
        assert rel(got, ref) < TOL


def test_force_atomic_vs_oracle(pa, orc, gpu, small):
    pos, box, edges = small
    n = len(pos)
    p4 = pa.DeviceArray.from_numpy(pa.to_float4(pos))
    f4 = pa.DeviceArray((n, 4), np.float32)
    ed = pa.DeviceArray.from_numpy(edges)
    pa.call("pat_lj_force_atomic_dev", p4.ptr, ed.ptr, len(edges), n, box, 1, 2.5, f4.ptr, None)
    pa.sync()
    ref, _ = orc.lj_force(pos, edges, box, 2.5)
    assert rel(f4.numpy()[:, :3], ref) < TOL


def test_two_particle_and_cutoff(pa, gpu):
    pos = np.array([[5.0, 5, 5], [6.1, 5, 5], [5.0, 7.6, 5]], dtype=np.float32)
    edges = np.array([[0, 1], [0, 2]], dtype=np.int32)
    ir = pa.IrregularRuntime(edges, 3, pa.PAT_IR_LJ_OPEN, np.array([2.5 * 2.5]))
    ir.reset_node_data(pos)
    ir.start()
    f = ir.get_reduction_result()
    r = float(pos[1, 0]) - float(pos[0, 0])  # the distance the FP32 inputs encode
    F = 48 * (r ** -13 - 0.5 * r ** -7)
    assert abs(f[0, 0] + F) < 2e-6 * abs(F) and abs(f[1, 0] - F) < 2e-6 * abs(F)
    assert f[2].tolist() == [0, 0, 0]  # 2.6 apart: beyond the cutoff
    ir.close()


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("variant,lanes", [(0, 0), (0, 1), (0, 64), (1, 0), (2, 0)])
def test_irregular_runtime_handle(pa, orc, gpu, small, variant, lanes, order):
    pos, box, edges = small
    para = np.array([box, 2.5], dtype=np.float64)  # struct para {side, cutoffRadius}
    ir = pa.IrregularRuntime(edges, len(pos), pa.PAT_IR_LJ_MINIMAGE, para)
    ir.set_row_order(order)
    ir.set_variant(variant, lanes)
    ir.reset_node_data(pos.astype(np.float64))  # moldyn's double3 point_data
    ir.start()
    got = ir.get_reduction_result()
    ref, _ = orc.lj_force(pos, oracle_edges(pa, edges, order), box, 2.5)
    if variant == 0:
        assert np.array_equal(got, ref)
    else:
        assert rel(got, ref) < TOL
    ir.close()


def test_sliced_layout_refused_when_padding_is_large(pa, gpu):
    # a hub of degree 199 pads its whole 64-row slice to 199 entries
    n = 300
    edges = np.array([(0, j) for j in range(1, 200)] + [(250, 251)], dtype=np.int32)
    pos = np.random.default_rng(3).uniform(0, 20, (n, 3)).astype(np.float32)
    ir = pa.IrregularRuntime(edges, n, pa.PAT_IR_LJ_OPEN, np.array([6.25]))
    with pytest.raises(RuntimeError, match="sliced layout"):
        ir.set_variant(pa.PAT_IR_CSR, pa.PAT_IR_SLICED)
    ir.set_variant(pa.PAT_IR_CSR, 1)
    ir.reset_node_data(pos)
    ir.start()
    ir.close()


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("lanes", [1, 4, 8, 16, 32, 64])
def test_lanes_per_row_do_not_change_results(pa, gpu, small, lanes, order):
    pos, box, edges = small
    outs = []
    for l in (1, lanes):
        md = pa.MD(pos, None, edges, box, 2.5, 0.005, lanes_per_row=l, order=order)
        md.run(5)
        outs.append(md.get())
        md.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("variant,order", [(0, 1), (0, 0), (1, 1), (2, 1), (2, 0)])
def test_md_small_vs_oracle(pa, orc, gpu, small, variant, order):
    pos, box, edges = small
    vel = pa.gen_velocities(len(pos), 1.44, seed=5)
    md = pa.MD(pos, vel, edges, box, 2.5, 0.005, variant=variant, order=order)
    md.run(10)
    md.run(10)
    p, v, f = md.get()
    rp, rv, rf, rep = orc.md_run(pos, vel, oracle_edges(pa, edges, order), box, 2.5, 0.005, 20)
    if variant == 0:
        assert np.array_equal(f, rf) and np.array_equal(v, rv) and np.array_equal(p, rp)
        ep, ek = md.energy()
        assert abs(ep - rep) / abs(rep) < 1e-9
    else:
        assert rel(f, rf) < TOL and rel(v, rv) < TOL
        d = p - rp
        d -= box * np.rint(d / box)
        assert np.max(np.abs(d)) / box < TOL
    md.close()


def test_graph_timed_run_matches_run(pa, orc, gpu, small):
    """The timed graph replay advances the handle like run(): 7 + 7 + 6
    steps (odd and even step counts, so both start buffers are replayed),
    bit-identical to the oracle; each call reports a positive device time.
    A handle with neighbour rebuilds refuses it."""
    pos, box, edges = small
    vel = pa.gen_velocities(len(pos), 1.44, seed=6)
    md = pa.MD(pos, vel, edges, box, 2.5, 0.005)
    ms = [md.run_graph_timed(7), md.run_graph_timed(7), md.run_graph_timed(6)]
    p, v, f = md.get()
    assert md.stats()[1] == 20
    md.close()
    assert all(t > 0 for t in ms)
    rp, rv, rf, _ = orc.md_run(pos, vel, node_sorted(edges), box, 2.5, 0.005, 20)
    assert np.array_equal(p, rp) and np.array_equal(v, rv) and np.array_equal(f, rf)
    fp, fbox = pa.gen_fcc(6, 6, 6, rho=0.8442, jitter=0.0, seed=3)
    mb = pa.MD(fp, None, None, fbox[0], 2.5, 0.005, skin=0.3, every=20)
    with pytest.raises(pa.PatternError):
        mb.run_graph_timed(5)
    mb.close()


@pytest.mark.parametrize("order", [1, 0])
def test_moldyn_workload(pa, orc, gpu, order):
    """moldyn: 40^3 FCC cells (256,000 molecules), density 0.8442, jitter
    +-0.05, half list within 2.5 (random order), 50 VV steps, dt 0.005."""
    pos, box, edges = pa.moldyn_system(40, seed=42)
    assert 6.5e6 < len(edges) < 7.5e6
    md = pa.MD(pos, None, edges, box, 2.5, 0.005, order=order)
    md.run(50)
    p, v, f = md.get()
    ep, ek = md.energy()
    md.close()
    oe = oracle_edges(pa, edges, order)
    rp, rv, rf, _ = orc.md_run(pos, np.zeros_like(pos), oe, box, 2.5, 0.005, 50, threads=16)
    assert np.array_equal(p, rp) and np.array_equal(v, rv) and np.array_equal(f, rf)
    _, rep = orc.lj_force(rp, oe, box, 2.5)
    p64, v64, f64, _ = orc.md_run(pos.astype(np.float64), np.zeros((len(pos), 3)), edges, box, 2.5, 0.005, 50,
                                  dtype=np.float64)
    print(f"moldyn 256k x50: pairs={len(edges)} bit-exact vs FP32 oracle; rel err vs FP64: force "
          f"{rel(f, f64):.2e} vel {rel(v, v64):.2e}; epot {ep:.9e} (oracle {rep:.9e})")
    assert abs(ep - rep) / abs(rep) < 1e-9
    # size-independent properties: momentum stays ~0 (forces sum to ~0)
    assert np.max(np.abs(f.astype(np.float64).sum(axis=0))) / np.max(np.abs(f)) < 1e-3


@pytest.mark.parametrize("cells", [6, 16])
@pytest.mark.parametrize("form", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_sliced_forms_vs_oracle(pa, orc, gpu, cells, form, order):
    """Each loop form on the handle's sentinel layout (byte-offset columns, the
    minimum image taken per wavefront only on axes where a row lies within a
    cutoff of a face; forms 3 and 4 stage each block's neighbour positions in
    LDS), with the slices' rows in node, row-class or spatial order, gives
    the serial scatter's bits over a trajectory."""
    pos, box, edges = pa.moldyn_system(cells, seed=33)
    vel = pa.gen_velocities(len(pos), 1.44, seed=5)
    md = pa.MD(pos, vel, edges, box, 2.5, 0.005)
    md.set_sliced_form(form)
    md.set_slice_order(order)
    assert md.variant()[1] == pa.PAT_IR_SLICED and md.slice_order() == order
    md.run(7)
    md.run(8)
    p, v, f = md.get()
    md.close()
    rp, rv, rf, _ = orc.md_run(pos, vel, node_sorted(edges), box, 2.5, 0.005, 15)
    assert np.array_equal(f, rf) and np.array_equal(v, rv) and np.array_equal(p, rp)


@pytest.mark.parametrize("form", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("order", [0, 2])
def test_minimum_image_kept_for_positions_outside_the_box(pa, orc, gpu, form, order):
    """Rows 0..63 (one wavefront) lie well inside the box; node 64 starts at
    x = -2, outside it.  Pair (0, 64) is 9 apart in x unwrapped and 1 apart
    through the periodic image: its force is only right if the handle sees
    that not every position lies in [0, L] and keeps the minimum image for
    row 0's wavefront.  After the first drift node 64 is wrapped to x ~ 8,
    inside, and later runs may skip the image -- same bits throughout."""
    L, n = 10.0, 128
    pos = np.zeros((n, 3), np.float32)
    pos[:64] = np.stack([np.full(64, 7.0), 2.6 + 0.07 * np.arange(64), np.full(64, 5.0)], 1)
    pos[64] = (-2.0, 2.6, 5.0)
    pos[65:] = np.stack([np.full(63, 5.0), np.full(63, 7.0), 2.6 + 0.07 * np.arange(63)], 1)
    edges = np.array([[0, 64], [65, 80]], np.int32)
    md = pa.MD(pos, None, edges, L, 2.5, 0.001)
    md.set_sliced_form(form)
    md.set_slice_order(order)
    _, _, f0 = md.get()
    rf0, _ = orc.lj_force(pos, edges, L, 2.5)
    assert np.array_equal(f0, rf0) and f0[0, 0] != 0.0
    md.run(3)
    md.run(3)
    p, v, f = md.get()
    md.close()
    rp, rv, rf, _ = orc.md_run(pos, np.zeros_like(pos), edges, L, 2.5, 0.001, 6)
    assert 7.5 < rp[64, 0] < 8.5
    assert np.array_equal(f, rf) and np.array_equal(v, rv) and np.array_equal(p, rp)


def test_reciprocal_matches_ieee_division(pa, gpu):
    """The force kernels' reciprocal (one Newton step on v_rcp_f32) equals the
    IEEE quotient 1.0f / b for every FP32 value in [2^-126, 2^125]: checked
    exhaustively (about 2.1e9 values)."""
    import ctypes
    one, two = ctypes.c_int64(-1), ctypes.c_int64(-1)
    pa.call("pat_lj_rcp_selftest", 2.0 ** -126, 2.0 ** 125, ctypes.byref(one), ctypes.byref(two))
    assert one.value == 0 and two.value == 0


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_force_sliced_kinds_vs_oracle(pa, orc, gpu, small, kind, order):
    """Every form of the sliced kernel sums each row in CSR order: bit-identical
    to the serial scatter over the list in the row order."""
    pos, box, edges = small
    n = len(pos)
    rowptr, cols = pa.csr_build_dev(pa.DeviceArray.from_numpy(edges), len(edges), n, order=order)
    sptr, scol = pa.sell_build_dev(rowptr, cols, n)
    p4 = pa.DeviceArray.from_numpy(pa.to_float4(pos))
    f4 = pa.DeviceArray((n, 4), np.float32)
    pa.call("pat_lj_force_sell_dev", p4.ptr, rowptr.ptr, sptr.ptr, scol.ptr, n, box, 1, 2.5, kind, f4.ptr, None)
    pa.sync()
    ref, _ = orc.lj_force(pos, oracle_edges(pa, edges, order), box, 2.5)
    assert np.array_equal(f4.numpy()[:, :3], ref)


@pytest.mark.parametrize("order", [0, 1])
def test_open_boundary_minimd_vs_oracle(pa, orc, gpu, order):
    """miniMD's map1 (no periodic images; miniMD/irregular.h:109-163) on a
    16^3-cell FCC system (16,384 atoms, jitter 0.05) with its half list of
    the pairs within 2.8 and no images, in seeded random order: the in-order
    kernel equals the oracle's serial scatter (no minimum image) bit for
    bit, rows in edge order and in node order."""
    pos, box = pa.gen_fcc(16, 16, 16, rho=0.8442, jitter=0.05, seed=31)
    edges = orc.pairs_brute(pos, box[0], 2.8, pbc=False)
    edges = edges[np.random.default_rng(32).permutation(len(edges))]
    ir = pa.IrregularRuntime(edges, len(pos), pa.PAT_IR_LJ_OPEN, np.array([2.5 * 2.5]))
    ir.set_row_order(order)
    ir.reset_node_data(pos)
    ir.start()
    got = ir.get_reduction_result()
    ir.close()
    ref, _ = orc.lj_force(pos, oracle_edges(pa, edges, order), box[0], 2.5, pbc=False)
    assert np.array_equal(got, ref)
    # without images the faces feel no pull from across the box: the atoms
    # nearest the x = 0 face are pulled inwards on average
    face = pos[:, 0] < 1.0
    assert np.mean(ref[face, 0]) > 0
    f64, _ = orc.lj_force(pos.astype(np.float64), edges, box[0], 2.5, pbc=False, dtype=np.float64)
    print(f"miniMD map1 16^3 ({len(edges)} pairs, order {order}): FP32 vs FP64 {rel(got, f64):.2e} relative")
    assert rel(got, f64) < 1e-5
