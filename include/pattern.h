/*
 * pattern.h -- C ABI of libpattern, the MI355X (gfx950) engines behind the
 * IPDPS'15 pattern framework's API.
 *
 * Every entry point returns an int status (PAT_OK or a negative PAT_ERR_*
 * code); pat_last_error() returns the calling thread's last message.  Sizes
 * are int64_t, handles are opaque, buffers are plain pointers.  Entry points
 * named *_dev take device pointers and an optional hipStream_t passed as
 * void* (NULL = the null stream); handle-level entry points take host
 * buffers, own device copies, and never modify caller memory.
 *
 * Each declaration names the reference interface it replaces
 * (file:line in lcchen008/IPDPS15-Pattern).
 */
#ifndef PATTERN_H_
#define PATTERN_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PAT_OK 0
#define PAT_ERR_INVALID -1     /* bad argument or size */
#define PAT_ERR_HIP -2         /* a HIP runtime call failed */
#define PAT_ERR_NOMEM -3       /* device allocation failed */
#define PAT_ERR_BUDGET -4      /* job does not fit the HBM budget */
#define PAT_ERR_UNSUPPORTED -5 /* functor index / variant not available */
#define PAT_ERR_STATE -6       /* call out of order (e.g. run before upload) */

/* Last error message of the calling thread ("" when none). */
const char *pat_last_error(void);

/* ------------------------------------------------------------------------
 * Runtime environment
 * ---------------------------------------------------------------------- */

/* Replaces RuntimeInit (kmeans/include/runtime.cpp:4-8; MPI_Init there):
 * selects the HIP device and creates the library's streams. */
int pat_init(int device);
/* Replaces RuntimeFinalize (kmeans/include/runtime.cpp:10-13). */
int pat_finalize(void);
/* Library version, build target and device name into buf. */
int pat_info(char *buf, int64_t buf_len);
/* Blocks until all work queued by the library on the current device ends. */
int pat_device_sync(void);

/* ------------------------------------------------------------------------
 * Device memory, streams and events for host harnesses (no counterpart in
 * the reference, which moves data with cudaMemcpy/mapped host memory inside
 * its runtimes; exposed so that callers in other languages can keep inputs
 * resident and time work without linking a second HIP runtime).
 * ---------------------------------------------------------------------- */
int pat_dev_malloc(void **ptr, int64_t bytes);
int pat_dev_free(void *ptr);
int pat_memcpy_h2d(void *dst, const void *src, int64_t bytes, void *stream);
int pat_memcpy_d2h(void *dst, const void *src, int64_t bytes, void *stream);
int pat_memcpy_d2d(void *dst, const void *src, int64_t bytes, void *stream);
int pat_memset_dev(void *ptr, int value, int64_t bytes, void *stream);
/* Default-priority non-blocking stream. */
int pat_stream_create(void **stream);
int pat_stream_destroy(void *stream);
int pat_stream_sync(void *stream);
int pat_event_create(void **event);
int pat_event_destroy(void *event);
int pat_event_record(void *event, void *stream);
/* Milliseconds between two recorded events (waits for 'end'). */
int pat_event_elapsed_ms(void *start, void *end, float *ms);
/* Device properties: CU count, total HBM bytes, free HBM bytes. */
int pat_device_query(int *num_cus, int64_t *total_bytes, int64_t *free_bytes);

/* ------------------------------------------------------------------------
 * Seeded workload generators (host code, deterministic for any thread
 * count; the app drivers' data set-up, e.g. kmeans/km.cu:28-65,
 * kmeans/moldyn.cu:637-741, miniMD/setup.cpp:61-147).
 * ---------------------------------------------------------------------- */

/* Halo-inclusive (nx+2)(ny+2)(nz+2) FP32 grid, x fastest; interior uniform
 * in [0,1) from seed, halo 0. */
int pat_gen_heat3d(int64_t nx, int64_t ny, int64_t nz, uint64_t seed, float *grid);
/* FCC lattice of ncx*ncy*ncz unit cells at reduced density rho, atoms
 * numbered as miniMD create_atoms visits lattice sites (miniMD/setup.cpp:
 * 92-113), each coordinate jittered uniformly in [-jitter, jitter) and
 * wrapped into [0, L).  pos: 4*ncx*ncy*ncz x 3.  box_out: L per axis (3). */
int pat_gen_fcc(int64_t ncx, int64_t ncy, int64_t ncz, double rho, double jitter, uint64_t seed, float *pos,
                double *box_out);
/* Maxwell velocities at temperature t (unit mass, k_B = 1): Gaussian
 * components, zero centre-of-mass motion, rescaled so that
 * sum v^2 / (3(N-1)) = t (miniMD create_velocity, miniMD/setup.cpp:150-190). */
int pat_gen_velocities(int64_t n, double t, uint64_t seed, float *vel);
/* Half list of all pairs i<j whose minimum-image distance is below cut
 * (cell-list build; the moldyn program's RebuildNeighbor,
 * kmeans/moldyn.cu:552-594, is the O(N^2) form).  Pairs come out in (i, j)
 * order, or shuffled with shuffle_seed when shuffle != 0.  Call with
 * edges == NULL to get the count in *num_edges; otherwise *num_edges is the
 * capacity on input and the count on output. */
int pat_gen_pairs(int64_t n, const float *pos, double box, double cut, int shuffle, uint64_t shuffle_seed,
                  int32_t *edges, int64_t *num_edges);
/* Gaussian mixture: ncomp centres uniform in [lo, hi]^d from seed, each
 * point picks a component uniformly and adds sigma * N(0, 1) per axis.
 * x: n x d row-major. */
int pat_gen_gmm(int64_t n, int d, int ncomp, double sigma, double lo, double hi, uint64_t seed, float *x);
/* Device forms of the generators, writing device memory on 'stream' (not
 * waited for): the heat3D grid (halo-inclusive, as pat_gen_heat3d) and the
 * FCC lattice (float4 x, y, z, 0 per atom) are bit-identical to the host
 * generators; the mixture points use the device's FP64 log/cos/sqrt. */
int pat_gen_heat3d_dev(int64_t nx, int64_t ny, int64_t nz, uint64_t seed, float *grid, void *stream);
int pat_gen_fcc_dev(int64_t ncx, int64_t ncy, int64_t ncz, double rho, double jitter, uint64_t seed, float *pos4,
                    double *box_out, void *stream);
int pat_gen_gmm_dev(int64_t n, int d, int ncomp, double sigma, double lo, double hi, uint64_t seed, float *x,
                    void *stream);

/* ------------------------------------------------------------------------
 * Stencil engine (replaces compute_cuda_internal, kmeans/lib/compute_cuda.cu:
 * 30-100 / heat_3D/lib/compute_cuda.cu:30-105, driven by StencilRuntime,
 * kmeans/lib/stencil_runtime.h:13-93)
 * ---------------------------------------------------------------------- */

/* Stencil functor indices (stencil_function_table, heat_3D/stencil.h:119). */
#define PAT_STENCIL_HEAT 0 /* u' = u + alpha*(sum6 - 6u), parameter = float alpha */

/* Kernel variants. */
#define PAT_STENCIL_AUTO 0    /* autotuned choice */
#define PAT_STENCIL_LDS25D 1  /* LDS-staged 2.5D z-marching tiles */
#define PAT_STENCIL_GENERIC 2 /* one thread per cell calling the functor */
#define PAT_STENCIL_TEMPORAL2 3 /* two sweeps fused per pass (temporal blocking, 64-wide windows of 8,
                                   16 or 32 rows); an odd last step runs the LDS25D kernel.  Same bits */
#define PAT_STENCIL_CUBES 4 /* several sweeps per launch on blocks held in LDS; bx = shape: 0, 1, 2 =
                               16^3 blocks, 2, 3 or 4 sweeps; 3 = 32x16x8, 2 sweeps; 4 = 32x8x8, 2
                               sweeps.  Steps left over run the LDS25D kernel.  Same bits */

/* nsteps ping-pong heat sweeps over device grids a, b (halo-inclusive,
 * x fastest, both halos already 0).  *result_in_b = 1 when the final field
 * is in b.  bx, by, zc: tile (x threads, y rows, z-chunk), 0 = tuned. */
int pat_heat3d_run_dev(float *a, float *b, int64_t nx, int64_t ny, int64_t nz, int nsteps, float alpha,
                       int variant, int bx, int by, int zc, int *result_in_b, void *stream);

typedef struct pat_stencil pat_stencil;
/* StencilCreator<T>(...) + CreateRuntime + StencilInit
 * (heat_3D/lib/stencil_creator.h:22-30, kmeans/lib/stencil_runtime.cu:34-67):
 * a device-resident double-buffered grid of ndims (2 or 3) with the given
 * interior size (x, y[, z]), halo width 'width' (1..16), element size
 * unit_size (4 = float, 8 = double).  The library's own engine runs 3-D FP32
 * grids of width 1; the rest need an application table
 * (pat_stencil_set_table). */
int pat_stencil_create(int ndims, const int64_t *interior, int width, int unit_size, const void *param,
                       int64_t param_size, pat_stencil **out);
/* set_stencil_idx (kmeans/lib/stencil_runtime.h:36): an index of the
 * registered application table, or PAT_STENCIL_HEAT without one. */
int pat_stencil_set_idx(pat_stencil *s, int idx);
/* An application's stencil functor table compiled for the device: the
 * stencil_function_table of its functor header (heat_3D/stencil.h:119,
 * kmeans/stencil.h:251), built by hipcc with GPU_KERNEL defined through the
 * templates of api/gpu/stencil_table.hpp, which also define
 * pat_app_stencil_table() returning this struct.  Replaces the reference's
 * __constant__ function-pointer call per cell (kmeans/lib/compute_cuda.cu:95). */
typedef struct pat_stencil_table {
    int count;             /* functors in the table */
    const int *unit_size;  /* per functor: bytes per grid element it reads and writes */
    const int *same_as;    /* per functor: the library functor giving the same bits
                              (PAT_STENCIL_HEAT: the tuned engine runs it), or -1 */
    /* One sweep of functor idx: out = f(in) at every interior cell of a
     * halo-inclusive grid of extents size[0..ndims-1] (dimension 0 fastest)
     * with halo width 'width'; param is a device pointer (may be NULL);
     * queued on stream, not waited for. */
    int (*sweep)(int idx, const void *in, void *out, const int32_t *size, int ndims, int width, const void *param,
                 void *stream);
} pat_stencil_table;
/* Registers (or with NULL removes) the application's table for the handle.
 * Its entries then run on grids of any halo width and of the element size
 * they declare; entries declared the same as PAT_STENCIL_HEAT keep the tuned
 * engine on 3-D FP32 grids of width 1 unless PAT_STENCIL_GENERIC is set.
 * The table must outlive the handle. */
int pat_stencil_set_table(pat_stencil *s, const pat_stencil_table *table);
/* Selects the kernel variant and tile (0 = tuned). */
int pat_stencil_set_variant(pat_stencil *s, int variant, int bx, int by, int zc);
/* Copies a halo-inclusive host grid in (the app writes grid()->data(),
 * heat_3D/main.cpp:72-75). */
int pat_stencil_upload(pat_stencil *s, const void *host);
/* StencilBegin for nsteps iterations (kmeans/lib/stencil_runtime.cu:576-614). */
int pat_stencil_run(pat_stencil *s, int nsteps);
/* pat_stencil_run with device_ms = the time between HIP events recorded on
 * the handle's stream around the replay of the captured sweep graph (every
 * launch and the gaps between them).  Timing only; the reference has no
 * counterpart. */
int pat_stencil_run_graph_timed(pat_stencil *s, int nsteps, float *device_ms);
/* nsteps sweeps launched one by one with HIP events around each kernel:
 * total_ms = first kernel start to last kernel end, kernel_ms[k] = sweep k
 * (array of nsteps, may be NULL).  Blocking. */
int pat_stencil_run_timed(pat_stencil *s, int nsteps, float *total_ms, float *kernel_ms);
/* Copies the current halo-inclusive grid out. */
int pat_stencil_download(pat_stencil *s, void *host);
/* Tile the runs use (tuned on the first run when none was set): x width,
 * tile height in rows, z-chunk in planes. */
int pat_stencil_tile(pat_stencil *s, int *bx, int *by, int *zc);
/* Device pointer of the current grid. */
int pat_stencil_device_ptr(pat_stencil *s, void **dptr);
/* StencilFinalize + delete. */
int pat_stencil_destroy(pat_stencil *s);

/* ------------------------------------------------------------------------
 * Irregular-reduction engine (replaces irregular compute_gpu,
 * kmeans/irregular/gpu_kernel.cu:21-103, with the LDS/device hash reduction
 * objects ros.cu/rog.cu, driven by IrregularRuntime,
 * kmeans/irregular/irregular_runtime.h:12-112)
 * ---------------------------------------------------------------------- */

/* Edge functor indices (map_function_table). */
#define PAT_IR_LJ_MINIMAGE 0 /* moldyn map, kmeans/irregular.h:26-107: minimum image;
                                param = struct para {double side; double cutoffRadius;} */
#define PAT_IR_LJ_OPEN 1     /* miniMD map1, miniMD/irregular.h:109-163: no images;
                                param = double cutforcesq */
#define PAT_IR_NO_FUNCTOR -1 /* pat_ir_create before pat_ir_set_table: no library functor */
/* Kernel variants. */
#define PAT_IR_CSR 0      /* by-destination CSR, each row summed in its stored order: bit-identical
                             to a serial scatter over the edge list in that order (default) */
#define PAT_IR_ATOMIC 1   /* edge-parallel, float atomics into both endpoints */
#define PAT_IR_CSR_TREE 2 /* CSR with butterfly (tree-order) row sums: FP32-tolerance parity */
#define PAT_IR_TABLE_KERNEL 3 /* the registered application table's own per-node kernel, also for
                                 functors a library kernel computes (pat_ir_set_table) */
/* lanes_per_row value of PAT_IR_CSR selecting the sliced layout: one lane per
 * row over 64-row slices stored entry-major (SELL-64), so a wavefront's k-th
 * column read is one contiguous load.  Same sums, bit for bit, as every other
 * in-order choice; available when padding stays within 1.5x the entries. */
#define PAT_IR_SLICED 64
/* CSR row orders.  A row's order is the order its pair forces are added in. */
#define PAT_IR_ORDER_EDGE 0 /* as the edges appear in the list: the serial scatter over the list as given */
#define PAT_IR_ORDER_NODE 1 /* by neighbour index: the serial scatter over the list sorted by
                               (min(i, j), max(i, j)); keeps gathers of neighbouring rows
                               close in memory (handle default) */

/* Bytes of device workspace pat_ir_csr_build_dev needs. */
int pat_ir_csr_workspace(int64_t num_nodes, int64_t num_edges, int64_t *bytes);
/* Reorders a half edge list (int32 pairs i, j, i != j) once into a CSR by
 * destination: row r lists every neighbour of r (both directions), in edge
 * list order (PAT_IR_ORDER_EDGE) or by neighbour index (PAT_IR_ORDER_NODE).
 * rowptr: num_nodes+1, cols: 2*num_edges.  Replaces the host Partitioner /
 * Reorder (kmeans/irregular/partitioner.cpp:20-499, reorder.cpp:13-281). */
int pat_ir_csr_build_dev(const int32_t *edges, int64_t num_edges, int64_t num_nodes, int order, int32_t *rowptr,
                         int32_t *cols, void *work, int64_t work_bytes, void *stream);
/* pat_ir_csr_build_dev without allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The work
 * space is a static 128 MiB device array that every GPU the library runs on
 * holds until the process ends; PAT_ERR_UNSUPPORTED when the build needs more
 * (pat_ir_csr_workspace bytes: about 16 per edge + 4 per node, so up to ~8.3M
 * edges).  It calls no allocator and records no memset.  Calls that overlap
 * on one GPU share the array and race: serialise them. */
int pat_ir_csr_build_dev_noalloc(const int32_t *edges, int64_t num_edges, int64_t num_nodes, int order,
                                 int32_t *rowptr, int32_t *cols, void *stream);
/* LJ forces from float4 positions (x, y, z, unused) into float4 forces over
 * the CSR; pbc selects the minimum image in a cubic box of side 'box'.
 * lanes_per_row 1, 4, 8, 16, 32: rows summed in CSR order (bit-identical for
 * every choice); -4 ... -64: tree-order sums over that many lanes. */
int pat_lj_force_csr_dev(const float *pos4, const int32_t *rowptr, const int32_t *cols, int64_t num_nodes, float box,
                         int pbc, float cutoff, int lanes_per_row, float *force4, void *stream);
/* Sliced copy (SELL-64) of a CSR for pat_lj_force_sell_dev: slice s covers
 * rows 64s..64s+63 and is as wide as its longest row; entry k of row 64s+l
 * is scol[sptr[s] + 64k + l] (padding repeats the row's own index).  sptr:
 * num_nodes/64 + 1 ints (rounded up); scol: scol_capacity ints.
 * *num_entries = the padded size; PAT_ERR_INVALID when the capacity is
 * short (call again with that much). */
int pat_ir_sell_build_dev(const int32_t *rowptr, const int32_t *cols, int64_t num_nodes, int32_t *sptr,
                          int32_t *scol, int64_t scol_capacity, int64_t *num_entries, void *stream);
/* LJ forces over the sliced copy, each row summed in CSR order (the same
 * bits as pat_lj_force_csr_dev); kind 0, 1, 2 = plain, software-pipelined by
 * 4, by 8 entries. */
int pat_lj_force_sell_dev(const float *pos4, const int32_t *rowptr, const int32_t *sptr, const int32_t *scol,
                          int64_t num_nodes, float box, int pbc, float cutoff, int kind, float *force4, void *stream);
/* Compares, for every FP32 value in [lo, hi], the IEEE quotient 1.0f / b
 * with the hardware reciprocal refined by one and by two Newton steps, and
 * returns how many results differ for each (the check behind the LJ
 * kernels' reciprocal; no counterpart in the reference). */
int pat_lj_rcp_selftest(float lo, float hi, int64_t *bad_one_step, int64_t *bad_two_step);
/* The same forces from the half edge list with float atomics (zeroes force4
 * first). */
int pat_lj_force_atomic_dev(const float *pos4, const int32_t *edges, int64_t num_edges, int64_t num_nodes, float box,
                            int pbc, float cutoff, float *force4, void *stream);

typedef struct pat_ir pat_ir;
/* IrregularRuntime(num_edges, num_nodes, edges, ...) + IrregularInit
 * (irregular_runtime.h:71-111, irregular_runtime.cu:17-193): uploads the
 * half edge list and builds the CSR on the device.  The caller's arrays are
 * not modified (the reference permutes them in place).  map_idx: a library
 * functor with its parameter, or PAT_IR_NO_FUNCTOR (param may be NULL) when
 * an application table follows. */
int pat_ir_create(int64_t num_edges, int64_t num_nodes, const int32_t *edges, int map_idx, const void *param,
                  int64_t param_size, pat_ir **out);
/* set_map_idx (irregular_runtime.h:110): a library functor (PAT_IR_LJ_*), or
 * with an application table registered an index of that table. */
int pat_ir_set_map_idx(pat_ir *h, int idx);
/* set_reduce_idx (irregular_runtime.h:111): an index of the application
 * table's reduce functors (the library's own functors add float3 values). */
int pat_ir_set_reduce_idx(pat_ir *h, int idx);
/* An application's irregular functor tables compiled for the device: the
 * map_function_table / reduce_function_table of its functor header
 * (kmeans/irregular.h:116-118, miniMD/irregular.h:172-174), built by hipcc
 * with GPU_KERNEL defined through api/gpu/irregular_table.hpp, which also
 * defines pat_app_ir_table() returning this struct.  Replaces the reference's
 * per-partition kernel with its LDS/device hashes
 * (kmeans/irregular/gpu_kernel.cu:21-103). */
typedef struct pat_ir_table {
    int map_count, reduce_count; /* functors in the two tables */
    int value_size;              /* bytes of the reduction VALUE */
    const int *same_as;          /* per map: the library functor giving the same bits, or -1 */
    /* For every node r: map map_idx over the edges in row r of the CSR by
     * destination (rowptr; eids = edge ids, increasing within a row) with
     * part_index = identity and part_id = r, into node r's object (first
     * insert copied, later ones folded by reduce reduce_idx; inserts for
     * other nodes refused and flagged in *foreign), written to out[r]
     * (value_size bytes; zero for a node nothing was inserted for).  Device
     * pointers; queued on stream, not waited for. */
    int (*reduce_rows)(int map_idx, int reduce_idx, const int32_t *rowptr, const int32_t *eids, int64_t num_nodes,
                       const void *point_data, const int32_t *part_index, const void *edge_data, const void *param,
                       void *out, int32_t *foreign, void *stream);
} pat_ir_table;
/* Registers (or with NULL removes) the application's tables; map and reduce
 * indices restart at 0 and the node data must be set again.  Maps declared
 * the same as a library functor run on its kernels unless
 * PAT_IR_TABLE_KERNEL is selected.  The table must outlive the handle. */
int pat_ir_set_table(pat_ir *h, const pat_ir_table *table);
/* The functor parameter (the runtime's 'parameter' argument). */
int pat_ir_set_param(pat_ir *h, const void *param, int64_t param_size);
/* Kernel variant and lanes per row (0 = tuned for PAT_IR_CSR). */
int pat_ir_set_variant(pat_ir *h, int variant, int lanes_per_row);
/* Row order of the CSR (rebuilt from the uploaded edges when it changes);
 * handles start in PAT_IR_ORDER_NODE.  The reference's runtime also sums in
 * an order of its own, the edge order of its partitions
 * (kmeans/irregular/reorder.cpp:13-281). */
int pat_ir_set_row_order(pat_ir *h, int order);
/* reset_node_data (irregular_runtime.cu:427-439): node data of node_elm_size
 * bytes per node: float3 (12) or double3 (24, moldyn's point_data) for the
 * library's functors; any size, passed to the functors as given, with an
 * application table. */
int pat_ir_reset_node_data(pat_ir *h, const void *node_data, int node_elm_size);
/* IrregularStart (irregular_runtime.cu:311-391): one reduction pass. */
int pat_ir_start(pat_ir *h);
/* get_reduction_result (irregular_runtime.cu:393-411): N VALUEs in the
 * caller's node order: float3 (kmeans/irregular/parameters.h:19-28), or the
 * application table's value_size bytes each. */
int pat_ir_get_result(pat_ir *h, float *out);
int pat_ir_destroy(pat_ir *h);

/* Device-resident molecular dynamics around the LJ reduction: the moldyn
 * and miniMD drivers' step loops (kmeans/moldyn.cu:331-403,
 * miniMD/ljs.cu:171-184) as velocity-Verlet with unit mass:
 * v += dt/2 f; x += dt v; wrap; f = F(x); v += dt/2 f.  The CSR variant
 * fuses the update into the force kernel; run() replays a captured
 * hipGraph.  pos, vel: N x 3 FP32 (vel may be NULL = 0). */
typedef struct pat_md pat_md;
int pat_md_create(int64_t num_nodes, const float *pos, const float *vel, int64_t num_edges, const int32_t *edges,
                  double box, int pbc, double cutoff, double dt, pat_md **out);
/* The same loop without a caller edge list: the neighbour CSR (all pairs
 * within cutoff + skin, periodic images) is built on the device from the
 * positions by binning, and rebuilt before the force of every step whose
 * index is a multiple of 'every' -- miniMD's Neighbor::setup/build
 * (miniMD/neighbor.cpp:76-427) with its reneighbouring interval and skin
 * (miniMD/lj.in:9-10).  Needs box >= 3 * (cutoff + skin), or box >= 2 *
 * (cutoff + skin) for at most 65,536 atoms (every pair is then tested). */
int pat_md_create_binned(int64_t num_nodes, const float *pos, const float *vel, double box, double cutoff,
                         double skin, int every, double dt, pat_md **out);
/* The same from device float4 arrays (vel4 may be NULL = 0), copied. */
int pat_md_create_binned_dev(int64_t num_nodes, const float *pos4, const float *vel4, double box, double cutoff,
                             double skin, int every, double dt, pat_md **out);
int pat_md_set_variant(pat_md *h, int variant, int lanes_per_row);
/* Row order (see pat_ir_set_row_order); recomputes the stored force.  Binned
 * handles only have PAT_IR_ORDER_NODE. */
int pat_md_set_row_order(pat_md *h, int order);
/* The kernel in use: variant, lanes per row (after tuning; PAT_IR_SLICED =
 * the sliced layout) and row order. */
int pat_md_get_variant(pat_md *h, int *variant, int *lanes_per_row, int *order);
/* Selects the sliced layout with loop form 0 (plain), 1 (software-pipelined
 * by 4), 2 (by 8), or 3, 4 (neighbour positions staged per 256-row block in
 * LDS, 4 or 8 entries ahead; PAT_ERR_UNSUPPORTED when a block's neighbour
 * list does not fit) in place of the tuner's choice; recomputes the stored
 * force.  All forms give the same bits.  Tuning override of
 * pat_md_set_variant(h, PAT_IR_CSR, PAT_IR_SLICED); no reference
 * counterpart (the reference has one kernel form). */
int pat_md_set_sliced_form(pat_md *h, int form);
/* The sliced loop form in use (0..4 as above), or -1 when the handle runs
 * another kernel.  No reference counterpart. */
int pat_md_get_sliced_form(pat_md *h, int *form);
/* Order of rows in the sliced layout's 64-row slices: 0 node order, 1 row
 * class (nodes near a box face grouped within 256-row windows), 2 spatial
 * (rows sorted by the Morton code of a half-cutoff bin, so each slice's
 * rows share one compact gather neighbourhood; recomputed at every
 * rebuild).  Rows keep their own entries in order, so every order gives the
 * same bits.  In place of the tuner's choice; PAT_ERR_UNSUPPORTED without
 * the sliced sentinel layout of a periodic box.  No reference counterpart
 * (the reference has no device layout). */
int pat_md_set_slice_order(pat_md *h, int order);
/* The slice order of the current layout (0, 1, 2 as above). */
int pat_md_get_slice_order(pat_md *h, int *order);
/* Current half-pair count, steps run, neighbour rebuilds and their total
 * device time (HIP events around each rebuild's launches, the read-back of
 * the pair count between them included). */
int pat_md_stats(pat_md *h, int64_t *num_pairs, int64_t *steps, int64_t *rebuilds, float *rebuild_ms);
/* Current CSR by destination (rowptr: N+1, cols: 2 * num_pairs). */
int pat_md_csr(pat_md *h, int32_t *rowptr, int32_t *cols);
/* nsteps steps, blocking. */
int pat_md_run(pat_md *h, int nsteps);

/* Many independent periodic systems advanced as one: system s owns nodes
 * [offsets[s], offsets[s+1]) of the concatenated pos/vel arrays (N x 3) and
 * has cubic box box[s].  Each node carries its box in the float4 w lane, so
 * one fused force+update launch covers every system whatever its size, and
 * the steps between neighbour rebuilds replay one captured hipGraph.
 * Rebuilds (cutoff + skin, every 'every' steps) are per system.  The batched
 * form of the miniMD driver (miniMD/ljs.cu:55-184) for many small systems. */
typedef struct pat_mdb pat_mdb;
int pat_mdb_create(int nsys, const int64_t *offsets, const float *pos, const float *vel, const double *box,
                   double cutoff, double skin, int every, double dt, pat_mdb **out);
/* Lanes per CSR row (0 = default); recomputes the stored force. */
int pat_mdb_set_lanes(pat_mdb *h, int lanes_per_row);
/* pat_md_set_slice_order for the batch (the spatial order sorts system by
 * system). */
int pat_mdb_set_slice_order(pat_mdb *h, int order);
/* nsteps steps (blocking); total_ms = host wall time, kernel_ms[k] = step
 * k's force kernel when non-NULL (then launched one by one with events). */
int pat_mdb_run(pat_mdb *h, int nsteps, float *total_ms, float *kernel_ms);
/* System sys's positions, velocities and forces (n_s x 3 each, may be NULL). */
int pat_mdb_get(pat_mdb *h, int sys, float *pos, float *vel, float *force);
int pat_mdb_stats(pat_mdb *h, int64_t *num_pairs, int64_t *steps, int64_t *rebuilds, float *rebuild_ms);
int pat_mdb_destroy(pat_mdb *h);
/* nsteps steps queued on 'stream' (NULL = the handle's stream), not waited for. */
int pat_md_run_async(pat_md *h, int nsteps, void *stream);
/* pat_md_run for a handle without rebuilds, with device_ms = the time
 * between HIP events recorded on the handle's stream around the replay of
 * the captured step graph (every force + update launch and the gaps between
 * them).  Timing only; the reference has no counterpart. */
int pat_md_run_graph_timed(pat_md *h, int nsteps, float *device_ms);
/* nsteps steps launched one by one on the handle's stream with HIP events
 * around each force kernel: total_ms = whole run, kernel_ms[k] = step k's
 * force kernel (array of nsteps, may be NULL).  Blocking. */
int pat_md_run_timed(pat_md *h, int nsteps, float *total_ms, float *kernel_ms);
int pat_md_get(pat_md *h, float *pos, float *vel, float *force);
/* Potential energy sum_pairs 4(r^-12 - r^-6) and kinetic energy sum v^2/2. */
int pat_md_energy(pat_md *h, double *epot, double *ekin);
/* Current device buffers (float4 per node). */
int pat_md_device_buffers(pat_md *h, void **pos4, void **vel4, void **force4);
int pat_md_destroy(pat_md *h);

/* ------------------------------------------------------------------------
 * Generalized-reduction engine, kmeans (replaces the regular compute_gpu,
 * kmeans/regular/gpu_kernel.cu:39-183, with the LDS/device hash reduction
 * objects ros.cu/rog.cu, running map1/reduce1, kmeans/regular.h:75-136,
 * driven by RegularRuntime, kmeans/regular/regular_runtime.h:11-106)
 * ---------------------------------------------------------------------- */

#define PAT_KM_PLAIN 0 /* FP32 difference form over every centre (map1's arithmetic) */
#define PAT_KM_MFMA 1  /* BF16 MFMA distances + exact FP32 rescoring inside the error margin */

/* Fixed-point shift s used for the order-independent sums of n values of
 * magnitude below maxabs (maxabs * 2^s <= 2^30 and n * maxabs * 2^s < 2^62). */
int pat_kmeans_fixed_shift(int64_t n, float maxabs, int *shift);

typedef struct pat_kmeans pat_kmeans;
/* RegularRuntime(num_procs, input, input_size, offsets, n, params, ...)
 * + RegularInit + set_map_idx(1)/set_reduce_idx(1) (regular_runtime.cu:
 * 12-126, km.cu:76-80) for n points of dimension d (row-major FP32, the
 * implicit stride replaces the Offset array) and k centres, initialised to
 * the first k points.  x is a host pointer (copied) or, with x_on_device, a
 * device pointer the handle reads in place. */
int pat_kmeans_create(const float *x, int64_t n, int d, int k, int mode, int x_on_device, pat_kmeans **out);
int pat_kmeans_set_mode(pat_kmeans *h, int mode);
/* BF16 terms of the MFMA path's distance values: 3 (x and -2c split into
 * high and low parts) or 1 (rounded once, a wider rescoring margin).  Both
 * give the plain path's assignments bit for bit; by default the faster one
 * on the handle's data is kept (timed at creation when both fit in LDS and
 * N >= 2^20; three terms otherwise).
 * PAT_ERR_UNSUPPORTED when the shape's rows do not fit the requested form.
 * No reference counterpart (the reference has no MFMA path). */
int pat_kmeans_set_terms(pat_kmeans *h, int terms);
int pat_kmeans_get_terms(pat_kmeans *h, int *terms);
/* Replaces the centres (k x d host array). */
int pat_kmeans_set_centroids(pat_kmeans *h, const float *c);
/* iters x (RegularStart: assign + reduce; then the Lloyd update
 * c_k = sum_k / count_k, empty clusters keep their centre).  Replays a
 * captured hipGraph.  Blocking. */
int pat_kmeans_run(pat_kmeans *h, int iters);
/* Same, launched one by one with HIP events: total_ms = whole run,
 * assign_ms[i] = iteration i's assignment kernel (may be NULL). */
int pat_kmeans_run_timed(pat_kmeans *h, int iters, float *total_ms, float *assign_ms);
/* Current centres (k x d), last assignment (n) and its min_dist (n); any
 * pointer may be NULL. */
int pat_kmeans_get(pat_kmeans *h, float *centroids, int32_t *assign, float *mindist);
/* RegularGetOutput (regular_runtime.cu:288-291): the last pass's reduction
 * object -- per key the fixed-point coordinate sums (k x d), the point
 * counts (k) and the fixed-point min_dist sums (k), with their shifts
 * (value = sum * 2^-shift). */
int pat_kmeans_get_object(pat_kmeans *h, int64_t *sums, int64_t *counts, int64_t *dsums, int *shift, int *dshift);
/* Points rescored in FP32 so far (MFMA mode) and max |x| of the data. */
int pat_kmeans_stats(pat_kmeans *h, int64_t *rescored_total, float *maxabs);
int pat_kmeans_destroy(pat_kmeans *h);

/* Generalized reduction with the applications' additive map/reduce pairs
 * (reduce adds the emitted values), the RegularRuntime work beside kmeans:
 * RegularRuntime(num_procs, input, input_size, offsets, n, params,
 * param_size) + RegularInit + set_map_idx/set_reduce_idx + RegularStart +
 * RegularGetOutput (kmeans/regular/regular_runtime.h:79-105) with
 *   PAT_GR_VSTATS (0): kmeans/regular.h:30-73 map/reduce as moldyn runs it
 *     (kmeans/moldyn.cu:440-470): units = velocities (double[3]), param =
 *     float vaverh; per unit {(float)|v|, |v| > vaverh}; 1 key, 2 values.
 *   PAT_GR_LJ_ENERGY (2): miniMD/regular.h:139-181 map2/reduce2: units =
 *     edges (int[2]), param = {double cutsq, double x[3N]}; per edge within
 *     the cutoff phi = sr6 (sr6 - 1); 1 key, 1 value.
 * The reference's ANY_KEY buckets are folded into key 0 and the values are
 * summed in FP64 in a fixed order (the same bits every run).  Units and
 * parameter are copied at create.  kmeans map1 (index 1) runs on
 * pat_kmeans_*. */
#define PAT_GR_VSTATS 0
#define PAT_GR_LJ_ENERGY 2
typedef struct pat_gr pat_gr;
int pat_gr_create(const void *units, int64_t num_units, int64_t unit_bytes, int map_idx, const void *param,
                  int64_t param_bytes, pat_gr **out);
/* set_map_idx / set_reduce_idx: the handle's own map and its adding reduce
 * only (PAT_ERR_UNSUPPORTED otherwise: units are laid out per map). */
int pat_gr_set_map_idx(pat_gr *h, int idx);
int pat_gr_set_reduce_idx(pat_gr *h, int idx);
int pat_gr_set_param(pat_gr *h, const void *param, int64_t param_bytes);
/* RegularStart: one map + reduce pass over every unit.  Blocking. */
int pat_gr_run(pat_gr *h);
/* RegularGetOutput: the reduction object, keys x values FP64 sums
 * (values may be NULL) and the emits per key (may be NULL). */
int pat_gr_get_output(pat_gr *h, int *num_keys, int *num_values, double *values, int64_t *emits);
int pat_gr_destroy(pat_gr *h);

/* Generalized reduction with an application's own map/reduce functors and
 * keyed reduction objects: RegularRuntime(num_procs, input, input_size,
 * offsets, n, params, param_size) + RegularInit + set_map_idx /
 * set_reduce_idx + RegularStart + RegularGetOutput
 * (kmeans/regular/regular_runtime.h:79-105) for any map_fp / reduce_fp
 * (kmeans/regular.h:25-27) with any key and value bytes, their tables
 * compiled for the device by hipcc with GPU_KERNEL defined through
 * api/gpu/regular_table.hpp, which defines pat_app_gr_table().  Replaces the
 * reference's regular compute_gpu with its LDS and device hash objects
 * (kmeans/regular/gpu_kernel.cu:39-183, ros.cu:15-170, rog.cu:17-310).
 * Each wavefront folds its units into an LDS hash table (lanes inserting in
 * turn); a full table is merged into the wavefront's table in device memory
 * and the refused units' maps run again; the wavefronts' tables are then
 * folded in a fixed order into the final table, so every value is the same
 * fold every run. */
typedef struct pat_gr_object { /* device memory of one run, set up by the library */
    const void *data;          /* the units' bytes */
    const int64_t *offsets;    /* byte offset of each unit in data, or NULL: unit i at i * unit_bytes */
    int64_t unit_bytes, num_units;
    const void *param;         /* the functors' parameter (may be NULL) */
    uint32_t *wave_tables;     /* waves tables of (table_buckets + 1 + table_pool_words) words */
    uint32_t *final_table;     /* one table of that size */
    int32_t *status;           /* [0] error bits (1 a wavefront's table full, 2 a unit's values do not fit
                                  an empty LDS table, 4 a map's insert refused after another was taken,
                                  8 the final table full), [1] LDS merges, [2] maps run again */
    int32_t waves, table_buckets, table_pool_words, lds_buckets, lds_pool_words;
} pat_gr_object;
typedef struct pat_gr_table {
    int map_count, reduce_count;
    int max_lds_buckets, max_lds_pool_words; /* the kernels' LDS tables hold at most this much */
    /* One pass of map map_idx over every unit with reduce reduce_idx into the
     * final table; queued on stream, not waited for. */
    int (*map_reduce)(int map_idx, int reduce_idx, const pat_gr_object *obj, void *stream);
} pat_gr_table;
typedef struct pat_regular pat_regular;
/* RegularRuntime(..., input, input_size, offsets, n, ...) + RegularInit: the
 * input (copied to the device) and its units: unit i is unit_bytes bytes at
 * offsets[i] (Offset.offset, kmeans/regular/data_type.h:21-27), or at
 * i * unit_bytes when offsets is NULL. */
int pat_regular_create(const void *input, int64_t input_bytes, int64_t num_units, int64_t unit_bytes,
                       const int64_t *offsets, pat_regular **out);
/* The application's table (required before a start; must outlive the handle). */
int pat_regular_set_table(pat_regular *h, const pat_gr_table *table);
/* Capacity of the reduction objects: distinct keys and bytes of key + value
 * per record (default 4096 keys of 256 bytes), and the keys an LDS table
 * takes before it is merged (0: as many as fit; a small number forces the
 * merge-and-run-again path). */
int pat_regular_set_capacity(pat_regular *h, int64_t max_keys, int64_t max_record_bytes, int lds_keys);
int pat_regular_set_map_idx(pat_regular *h, int idx);
int pat_regular_set_reduce_idx(pat_regular *h, int idx);
int pat_regular_set_param(pat_regular *h, const void *param, int64_t param_bytes);
/* RegularStart: one map + reduce pass over every unit.  Blocking; fails with
 * the status error bits named above. */
int pat_regular_start(pat_regular *h);
/* RegularGetOutput: the final object's records sorted by key bytes.  Call
 * with NULL buffers for *num_keys and the byte totals; then keys (key bytes
 * back to back), key_index (byte offset of each key), key_size, vals,
 * val_index, val_size are filled (each array num_keys long). */
int pat_regular_get_output(pat_regular *h, int64_t *num_keys, int64_t *key_bytes, int64_t *val_bytes, void *keys,
                           uint32_t *key_index, uint32_t *key_size, void *vals, uint32_t *val_index,
                           uint32_t *val_size);
/* LDS merges and maps run again in the last start (the reference's
 * merge-and-retry), and the wavefronts used. */
int pat_regular_stats(pat_regular *h, int64_t *merges, int64_t *reruns, int *waves);
int pat_regular_destroy(pat_regular *h);

/* ------------------------------------------------------------------------
 * Reference input files (host only), so reference-shaped datasets replay.
 * pat_read_mold: moldyn's parameter file, which kmeans/moldyn.cu:193-201
 *   scanf()s from stdin: nB (4 nB^3 molecules), time steps, neighbour
 *   update period, cutoff, mesh and xyz file names, percent, start time,
 *   update frequency.
 * pat_read_ljin: miniMD's lj.in (miniMD/input.cpp:39-72): two header lines,
 *   then problem size, steps, bins, dt, temperature, density, reneighbour
 *   period, cutoffs, thermo period, one group per line.
 * pat_read_mesh: ReadInputGraph (kmeans/misc.c:76-145): leading '%' lines,
 *   "#nodes #edges", then per node its neighbours (1-based) up to a
 *   non-positive value or the line end; each edge once, from the node with
 *   the smaller number (neighbour >= own number, as the reference counts).
 *   Edges are 1-based, or 0-based with zero_based (the reference's
 *   CSTYLE).  edges = NULL: counts only.  The edge count must equal the
 *   header's.
 * pat_read_xyz: ReadCoord (misc.c:226-247): whitespace-separated triples,
 *   exactly num_nodes of them. */
typedef struct pat_mold {
    int box_cells, steps, neigh_update;
    double cutoff;
    char mesh[512], xyz[512];
    double percent;
    int start_time, update_freq;
} pat_mold;
typedef struct pat_ljin {
    int nx, ny, nz, ntimes, nbinx, nbiny, nbinz;
    double dt, t_request, rho;
    int every;
    double cutforce, cutneigh;
    int nstat;
} pat_ljin;
int pat_read_mold(const char *path, pat_mold *out);
int pat_read_ljin(const char *path, pat_ljin *out);
int pat_read_mesh(const char *path, int zero_based, int64_t *num_nodes, int64_t *num_edges, int32_t *edges,
                  int64_t capacity);
int pat_read_xyz(const char *path, int64_t num_nodes, double *xyz);

/* ------------------------------------------------------------------------
 * Serving loop.  No counterpart in the reference, whose mains run one
 * application per process (RuntimeInit .. RuntimeFinalize,
 * kmeans/include/runtime.cpp:4-13): here heat3D, LJ and kmeans jobs of any
 * size arrive over time, are admitted while their device-memory estimate
 * fits the budget and a slot is free, and run concurrently (one HIP stream
 * per slot), their inputs generated on the device from the job's seed.
 * ---------------------------------------------------------------------- */
#define PAT_JOB_HEAT3D 0 /* size = cube side; steps of u' = u + 0.1 (sum6 - 6u); interior U[0,1) */
#define PAT_JOB_LJ 1     /* size = FCC cells per side, 4..160 (rho 0.8442, jitter 0.05); steps of
                            velocity Verlet, cutoff 2.5, skin 0.3, neighbour list rebuilt on the
                            device every 20 steps, dt 0.005 */
#define PAT_JOB_KMEANS 2 /* size = points, dim, k (also the mixture's components, sigma 1,
                            centres in [-10, 10]^dim); steps = Lloyd iterations */
typedef struct pat_job {
    int kind;
    int64_t size;
    int dim, k;
    int steps;
    uint64_t seed;
    float *out;      /* optional host buffer for the full result, or NULL: heat3D the size^3
                        interior; LJ 4*size^3 x 3 positions; kmeans k x dim centres */
    int64_t out_len; /* floats available at out */
} pat_job;
typedef struct pat_job_result {
    int status;    /* PAT_OK or the job's error code */
    int kernel;    /* engine choice: kmeans PAT_KM_*, LJ lanes per row */
    double checksum;                          /* FP64 sum of the result (for a heat3D grid, of every cell) */
    double t_submit_ms, t_start_ms, t_end_ms; /* since pat_serve_start; end = result on the host */
    int64_t device_bytes;                     /* admission estimate held while it ran */
    int batch; /* jobs that ran fused with this one, itself included (1 = alone) */
} pat_job_result;
typedef struct pat_serve pat_serve;
#define PAT_SERVE_PLAIN 1 /* flag: kmeans on the plain FP32 kernel */
/* flag: every job runs alone.  Without it, a slot that admits an LJ or a
 * small heat3D job also admits the jobs of that kind waiting with the same
 * step count (each within the budget, up to 16 jobs) and advances them as
 * one batch:
 *   LJ (boxes of at least three neighbour bins, up to 2^21 atoms): pat_mdb,
 *   one fused force and update launch per step over every system, the steps
 *   between neighbour rebuilds replayed from a captured hipGraph;
 *   heat3D (sides up to 128, up to 2^23 cells): two sweeps per launch over
 *   every grid, the launches captured as one hipGraph.
 * Each job's result is the same bits either way; the batch's jobs share its
 * start and end times. */
#define PAT_SERVE_NO_FUSE 2
/* Starts 'slots' workers over an arena of hbm_budget device bytes. */
int pat_serve_start(int64_t hbm_budget, int slots, int flags, pat_serve **out);
/* Queues a job (non-blocking); *id identifies it. */
int pat_serve_submit(pat_serve *s, const pat_job *job, int64_t *id);
int pat_serve_poll(pat_serve *s, int64_t id, int *done, pat_job_result *res);
/* Blocks until job id has finished. */
int pat_serve_wait(pat_serve *s, int64_t id, pat_job_result *res);
/* Blocks until every submitted job has finished. */
int pat_serve_drain(pat_serve *s);
int pat_serve_stats(pat_serve *s, int64_t *submitted, int64_t *finished, int64_t *queued, int64_t *reserved_bytes,
                    int64_t *arena_in_use);
/* Runs what is queued, then stops the workers and frees the arena. */
int pat_serve_stop(pat_serve *s);

#ifdef __cplusplus
}
#endif

#endif /* PATTERN_H_ */
