/*
 * bp_hip.h -- C-ABI of the MI355X (gfx950) belief-propagation engine in
 * libbp_amd.so.  Plain C types only; every function returns 0 on success
 * and -1 on failure with the message in bp_last_error() (it never aborts,
 * whatever bp_set_abort_on_error() says).
 *
 * Two ways in:
 *  1. The drop-in graph API of graph.h (init_previous_edge,
 *     loopy_propagate_*, propagate_using_levels*, marginalize): each call
 *     uploads what it reads from the Graph_t, runs on the GPU and writes its
 *     results back, like the reference's CPU functions.
 *  2. A device-resident handle (bp_dev) for callers that keep the graph on
 *     the GPU across calls: upload once, iterate, download when needed.  A
 *     handle owns one HIP stream; calls on different handles may run from
 *     different threads.
 *
 * Reference interfaces replaced:
 *   loopy_propagate_until_hip       <- loopy_propagate_until_cuda
 *                                      (src/cuda_benchmark/belief-propagation.cu:529)
 *   loopy_propagate_until_hip_edge  <- loopy_propagate_until_cuda_edge
 *                                      (src/cuda_benchmark/belief-propagation.cu:664)
 *   bp_upload / bp_download         <- the cudaMalloc + cudaMemcpy blocks of
 *                                      belief-propagation.cu:567-600,634-636
 *   bp_loopy_init                   <- init_previous_edge (src/graph/graph.c:725)
 *   bp_loopy_iterate                <- loopy_propagate_one_iteration x n
 *                                      (graph.c:1121), plus the per-iteration
 *                                      delta of graph.c:1300-1309 in FP64
 *   bp_loopy_until                  <- loopy_propagate_until (graph.c:1278)
 *   bp_tree                         <- the tree driver sequence
 *                                      (src/c_benchmark/main.c:205-215)
 */
#ifndef BP_HIP_H_
#define BP_HIP_H_

#include <stdint.h>

#include "graph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bp_dev bp_dev;

/* Errors (shared with the drop-in API). */
const char *bp_last_error(void);
int bp_set_abort_on_error(int enabled); /* returns the previous setting */

/* Number of HIP devices visible (0 without a GPU); never fails. */
int bp_device_count(void);

/* Node numbering on the device for handles uploaded, and batch instances
 * added (bp_batch_add), from now on: 1 (the default) renumbers nodes in
 * Cuthill-McKee order for locality of the gathers and then, for handles,
 * groups each XCD's range of nodes by the set of XCDs that gather them
 * (kept where it lowers the number of 128-byte lines each XCD reads); 2
 * uses Cuthill-McKee order alone; 0 keeps the graph's order.  Results are
 * the same either way
 * (each node's in-edges keep the reference's order); only the FP64 delta's
 * summation order, and so its last bits, follow the numbering.  Returns the
 * previous setting. */
int bp_set_node_order(int mode);

/* GPU propagation with the contract of the reference's CUDA entry point:
 * the caller has run init_previous_edge (as src/cuda_benchmark/
 * belief-propagation.cu:1081 does); g is uploaded, iterated until the stop
 * test of graph.c:1313 fires on an FP64 delta (checked on the device after
 * every iteration) or max_iterations, and node_states and both message
 * arrays are downloaded with the pointer swaps the CPU path would have
 * made.  Returns what loopy_propagate_until returns (the index of the
 * converged iteration, or max_iterations); 0xffffffff on failure. */
unsigned int loopy_propagate_until_hip(Graph_t g, float convergence, unsigned int max_iterations);

/* The same through the edge-centric schedule of graph.c:1186-1276 with its
 * defects fixed (see graph.h). */
unsigned int loopy_propagate_until_hip_edge(Graph_t g, float convergence, unsigned int max_iterations);

/* ---- device-resident handle ---------------------------------------------- */

/* Uploads structure, tables, node states and both message arrays of g (its
 * CSR views must be set up).  device < 0 selects the current device. */
int bp_upload(Graph_t g, int device, bp_dev **out);
void bp_free(bp_dev *d);

/* Re-reads node_states and the message arrays from g (same structure). */
int bp_upload_state(bp_dev *d, Graph_t g);

/* Writes node_states, both message arrays and the message pointer swaps
 * back into g; for tree runs, edges_messages holds the tree messages. */
int bp_download(bp_dev *d, Graph_t g);

/* Node beliefs as [V][stride] floats (stride = bp_dev_info(...)[2]). */
int bp_download_states(bp_dev *d, float *out);

/* Restores the node states given to bp_upload / bp_upload_state (device to
 * device), so a run can start over without a host transfer. */
int bp_restore_states(bp_dev *d);

/* init_previous_edge on the device. */
int bp_loopy_init(bp_dev *d);

/* n loopy iterations; deltas (nullable) receives the FP64 delta of each.
 * The launches are replayed from a cached HIP graph. */
int bp_loopy_iterate(bp_dev *d, unsigned int n, double *deltas);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): n loopy iterations with the results of
 * bp_loopy_iterate, bit for bit, that call no hipMalloc, hipFree or other
 * allocating routine and create no HIP graph.  The FP64 delta partials and
 * deltas live in a fixed 4 MiB device array of the library, one per GPU:
 * the call fails with an error when the handle's delta partials per
 * iteration (one per wave of its grid: about one per 64-128 message edges)
 * plus n exceed 524,288 doubles, i.e. from about 30M message edges on.
 * Calls that overlap on
 * one GPU share that array and race.  deltas (nullable) receives the n FP64
 * deltas; without it the call returns without waiting for the GPU. */
int bp_loopy_iterate_noalloc(bp_dev *d, unsigned int n, double *deltas);

/* Like loopy_propagate_until: runs until the stop test fires or
 * max_iterations; *result receives the reference's return value. */
int bp_loopy_until(bp_dev *d, float convergence, unsigned int max_iterations, unsigned int *result);

/* n edge-centric iterations (loopy_propagate_edge_one_iteration,
 * graph.c:1186-1227, defects fixed). */
int bp_edge_iterate(bp_dev *d, unsigned int n);

/* Like loopy_propagate_until_edge (graph.c:1230-1276). */
int bp_edge_until(bp_dev *d, float convergence, unsigned int max_iterations, unsigned int *result);

/* Tree BP: start, up-sweep over levels 1..L-2, down-sweep over L-1..1,
 * marginalize.  The Graph_t given to bp_upload must have its levels set
 * (init_levels_to_nodes). */
int bp_tree(bp_dev *d);

/* Waits for all work queued on the handle. */
int bp_sync(bp_dev *d);

/* info[0]=V, info[1]=E (message edges), info[2]=state stride,
 * info[3]=max state count, info[4]=chunks of the node kernel,
 * info[5]=1 if state counts are mixed, info[6]=tree sub-waves,
 * info[7]=1 if the loopy kernels read one stored table per transposed pair */
int bp_dev_info(const bp_dev *d, uint64_t *info);

/* Algorithmic HBM bytes of one loopy iteration of this graph, the bytes an
 * iteration must move at minimum:
 *   distinct tables 4*x*y (a transposed reverse pair counts once)
 *   + per edge 4*(2x + y + 1): message write, previous message read (the
 *     delta), source buffer gather, source index
 *   + per node 12*n + 4: belief read and write, combine-buffer write, offset */
double bp_loopy_algorithmic_bytes(const bp_dev *d);

/* Times each of n iterations with HIP events taken at the start and end of
 * its kernel (hipExtLaunchKernelGGL, no graph): the kernel's duration alone.
 * ms[i] gets iteration i. */
int bp_profile_loopy(bp_dev *d, unsigned int n, float *ms);

#ifdef __cplusplus
}
#endif

#endif /* BP_HIP_H_ */
