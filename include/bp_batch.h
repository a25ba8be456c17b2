/*
 * bp_batch.h -- many networks in one launch (C-ABI of libbp_amd.so).
 *
 * No counterpart in the reference, which propagates one network per call
 * (src/graph/graph.c:1121-1183, src/cuda_benchmark/belief-propagation.cu:
 * 529).  A batch holds many independent networks -- typically several
 * evidence sets of each of many networks -- and runs every one of them
 * through init_previous_edge (graph.c:725-751) and the node-centric loopy
 * iteration with one kernel launch per iteration over all of them; the
 * iteration loop is replayed from a HIP graph.  Each network's results are
 * bit-identical to running it alone (bp_upload + bp_loopy_init +
 * bp_loopy_iterate): networks never share nodes, and every node's
 * arithmetic is the same.
 *
 * Tables are pooled by content: an edge table equal to an earlier one (the
 * same arc under another evidence set) points at it.  The pool keeps each
 * table in the orientation its edges read it -- a reverse edge's table, the
 * transpose of its forward edge's, is a second entry -- with rows padded to
 * four-float multiples so the kernels read them with aligned vector loads.
 * Instances over the same nodes (one network under several evidence sets)
 * are scheduled together, node range by node range, so their shared tables
 * are read while cached.
 *
 * The kernel's tile shape (slots per lane, i.e. edges per thread and chunk
 * size; lanes per slot, i.e. threads per edge) and scheduling (static
 * ranges, or a work queue handing out blocks of chunks) are chosen by
 * bp_batch_tune, which times the candidates on the batch itself and can
 * keep its choices in a JSON file.
 *
 * Every function returns 0, or -1 with the message in bp_last_error().
 */
#ifndef BP_BATCH_H_
#define BP_BATCH_H_

#include <stdint.h>

#include "graph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bp_batch bp_batch;

int bp_batch_create(bp_batch **out);
void bp_batch_free(bp_batch *b);

/* Appends a network (CSR views set up) as the next instance and copies what
 * the batch needs; g may be destroyed afterwards.  index (nullable)
 * receives the instance number. */
int bp_batch_add(bp_batch *b, Graph_t g, unsigned int *index);

/* Moves the batch to the device (device < 0: the current one); no more
 * networks can be added afterwards. */
int bp_batch_upload(bp_batch *b, int device);

/* Kernel shape: slots_per_lane 1 or 2 (edges per thread; 2 needs six or
 * fewer states); lanes_per_slot 1, 2, 4 or 8 (threads per edge, each
 * forming every L-th row of the message; more than 1 needs four or more
 * states and slots_per_lane 1); queue_block 0 for static chunk ranges, else
 * the number of chunks a wave takes from the work queue at a time. */
int bp_batch_set_config(bp_batch *b, unsigned int slots_per_lane, unsigned int lanes_per_slot,
                        unsigned int queue_block);
int bp_batch_get_config(const bp_batch *b, unsigned int *slots_per_lane, unsigned int *lanes_per_slot,
                        unsigned int *queue_block);

/* Times every candidate shape over `iterations` iterations and keeps the
 * fastest.  cache_path (nullable): a JSON file of earlier choices keyed by
 * the device and the batch's sizes; a hit skips the timing, a miss adds an
 * entry. */
int bp_batch_tune(bp_batch *b, unsigned int iterations, const char *cache_path);

/* Every instance from its uploaded beliefs: init_previous_edge. */
int bp_batch_init(bp_batch *b);

/* n iterations of every instance; deltas (nullable) receives [n][instances]
 * FP64 deltas (graph.c:1300-1309, summed in FP64). */
int bp_batch_iterate(bp_batch *b, unsigned int n, double *deltas);

/* Writes instance `index`'s beliefs and both message arrays into g (the
 * network that was added, or one with the same structure), as if
 * init_previous_edge and the iterations since had run on g itself,
 * message pointer swaps included. */
int bp_batch_download(bp_batch *b, unsigned int index, Graph_t g);

int bp_batch_sync(bp_batch *b);

/* info[0]=instances, info[1]=nodes, info[2]=message edges, info[3]=largest
 * state count, info[4]=distinct table floats, info[5]=chunks of the current
 * shape, info[6]=algorithmic bytes of one iteration (see
 * bp_loopy_algorithmic_bytes; tables counted once per distinct table, a
 * table and its transpose being one); info[4] counts the same way */
int bp_batch_info(const bp_batch *b, uint64_t *info);

/* Times n iterations of the current shape with HIP events around the
 * iteration kernels only; *ms receives the mean per iteration. */
int bp_batch_profile(bp_batch *b, unsigned int n, float *ms);

#ifdef __cplusplus
}
#endif

#endif /* BP_BATCH_H_ */
