/*
 * bp_internal.h -- private state shared by the host-side translation units of
 * libbp_amd.so.  Not installed; callers use include/graph.h and
 * include/bp_hip.h.
 */
#ifndef BP_INTERNAL_H_
#define BP_INTERNAL_H_

#include <stddef.h>
#include <stdint.h>

#include "../../include/graph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Every Graph_t handed out by create_graph() is the first member of this
 * record, so the public pointer converts back with bp_impl(). */
struct bp_graph_impl {
    struct graph pub;

    /* name -> first node index with that name (open addressing, power of 2) */
    unsigned int *name_slots;
    unsigned int name_cap;
    unsigned int name_count; /* number of nodes hashed so far */

    /* bumped on every structural change; the device cache compares it */
    uint64_t generation;

    /* device mirror used by the drop-in propagation entry points */
    void *dev_cache;
    uint64_t dev_cache_generation;
};

static inline struct bp_graph_impl *bp_impl(Graph_t g) { return (struct bp_graph_impl *)g; }

/* Error reporting (bp_error.c).  bp_fail() records the message for
 * bp_last_error(); in abort mode (the default, matching the reference's
 * asserts) it also prints it and aborts. */
void bp_fail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void bp_clear_error(void);
/* records a message without aborting (the C-ABI of bp_hip.h) */
void bp_set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
int bp_has_error(void);

/* End offset of node v's slice in a reference-style CSR: the reference keeps
 * no sentinel, the last node ends at current_num_edges (graph.c:395-400). */
static inline unsigned int bp_csr_end(const struct graph *g, const unsigned int *node_list, unsigned int v)
{
    return (v + 1 >= g->current_num_vertices) ? g->current_num_edges : node_list[v + 1];
}

/* Releases the device mirror of a graph (bp_device.hip); NULL-safe. */
void bp_dev_cache_release(struct bp_graph_impl *impl);

#ifdef __cplusplus
}
#endif

#endif /* BP_INTERNAL_H_ */
