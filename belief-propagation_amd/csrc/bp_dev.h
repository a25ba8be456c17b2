/*
 * bp_dev.h -- the device-resident graph shared by the loopy (bp_device.hip)
 * and tree (bp_tree.hip) translation units of libbp_amd.so.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "bp_internal.h"
#include "bp_kernels.h"
#include "../../include/bp_hip.h"

extern "C" void bp_set_error(const char *fmt, ...);

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t err_ = (expr);                                                            \
        if (err_ != hipSuccess) {                                                            \
            bp_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(err_), __FILE__, \
                         __LINE__);                                                          \
            return -1;                                                                       \
        }                                                                                    \
    } while (0)

struct bp_tree_plan;

struct bp_dev {
    int device = 0;
    hipStream_t stream = nullptr;
    unsigned V = 0, E = 0, S = 0, SP = 0;
    bool mixed = false;

    unsigned *in_off = nullptr;       /* [V+1] slot range per node */
    unsigned *slot_src = nullptr;     /* [E] */
    unsigned char *slot_xy = nullptr; /* [E] */
    float *J = nullptr;               /* [E][S*S]; mixed state counts: each slot's x*y floats, packed */
    unsigned *slot_toff = nullptr;    /* mixed state counts: [E] offset of each slot's table in J */
    /* uniform state counts, transposed pairs shared (the loopy kernels
     * only; tree BP and the adjoint read J): slot_pool[s] = index of the
     * slot's table in Jp, bit 31 set when the slot reads it transposed */
    bool shared = false;
    float *Jp = nullptr;              /* [<= E][S*S] */
    unsigned *slot_pool = nullptr;    /* [E] */
    std::vector<unsigned> partner;    /* [E] by slot: the slot of the reverse edge, or ~0u */
    std::vector<unsigned> toff;       /* the same on the host */
    std::vector<unsigned> in_off_h;   /* [V+1] in_off on the host */
    size_t jfloats = 0;               /* floats in J */
    float *h_stage = nullptr;         /* pinned host staging for uploads and downloads */
    size_t h_stage_n = 0;             /* its capacity in floats */
    unsigned char *nvars = nullptr;   /* [V] */
    float *states = nullptr;          /* [V][SP] */
    float *buf[2] = {nullptr, nullptr};
    float *msg[2] = {nullptr, nullptr};
    float *states0 = nullptr;         /* [V][SP] the states given at upload, for bp_restore_states */
    uint2 *chunk = nullptr;     /* [nchunks+1] (first node, first slot) */
    unsigned nchunks = 0, G = 0;
    unsigned NP = 0;            /* delta partials per iteration (one per wave) */
    unsigned part_cap = 1;      /* iterations whose delta partials fit the buffer */

    double *partials = nullptr; /* [part_cap][NP] */
    double *deltas = nullptr;   /* [max(part_cap,16)] */
    double *last = nullptr;
    unsigned *stop = nullptr;

    /* host mirrors for upload / download */
    std::vector<unsigned> perm;      /* node -> device node (the device numbers nodes in its own order) */
    std::vector<unsigned> inv;       /* device node -> node */
    std::vector<unsigned> slot_edge; /* slot -> reference edge index */
    std::vector<unsigned> edge_slot; /* reference edge index -> slot */
    std::vector<unsigned> node_n;    /* state count per device node */
    std::vector<unsigned> edge_x;    /* x per reference edge */
    double alg_bytes = 0.0;

    /* run state */
    int latest = 0;       /* msg[latest] holds the newest messages */
    int bcur = 0;         /* buf[bcur] holds the combine buffer of the next iteration */
    bool buf_valid = false;
    unsigned swaps = 0;   /* iterations since upload: host pointer swaps owed */
    /* The host keeps two message arrays, edges_messages (A) and
     * last_edges_messages, and swaps which one current_edge_messages points
     * at; tree BP always works on A.  cur_is_a records whether the host's
     * current pointer pointed at A at upload / last download. */
    bool cur_is_a = true;

    std::map<unsigned long long, hipGraphExec_t> graphs;

    bp_tree_plan *tree = nullptr;
};

inline unsigned stride_of(unsigned S) { return S <= 2 ? S : S <= 4 ? 4 : S <= 8 ? 8 : 10; }

template <typename T> static int dmalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0)
        count = 1;
    HIP_TRY(hipMalloc((void **)p, count * sizeof(T)));
    return 0;
}

/* Loopy kernel launch shared with the adjoint (bp_diff.hip): the handle's
 * arrays as kernel arguments, and one launch of k_loopy<S, MIXED, mode> on
 * stream st over the handle's chunks (ev: optional start/stop events). */
bpk::LoopyArgs bp_make_args(bp_dev *d);
void bp_launch_loopy(bp_dev *d, int mode, const bpk::LoopyArgs &a, hipStream_t st, const hipEvent_t *ev = nullptr);
int bp_upload_graph(Graph_t g, int device, bp_dev **out);

/* Device node numbering: bp_node_order_mode() is what bp_set_node_order
 * chose (1: Cuthill-McKee, 0: the graph's own order); bp_cm_order fills
 * order[i] = the graph node placed at device position i. */
int bp_node_order_mode();
void bp_cm_order(Graph_t g, std::vector<unsigned> &order);

/* Chunks over the slot ranges in_off[V+1]: consecutive nodes with at most
 * cap in-slots and cap nodes, or one node with more; entry c is chunk c's
 * (first node, first slot), and a sentinel (V, in_off[V]) closes the list. */
void bp_build_chunks(const std::vector<unsigned> &in_off, unsigned V, unsigned cap, std::vector<uint2> &chunks);

/* tree BP (bp_tree.hip) */
int bp_tree_plan_build(bp_dev *d, Graph_t g);
void bp_tree_plan_free(bp_tree_plan *p);
unsigned long long bp_tree_plan_waves(const bp_tree_plan *p);
/* The device buffer mirroring the host's edges_messages array. */
inline int a_index(const bp_dev *d)
{
    bool cur_points_a = d->cur_is_a ^ ((d->swaps & 1u) != 0);
    return cur_points_a ? (d->latest ^ 1) : d->latest;
}
