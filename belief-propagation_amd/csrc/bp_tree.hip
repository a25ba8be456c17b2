/*
 * bp_tree.hip -- tree belief propagation on the GPU (src/graph/graph.c:
 * 376-605 and the driver sequence of src/c_benchmark/main.c:205-215).
 *
 * The reference walks each level sequentially: a node marks itself visited,
 * multiplies its in-messages and sends on every out-edge whose destination
 * is not yet visited (graph.c:461-522).  Two things depend on that order --
 * which sends happen, and which in-messages a node sees -- and both are
 * fixed by the level lists alone, so the host works them out once:
 *   - an out-edge v->w of a node at position p of the level is active iff w
 *     was not visited before the level, is not an earlier node of the level,
 *     and is not v itself;
 *   - a node must run after every earlier node of its level that sends to
 *     it; its sub-wave is one more than the largest sub-wave of those nodes.
 * Sub-waves of a level then run as parallel launches, one thread per node,
 * and give the sequential result bit for bit (a node reading an in-edge
 * whose source comes later in the level reads the old value there too,
 * because that send is inactive).  The start phase (graph.c:376-424) is one
 * wave that sends on every out-edge from the node's own state.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bp_dev.h"

using namespace bpk;

/* sub-waves of at most this many nodes run several to a launch, in one
 * workgroup of SMALL_BLOCK threads (1M-node tree: 1024 threads 0.252 ms,
 * 256 threads 0.260 ms; more than 1024 nodes per small wave was slower) */
#ifndef BP_TREE_SMALL_WAVE
#define BP_TREE_SMALL_WAVE 1024
#endif
#ifndef BP_TREE_SMALL_BLOCK
#define BP_TREE_SMALL_BLOCK 1024
#endif
/* threads per workgroup of a sub-wave launch (at most 256): small
 * workgroups spread a sub-wave of a few thousand nodes over more CUs
 * (1M-node tree: 64 threads 0.242 ms, 128 0.245 ms, 256 0.250 ms) */
#ifndef BP_TREE_WAVE_BLOCK
#define BP_TREE_WAVE_BLOCK 64
#endif
static_assert(BP_TREE_WAVE_BLOCK >= 64 && BP_TREE_WAVE_BLOCK <= 256, "k_tree_wave is bounded to 256 threads");
constexpr unsigned WAVE_BLOCK = BP_TREE_WAVE_BLOCK;
constexpr unsigned SMALL_WAVE = BP_TREE_SMALL_WAVE;
constexpr unsigned SMALL_BLOCK = BP_TREE_SMALL_BLOCK;

struct TreeEntry;

struct bp_tree_plan {
    TreeEntry *ent = nullptr;     /* entries, waves concatenated */
    unsigned *act_slot = nullptr; /* active out-slots */
    unsigned *wave_begin_d = nullptr;
    unsigned char *wave_start_d = nullptr;
    std::vector<unsigned> wave_begin; /* [waves+1] entry ranges */
    std::vector<unsigned char> wave_start;
    hipGraphExec_t graph[2] = {nullptr, nullptr}; /* per message buffer it runs on */
};

/* One node of a sub-wave: its device node, its in-slot range, its range
 * of active out-slots (never empty) and its state count, in one 32-byte
 * record (one load before the message and table reads instead of a chain
 * of index loads). */
struct TreeEntry {
    unsigned v, in_b, in_e, act_b, act_e, n, pad0, pad1;
};

struct TreeArgs {
    const unsigned char *slot_xy;
    const float *J;
    const unsigned *slot_tab; /* mixed state counts: table offsets in J (x rows, stride mixed_row_stride(S)) */
    const float *states;
    float *msg;
    const TreeEntry *ent;
    const unsigned *act_slot;
    const unsigned *wave_begin;       /* [waves+1] entry ranges */
    const unsigned char *wave_start;  /* [waves] 1: a start wave (sends from the node's own state) */
};

template <int S, bool MIXED>
__device__ __forceinline__ void tree_entry(const TreeArgs &a, unsigned entry, int start)
{
    constexpr int SP = Stride<S>::value;
    const uint4 *e4 = reinterpret_cast<const uint4 *>(a.ent + entry);
    uint4 lo = e4[0], hi = e4[1];
    unsigned v = lo.x, in_b = lo.y, in_e = lo.z, act_b = lo.w, act_e = hi.x;
    unsigned n = MIXED ? hi.y : (unsigned)S;
    float buf[S], m[S];
    if (start) {
#pragma unroll
        for (int k = 0; k < S; ++k)
            buf[k] = a.states[(size_t)v * SP + k];
    } else {
#pragma unroll
        for (int k = 0; k < S; ++k)
            buf[k] = (!MIXED || (unsigned)k < n) ? 1.0f : 0.0f;
        unsigned s = in_b;
        /* four rows loaded before they are combined (in slot order), so a
         * node's in-messages are fetched together instead of one latency each */
        for (; s + 4 <= in_e; s += 4) {
            float m4[4][S];
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int k = 0; k < S; ++k)
                    m4[j][k] = a.msg[(size_t)(s + j) * SP + k];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                combine<S, MIXED>(buf, m4[j], n);
        }
        for (; s < in_e; ++s) {
#pragma unroll
            for (int k = 0; k < S; ++k)
                m[k] = a.msg[(size_t)s * SP + k];
            combine<S, MIXED>(buf, m, n);
        }
    }
    for (unsigned q = act_b; q < act_e; ++q) {
        unsigned s = a.act_slot[q];
        unsigned x = S, y = S;
        if (MIXED) {
            unsigned xy = a.slot_xy[s];
            x = xy & 15u;
            y = xy >> 4;
        }
        if constexpr (MIXED) {
            float Jd[S * S];
            load_mixed_rows<S>(a.J + a.slot_tab[s], Jd);
            send_padded<S>(Jd, buf, x, y, m);
        } else {
            send<S, MIXED>(a.J + (size_t)s * S * S, buf, x, y, m);
        }
#pragma unroll
        for (int k = 0; k < SP; ++k)
            a.msg[(size_t)s * SP + k] = (k < S) ? m[k] : 0.0f;
    }
}

/* one sub-wave, one thread per node */
template <int S, bool MIXED>
__global__ __launch_bounds__(256) void k_tree_wave(TreeArgs a, unsigned begin, unsigned count, int start)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        tree_entry<S, MIXED>(a, begin + i, start);
}

/* consecutive small sub-waves [w0, w1) in one workgroup: each wave's nodes,
 * then a workgroup barrier (which orders the messages one wave writes
 * before the next wave reads them) */
template <int S, bool MIXED>
__global__ __launch_bounds__(SMALL_BLOCK) void k_tree_waves_small(TreeArgs a, unsigned w0, unsigned w1)
{
    for (unsigned w = w0; w < w1; ++w) {
        unsigned b = a.wave_begin[w], e = a.wave_begin[w + 1];
        int start = a.wave_start[w];
        for (unsigned entry = b + threadIdx.x; entry < e; entry += blockDim.x)
            tree_entry<S, MIXED>(a, entry, start);
        __syncthreads();
    }
}

/* graph.c:540-596: the state is replaced by the product of the in-messages */
template <int S, bool MIXED>
__global__ __launch_bounds__(256) void k_tree_marginalize(const unsigned *in_off, const unsigned char *nvars,
                                                          float *states, const float *msg, unsigned V)
{
    constexpr int SP = Stride<S>::value;
    unsigned v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V)
        return;
    unsigned n = MIXED ? nvars[v] : S;
    float t[S], m[S], st[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
        t[k] = 1.0f;
        st[k] = states[(size_t)v * SP + k];
    }
    unsigned b = in_off[v], e = in_off[v + 1];
    for (unsigned s = b; s < e; ++s) {
#pragma unroll
        for (int k = 0; k < S; ++k)
            m[k] = msg[(size_t)s * SP + k];
        combine<S, MIXED>(t, m, n);
    }
    if (e > b) {
#pragma unroll
        for (int k = 0; k < S; ++k)
            if (!MIXED || (unsigned)k < n)
                st[k] = t[k];
    }
    normalise<S, MIXED>(st, n);
#pragma unroll
    for (int k = 0; k < S; ++k)
        if (!MIXED || (unsigned)k < n)
            states[(size_t)v * SP + k] = st[k];
}

/* ------------------------------------------------------------------------ */
/* plan                                                                      */

namespace {

struct PlanBuilder {
    Graph_t g;
    const std::vector<unsigned> &edge_slot;
    std::vector<unsigned> entries, act_off, act_slot, wave_begin;
    std::vector<unsigned char> wave_start;
    std::vector<unsigned> pos;  /* 1 + position in the current level, 0 otherwise */
    std::vector<unsigned> sub;

    PlanBuilder(Graph_t gg, const std::vector<unsigned> &es) : g(gg), edge_slot(es)
    {
        pos.assign(g->current_num_vertices, 0);
        sub.assign(g->current_num_vertices, 0);
        act_off.push_back(0);
        wave_begin.push_back(0);
    }

    unsigned src_end(unsigned v) const { return bp_csr_end(g, g->src_nodes_to_edges_node_list, v); }
    unsigned dst_end(unsigned v) const { return bp_csr_end(g, g->dest_nodes_to_edges_node_list, v); }

    /* graph.c:376-424: level 0 sends on every out-edge */
    void start(std::vector<char> &vis, unsigned b, unsigned e)
    {
        for (unsigned k = b; k < e; ++k) {
            unsigned v = g->levels_to_nodes[k];
            vis[v] = 1;
            entries.push_back(v);
            for (unsigned i = g->src_nodes_to_edges_node_list[v]; i < src_end(v); ++i)
                act_slot.push_back(edge_slot[g->src_nodes_to_edges_edge_list[i]]);
            act_off.push_back((unsigned)act_slot.size());
        }
        close_wave(1);
    }

    void close_wave(unsigned char is_start)
    {
        if (entries.size() != wave_begin.back()) {
            wave_begin.push_back((unsigned)entries.size());
            wave_start.push_back(is_start);
        }
    }

    /* graph.c:524-538 for one level */
    void level(std::vector<char> &vis, unsigned b, unsigned e)
    {
        unsigned maxsub = 0;
        for (unsigned k = b; k < e; ++k)
            pos[g->levels_to_nodes[k]] = k - b + 1;
        std::vector<std::vector<unsigned>> act(e - b);
        for (unsigned k = b; k < e; ++k) {
            unsigned v = g->levels_to_nodes[k], p = k - b;
            unsigned sv = 0;
            for (unsigned i = g->dest_nodes_to_edges_node_list[v]; i < dst_end(v); ++i) {
                unsigned u = g->edges_src_index[g->dest_nodes_to_edges_edge_list[i]];
                if (pos[u] && pos[u] - 1 < p && !vis[v])
                    sv = std::max(sv, sub[u] + 1);
            }
            sub[v] = sv;
            maxsub = std::max(maxsub, sv);
            for (unsigned i = g->src_nodes_to_edges_node_list[v]; i < src_end(v); ++i) {
                unsigned edge = g->src_nodes_to_edges_edge_list[i];
                unsigned w = g->edges_dest_index[edge];
                bool earlier = pos[w] && pos[w] - 1 < p;
                if (!vis[w] && !earlier && w != v)
                    act[p].push_back(edge_slot[edge]);
            }
        }
        for (unsigned w = 0; w <= maxsub; ++w) {
            for (unsigned k = b; k < e; ++k) {
                unsigned v = g->levels_to_nodes[k];
                if (sub[v] != w)
                    continue;
                entries.push_back(v);
                for (unsigned s : act[k - b])
                    act_slot.push_back(s);
                act_off.push_back((unsigned)act_slot.size());
            }
            close_wave(0);
        }
        for (unsigned k = b; k < e; ++k) {
            unsigned v = g->levels_to_nodes[k];
            pos[v] = 0;
            vis[v] = 1;
        }
    }

    unsigned level_end(unsigned l) const
    {
        return (l + 1 == g->num_levels) ? 2 * g->current_num_vertices : g->levels_to_nodes[l + 1];
    }
};

int upload_plan(bp_dev *d, PlanBuilder &pb, bp_tree_plan **out)
{
    /* Entry records in device numbering.  The entries of one sub-wave are
     * independent: each writes only its own out-slots and reads only slots
     * written by earlier sub-waves.  So an entry that sends on no slot (its
     * product would be read by nobody) is left out, and each sub-wave is
     * ordered by the first slot its entries write: consecutive threads then
     * read adjacent tables and write adjacent messages.  A sub-wave left
     * empty is dropped; the order of the rest is kept. */
    std::vector<TreeEntry> ent;
    std::vector<unsigned> act, wbeg{0}, idx;
    std::vector<unsigned char> wstart;
    ent.reserve(pb.entries.size());
    act.reserve(pb.act_slot.size());
    for (size_t w = 0; w + 1 < pb.wave_begin.size(); ++w) {
        idx.clear();
        for (unsigned i = pb.wave_begin[w]; i < pb.wave_begin[w + 1]; ++i)
            if (pb.act_off[i + 1] > pb.act_off[i])
                idx.push_back(i);
        std::sort(idx.begin(), idx.end(), [&](unsigned x, unsigned y) {
            return pb.act_slot[pb.act_off[x]] < pb.act_slot[pb.act_off[y]];
        });
        for (unsigned i : idx) {
            unsigned v = d->perm[pb.entries[i]];
            unsigned ab = (unsigned)act.size();
            act.insert(act.end(), pb.act_slot.begin() + pb.act_off[i], pb.act_slot.begin() + pb.act_off[i + 1]);
            ent.push_back(TreeEntry{v, d->in_off_h[v], d->in_off_h[v + 1], ab, (unsigned)act.size(), d->node_n[v], 0, 0});
        }
        if (ent.size() != wbeg.back()) {
            wbeg.push_back((unsigned)ent.size());
            wstart.push_back(pb.wave_start[w]);
        }
    }
    size_t ne = ent.size();
    bp_tree_plan *p = new bp_tree_plan();
    p->wave_begin = wbeg;
    p->wave_start = wstart;
    if (dmalloc(&p->ent, ne) || dmalloc(&p->act_slot, act.size()) || dmalloc(&p->wave_begin_d, wbeg.size()) ||
        dmalloc(&p->wave_start_d, wstart.size())) {
        bp_tree_plan_free(p);
        return -1;
    }
    bool ok = (ne == 0 || hipMemcpy(p->ent, ent.data(), ne * sizeof(TreeEntry), hipMemcpyHostToDevice) == hipSuccess) &&
              (act.empty() || hipMemcpy(p->act_slot, act.data(), act.size() * 4, hipMemcpyHostToDevice) == hipSuccess) &&
              hipMemcpy(p->wave_begin_d, wbeg.data(), wbeg.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
              (wstart.empty() ||
               hipMemcpy(p->wave_start_d, wstart.data(), wstart.size(), hipMemcpyHostToDevice) == hipSuccess);
    if (!ok) {
        bp_set_error("upload of the tree schedule failed");
        bp_tree_plan_free(p);
        return -1;
    }
    *out = p;
    return 0;
}

int check_levels(Graph_t g)
{
    unsigned V = g->current_num_vertices;
    if (g->num_levels == 0 || g->num_levels > V + 1) {
        bp_set_error("levels are not set (call init_levels_to_nodes)");
        return -1;
    }
    for (unsigned l = 0; l < g->num_levels; ++l) {
        unsigned b = g->levels_to_nodes[l];
        unsigned e = (l + 1 == g->num_levels) ? 2 * V : g->levels_to_nodes[l + 1];
        if (b < V || b > e || e > 2 * V) {
            bp_set_error("levels_to_nodes is inconsistent (level %u)", l);
            return -1;
        }
    }
    std::vector<unsigned char> seen(V, 0);
    for (unsigned k = V; k < 2 * V; ++k) {
        unsigned v = g->levels_to_nodes[k];
        if (v >= V) {
            bp_set_error("levels_to_nodes names node %u of %u", v, V);
            return -1;
        }
        /* a node listed twice would break the sub-wave order the plan
         * derives from each node's position in its level */
        if (seen[v]) {
            bp_set_error("levels_to_nodes lists node %u twice", v);
            return -1;
        }
        seen[v] = 1;
    }
    return 0;
}

} // namespace

void bp_tree_plan_free(bp_tree_plan *p)
{
    if (!p)
        return;
    for (hipGraphExec_t &x : p->graph)
        if (x)
            (void)hipGraphExecDestroy(x);
    if (p->ent)
        (void)hipFree(p->ent);
    if (p->wave_begin_d)
        (void)hipFree(p->wave_begin_d);
    if (p->wave_start_d)
        (void)hipFree(p->wave_start_d);
    if (p->act_slot)
        (void)hipFree(p->act_slot);
    delete p;
}

unsigned long long bp_tree_plan_waves(const bp_tree_plan *p) { return p ? p->wave_begin.size() - 1 : 0; }

/* The whole driver sequence: start, levels 1..L-2, reset, levels L-1..1. */
int bp_tree_plan_build(bp_dev *d, Graph_t g)
{
    if (check_levels(g))
        return -1;
    unsigned V = g->current_num_vertices;
    PlanBuilder pb(g, d->edge_slot);
    std::vector<char> vis(V, 0);
    unsigned e0 = (g->num_levels == 1) ? 2 * V : g->levels_to_nodes[1];
    pb.start(vis, g->levels_to_nodes[0], e0);
    for (unsigned l = 1; l + 1 < g->num_levels; ++l)
        pb.level(vis, g->levels_to_nodes[l], pb.level_end(l));
    std::fill(vis.begin(), vis.end(), 0);
    for (unsigned l = g->num_levels - 1; l > 0; --l)
        pb.level(vis, g->levels_to_nodes[l], pb.level_end(l));
    bp_tree_plan_free(d->tree);
    d->tree = nullptr;
    return upload_plan(d, pb, &d->tree);
}

/* ------------------------------------------------------------------------ */
/* launches                                                                  */

static TreeArgs tree_args(bp_dev *d, const bp_tree_plan *p)
{
    TreeArgs a;
    a.slot_xy = d->slot_xy;
    a.J = d->J;
    a.slot_tab = d->mixed ? d->slot_toff : nullptr;
    a.states = d->states;
    a.msg = d->msg[a_index(d)];
    a.ent = p->ent;
    a.act_slot = p->act_slot;
    a.wave_begin = p->wave_begin_d;
    a.wave_start = p->wave_start_d;
    return a;
}

/* the sub-waves in order: a wave of more than SMALL_WAVE nodes is one
 * launch, one thread per node; a run of smaller ones is one launch of one
 * workgroup */
template <int S, bool MIXED> static void waves_s(bp_dev *d, const bp_tree_plan *p)
{
    TreeArgs a = tree_args(d, p);
    size_t nw = p->wave_begin.size() - 1;
    auto size = [&](size_t w) { return p->wave_begin[w + 1] - p->wave_begin[w]; };
    for (size_t w = 0; w < nw;) {
        if (size(w) > SMALL_WAVE) {
            unsigned b = p->wave_begin[w], n = size(w);
            hipLaunchKernelGGL((k_tree_wave<S, MIXED>), dim3((n + WAVE_BLOCK - 1) / WAVE_BLOCK), dim3(WAVE_BLOCK), 0,
                               d->stream, a, b, n, (int)p->wave_start[w]);
            ++w;
            continue;
        }
        size_t w1 = w + 1;
        while (w1 < nw && size(w1) <= SMALL_WAVE)
            ++w1;
        hipLaunchKernelGGL((k_tree_waves_small<S, MIXED>), dim3(1), dim3(SMALL_BLOCK), 0, d->stream, a, (unsigned)w,
                           (unsigned)w1);
        w = w1;
    }
}

template <int S, bool MIXED> static void marg_s(bp_dev *d)
{
    if (d->V == 0)
        return;
    hipLaunchKernelGGL((k_tree_marginalize<S, MIXED>), dim3((d->V + 255) / 256), dim3(256), 0, d->stream, d->in_off,
                       d->nvars, d->states, d->msg[a_index(d)], d->V);
}

template <bool MIXED> static void dispatch(bp_dev *d, const bp_tree_plan *p, bool marginal)
{
#define BP_TREE_CASE(N)                \
    case N:                            \
        if (p)                         \
            waves_s<N, MIXED>(d, p);   \
        if (marginal)                  \
            marg_s<N, MIXED>(d);       \
        break;
    switch (d->S) {
        BP_TREE_CASE(1)
        BP_TREE_CASE(2)
        BP_TREE_CASE(3)
        BP_TREE_CASE(4)
        BP_TREE_CASE(5)
        BP_TREE_CASE(6)
        BP_TREE_CASE(7)
        BP_TREE_CASE(8)
        BP_TREE_CASE(9)
    default:
        if (p)
            waves_s<10, MIXED>(d, p);
        if (marginal)
            marg_s<10, MIXED>(d);
        break;
    }
#undef BP_TREE_CASE
}

static void enqueue(bp_dev *d, const bp_tree_plan *p, bool marginal)
{
    if (d->mixed)
        dispatch<true>(d, p, marginal);
    else
        dispatch<false>(d, p, marginal);
}

extern "C" int bp_tree(bp_dev *d)
{
    bp_clear_error();
    if (!d) {
        bp_set_error("bp_tree: NULL handle");
        return -1;
    }
    if (!d->tree) {
        bp_set_error("bp_tree: the graph had no levels when it was uploaded (call init_levels_to_nodes first)");
        return -1;
    }
    HIP_TRY(hipSetDevice(d->device));
    /* tree BP works on the buffer mirroring the host's edges_messages */
    hipGraphExec_t &exec = d->tree->graph[a_index(d)];
    if (!exec) {
        hipGraph_t graph;
        HIP_TRY(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
        enqueue(d, d->tree, true);
        HIP_TRY(hipStreamEndCapture(d->stream, &graph));
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            exec = nullptr;
            bp_set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e));
            return -1;
        }
    }
    HIP_TRY(hipGraphLaunch(exec, d->stream));
    d->buf_valid = false; /* messages and states changed under the loopy combine buffer */
    return 0;
}

/* ------------------------------------------------------------------------ */
/* drop-in per-level entry points                                            */

extern "C" bp_dev *bp_dropin_device(Graph_t g, const char *who);

static void run_partial(Graph_t g, const char *who, int mode, unsigned level)
{
    bp_dev *d = bp_dropin_device(g, who);
    if (!d)
        return;
    if (mode != 2 && check_levels(g)) {
        bp_fail("%s: %s", who, bp_last_error());
        return;
    }
    bp_tree_plan *p = nullptr;
    std::vector<char> vis(g->visited, g->visited + g->current_num_vertices);
    if (mode != 2) {
        PlanBuilder pb(g, d->edge_slot);
        if (mode == 0) {
            unsigned e0 = (g->num_levels == 1) ? 2 * g->current_num_vertices : g->levels_to_nodes[1];
            pb.start(vis, g->levels_to_nodes[0], e0);
        } else {
            if (level >= g->num_levels) {
                bp_fail("%s: level %u of %u", who, level, g->num_levels);
                return;
            }
            pb.level(vis, g->levels_to_nodes[level], pb.level_end(level));
        }
        if (upload_plan(d, pb, &p)) {
            bp_fail("%s: %s", who, bp_last_error());
            return;
        }
    }
    if (hipSetDevice(d->device) != hipSuccess) {
        bp_tree_plan_free(p);
        bp_fail("%s: hipSetDevice failed", who);
        return;
    }
    enqueue(d, p, mode == 2);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(d->stream);
    bp_tree_plan_free(p);
    if (e != hipSuccess) {
        bp_fail("%s: %s", who, hipGetErrorString(e));
        return;
    }
    d->buf_valid = false;
    if (bp_download(d, g)) {
        bp_fail("%s: %s", who, bp_last_error());
        return;
    }
    /* the visited flags the sequential walk leaves behind */
    memcpy(g->visited, vis.data(), g->current_num_vertices);
}

extern "C" void propagate_using_levels_start(Graph_t g) { run_partial(g, "propagate_using_levels_start", 0, 0); }

extern "C" void propagate_using_levels(Graph_t g, unsigned int level)
{
    run_partial(g, "propagate_using_levels", 1, level);
}

extern "C" void marginalize(Graph_t g) { run_partial(g, "marginalize", 2, 0); }
