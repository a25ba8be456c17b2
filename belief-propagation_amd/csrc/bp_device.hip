/*
 * bp_device.hip -- device-resident graph, loopy BP launches, the C-ABI of
 * include/bp_hip.h and the GPU-executed propagation entry points of the
 * drop-in graph API (include/graph.h).
 *
 * Replaces the host orchestration of src/cuda_benchmark/belief-propagation.cu
 * (:529-661 node variant, :664-783 edge variant): one upload of a
 * destination-grouped structure-of-arrays graph, one kernel per loopy
 * iteration replayed from a cached HIP graph, the FP64 delta reduced once per
 * batch (fixed-iteration runs) or after every iteration with an on-device
 * stop flag (convergence runs), and one download.
 */
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "bp_dev.h"
#include "../../include/bp_hip.h"

using namespace bpk;

/* Sums the G per-wave partials of each of `count` launches, one workgroup
 * per launch (fixed order, so the result is reproducible).  until != 0 (count
 * 1): evaluates the stop test of graph.c:1313 for launch `index` against
 * *last and raises *stop. */
__global__ __launch_bounds__(BLOCK) void k_finish(const double *partials, unsigned G, unsigned count,
                                                 double *deltas, double *last, unsigned *stop, float conv,
                                                 int until, unsigned index)
{
    __shared__ double red[BLOCK];
    if (until && *stop)
        return;
    unsigned it = blockIdx.x;
    if (it >= count)
        return;
    const double *p = partials + (size_t)it * G;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    unsigned i = threadIdx.x;
    for (; i + 3 * BLOCK < G; i += 4 * BLOCK) {
        acc[0] += p[i];
        acc[1] += p[i + BLOCK];
        acc[2] += p[i + 2 * BLOCK];
        acc[3] += p[i + 3 * BLOCK];
    }
    for (; i < G; i += BLOCK)
        acc[0] += p[i];
    red[threadIdx.x] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    for (unsigned w = BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double d = red[0];
        deltas[until ? index : it] = d;
        if (until) {
            double prev = *last;
            double change = d - prev;
            if (change < 0.0)
                change = -change;
            if (d < (double)conv || change < (double)conv)
                *stop = index + 1;
            *last = d;
        }
    }
}

__global__ void k_copy_floats(const float *__restrict__ src, float *__restrict__ dst, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = src[i];
}

__global__ void k_until_reset(double *last, unsigned *stop)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *last = -1.0;
        *stop = 0;
    }
}

/* One message on one edge (the drop-in send_message). */
template <int S>
__global__ void k_send_one(const float *J, const float *b, unsigned x, unsigned y, float *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float m[S];
        send<S, true>(J, b, x, y, m);
        for (unsigned i = 0; i < x; ++i)
            out[i] = m[i];
    }
}



static void free_dev_arrays(bp_dev *d)
{
    void *ptrs[] = {d->Jp, d->slot_pool, d->slot_toff, d->in_off, d->slot_src, d->slot_xy, d->J, d->nvars, d->states, d->buf[0], d->buf[1],
                    d->msg[0], d->msg[1], d->states0, d->chunk, d->partials, d->deltas, d->last, d->stop};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
}

extern "C" void bp_free(bp_dev *d)
{
    if (!d)
        return;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(d->device);
    if (d->stream)
        (void)hipStreamSynchronize(d->stream);
    for (auto &kv : d->graphs)
        (void)hipGraphExecDestroy(kv.second);
    bp_tree_plan_free(d->tree);
    free_dev_arrays(d);
    if (d->h_stage)
        (void)hipHostFree(d->h_stage);
    d->h_stage = nullptr;
    if (d->stream)
        (void)hipStreamDestroy(d->stream);
    (void)hipSetDevice(prev);
    delete d;
}

extern "C" int bp_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

/* f(begin, end) over [0, n) split across up to 16 host threads (the
 * graph's host arrays are MAX_STATES-strided: the repacking is bound by
 * memory latency, not arithmetic) */
template <typename F> static void par_for(size_t n, F f)
{
    unsigned T = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (n < (1u << 16) || T == 1) {
        f((size_t)0, n);
        return;
    }
    size_t step = (n + T - 1) / T;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t) {
        size_t a = t * step, b = std::min(n, a + step);
        if (a < b)
            th.emplace_back(f, a, b);
    }
    for (auto &x : th)
        x.join();
}

/* Chunks: consecutive nodes with at most cap in-slots and cap nodes (cap =
 * 64 x slots per lane), or one node with more in-slots.  Each entry holds
 * the chunk's first node and first slot. */
void bp_build_chunks(const std::vector<unsigned> &in_off, unsigned V, unsigned cap, std::vector<uint2> &chunks)
{
    chunks.clear();
    unsigned v = 0;
    while (v < V) {
        chunks.push_back(make_uint2(v, in_off[v]));
        unsigned s0 = in_off[v];
        unsigned w = v;
        while (w < V && w - v < cap && in_off[w + 1] - s0 <= cap)
            ++w;
        if (w == v)
            w = v + 1; /* a node with more in-slots than a chunk holds gets one of its own */
        v = w;
    }
    chunks.push_back(make_uint2(V, in_off[V]));
}

/* Device node order.  Nodes are renumbered on the device in Cuthill-McKee
 * order of the undirected graph: breadth-first from a lowest-degree
 * unvisited node, neighbours in increasing degree (reversing the sequence,
 * the usual variant, measured about 1% slower).  Nodes whose arcs are close
 * then sit close, so the
 * waves that run at the same time gather from a narrow range of combine
 * buffers that the XCD's L2 keeps (on the random 1M-arc DAG the iteration
 * takes 18% less time).  Each node keeps its in-edges in the reference's
 * order, so every node's arithmetic, and every result, is unchanged; the
 * host sees its own numbering through upload and download. */
static int g_node_order = 1;

extern "C" int bp_set_node_order(int mode)
{
    int prev = g_node_order;
    g_node_order = (mode == 0 || mode == 2) ? mode : 1;
    return prev;
}

int bp_node_order_mode() { return g_node_order; }

void bp_cm_order(Graph_t g, std::vector<unsigned> &order)
{
    unsigned V = g->current_num_vertices, E = g->current_num_edges;
    std::vector<unsigned> off(V + 1, 0), adj(2 * (size_t)E);
    for (unsigned e = 0; e < E; ++e) {
        off[g->edges_src_index[e] + 1]++;
        off[g->edges_dest_index[e] + 1]++;
    }
    for (unsigned v = 0; v < V; ++v)
        off[v + 1] += off[v];
    {
        std::vector<unsigned> fill(off.begin(), off.end() - 1);
        for (unsigned e = 0; e < E; ++e) {
            unsigned a = g->edges_src_index[e], b = g->edges_dest_index[e];
            adj[fill[a]++] = b;
            adj[fill[b]++] = a;
        }
    }
    auto deg = [&](unsigned v) { return off[v + 1] - off[v]; };
    /* candidates for each component's start: ascending degree, then index */
    unsigned maxd = 0;
    for (unsigned v = 0; v < V; ++v)
        maxd = std::max(maxd, deg(v));
    std::vector<unsigned> cnt(maxd + 2, 0), cand(V);
    for (unsigned v = 0; v < V; ++v)
        cnt[deg(v) + 1]++;
    for (unsigned k = 0; k <= maxd; ++k)
        cnt[k + 1] += cnt[k];
    for (unsigned v = 0; v < V; ++v)
        cand[cnt[deg(v)]++] = v;
    std::vector<char> seen(V, 0);
    order.clear();
    order.reserve(V);
    std::vector<unsigned> nb;
    for (unsigned c : cand) {
        if (seen[c])
            continue;
#ifdef BP_ORDER_PSEUDO_PERIPHERAL
        /* George-Liu: restart from a lowest-degree node of the last BFS
         * level while the eccentricity grows */
        {
            static thread_local std::vector<unsigned> lvl, q;
            if (lvl.size() < V)
                lvl.assign(V, 0xffffffffu);
            unsigned ecc = 0;
            for (int round = 0; round < 8; ++round) {
                q.clear();
                q.push_back(c);
                lvl[c] = 0;
                size_t h = 0;
                unsigned last = 0;
                while (h < q.size()) {
                    unsigned v = q[h++];
                    last = lvl[v];
                    for (unsigned k = off[v]; k < off[v + 1]; ++k)
                        if (lvl[adj[k]] == 0xffffffffu) {
                            lvl[adj[k]] = lvl[v] + 1;
                            q.push_back(adj[k]);
                        }
                }
                unsigned best = c;
                for (unsigned v : q)
                    if (lvl[v] == last && (best == c || deg(v) < deg(best)))
                        best = v;
                for (unsigned v : q)
                    lvl[v] = 0xffffffffu;
                if (last <= ecc)
                    break;
                ecc = last;
                c = best;
            }
        }
#endif
        size_t head = order.size();
        seen[c] = 1;
        order.push_back(c);
        while (head < order.size()) {
            unsigned v = order[head++];
            nb.clear();
            for (unsigned k = off[v]; k < off[v + 1]; ++k)
                if (!seen[adj[k]]) {
                    seen[adj[k]] = 1;
                    nb.push_back(adj[k]);
                }
            std::sort(nb.begin(), nb.end(), [&](unsigned a, unsigned b) {
                return deg(a) != deg(b) ? deg(a) < deg(b) : a < b;
            });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
#ifdef BP_ORDER_REVERSE
    std::reverse(order.begin(), order.end());
#endif
}

/* Device node order, second step: nodes grouped by the XCDs that gather
 * them.  The iteration kernel hands each of the 8 XCDs one contiguous range
 * of chunks (xcd_block), so XCD x reads the combine buffers of its slots'
 * sources into its own L2 in 128-byte lines.  In Cuthill-McKee order the
 * sources an XCD needs from outside its range sit scattered between nodes
 * it does not need, and the line floor (distinct lines each XCD touches)
 * is about twice the row floor (distinct rows) on random DAGs.  Within
 * each XCD's range, nodes are stably sorted by the set of XCDs that gather
 * them (the ones only their own XCD gathers first), so rows gathered by the
 * same XCDs share lines.  Every node keeps its in-edges in order, so
 * results do not change.  The grouping is kept only when it lowers the
 * line floor (on a lattice, whose gathers are local, it does not). */
static uint64_t line_floor(Graph_t g, const std::vector<unsigned> &inv, const std::vector<unsigned> &perm,
                           const unsigned bounds[9], unsigned SP)
{
    unsigned V = g->current_num_vertices;
    size_t nlines = ((size_t)V * SP * 4 + 127) / 128 + 1;
    uint64_t count[8] = {};
    /* one thread per XCD: the distinct lines its nodes' sources occupy */
    std::vector<std::thread> th;
    for (unsigned x = 0; x < 8; ++x)
        th.emplace_back([&, x]() {
            std::vector<unsigned char> seen(nlines, 0);
            uint64_t n = 0;
            for (unsigned i = bounds[x]; i < bounds[x + 1]; ++i) {
                unsigned h = inv[i];
                for (unsigned k = g->dest_nodes_to_edges_node_list[h];
                     k < bp_csr_end(g, g->dest_nodes_to_edges_node_list, h); ++k) {
                    unsigned e = g->dest_nodes_to_edges_edge_list[k];
                    size_t l = (size_t)perm[g->edges_src_index[e]] * SP * 4 / 128;
                    if (!seen[l]) {
                        seen[l] = 1;
                        ++n;
                    }
                }
            }
            count[x] = n;
        });
    for (auto &t : th)
        t.join();
    uint64_t lines = 0;
    for (unsigned x = 0; x < 8; ++x)
        lines += count[x];
    return lines;
}

/* First device position of each XCD's range under the iteration kernel's
 * chunk split (bounds[8] = V) */
static void xcd_ranges(Graph_t g, const std::vector<unsigned> &inv, unsigned cap, unsigned bounds[9])
{
    unsigned V = g->current_num_vertices;
    std::vector<unsigned> in_off(V + 1, 0);
    for (unsigned i = 0; i < V; ++i) {
        unsigned h = inv[i];
        in_off[i + 1] = in_off[i] + (bp_csr_end(g, g->dest_nodes_to_edges_node_list, h) -
                                     g->dest_nodes_to_edges_node_list[h]);
    }
    std::vector<uint2> chunks;
    bp_build_chunks(in_off, V, cap, chunks);
    unsigned nch = (unsigned)chunks.size() - 1;
    unsigned want = (nch + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    unsigned long long W = (unsigned long long)std::max(8u, (want + 7u) & ~7u) * WAVES_PER_BLOCK;
    for (unsigned x = 0; x <= 8; ++x)
        bounds[x] = chunks[(unsigned)((x * W / 8) * nch / W)].x;
}

static void inverse_of(const std::vector<unsigned> &inv, std::vector<unsigned> &perm)
{
    perm.resize(inv.size());
    par_for(inv.size(), [&](size_t a, size_t b) {
        for (size_t i = a; i < b; ++i)
            perm[inv[i]] = (unsigned)i;
    });
}

static void xcd_group(Graph_t g, std::vector<unsigned> &inv, unsigned cap, unsigned SP)
{
    unsigned V = g->current_num_vertices;
    if (V < 2)
        return;
    unsigned bounds[9];
    xcd_ranges(g, inv, cap, bounds);
    std::vector<unsigned> perm;
    inverse_of(inv, perm);
    /* the XCDs that gather each node (those of its out-edges' destinations):
     * one thread per XCD marks the sources of its nodes' in-edges */
    std::vector<unsigned char> marks((size_t)8 * V, 0), mask(V, 0);
    {
        std::vector<std::thread> th;
        for (unsigned x = 0; x < 8; ++x)
            th.emplace_back([&, x]() {
                unsigned char *mk = marks.data() + (size_t)x * V;
                for (unsigned i = bounds[x]; i < bounds[x + 1]; ++i) {
                    unsigned h = inv[i];
                    for (unsigned k = g->dest_nodes_to_edges_node_list[h];
                         k < bp_csr_end(g, g->dest_nodes_to_edges_node_list, h); ++k)
                        mk[g->edges_src_index[g->dest_nodes_to_edges_edge_list[k]]] = 1;
                }
            });
        for (auto &t : th)
            t.join();
    }
    par_for(V, [&](size_t a, size_t b) {
        for (size_t h = a; h < b; ++h) {
            unsigned m = 0;
            for (unsigned x = 0; x < 8; ++x)
                m |= (unsigned)marks[(size_t)x * V + h] << x;
            mask[h] = (unsigned char)m;
        }
    });
    /* within each XCD's range a stable counting sort by mask, the nodes
     * only that XCD gathers first */
    std::vector<unsigned> grouped(V);
    {
        std::vector<std::thread> th;
        for (unsigned x = 0; x < 8; ++x)
            th.emplace_back([&, x]() {
                auto key = [&](unsigned v) { return mask[v] == (1u << x) ? 0u : 1u + mask[v]; };
                unsigned start[258] = {};
                for (unsigned i = bounds[x]; i < bounds[x + 1]; ++i)
                    ++start[key(inv[i]) + 1];
                start[0] = bounds[x];
                for (unsigned k = 1; k < 258; ++k)
                    start[k] += start[k - 1];
                for (unsigned i = bounds[x]; i < bounds[x + 1]; ++i)
                    grouped[start[key(inv[i])]++] = inv[i];
            });
        for (auto &t : th)
            t.join();
    }
    unsigned bounds2[9];
    xcd_ranges(g, grouped, cap, bounds2);
    std::vector<unsigned> perm2;
    inverse_of(grouped, perm2);
    if (line_floor(g, grouped, perm2, bounds2, SP) * 100 < line_floor(g, inv, perm, bounds, SP) * 99)
        inv.swap(grouped);
}

static int check_graph(Graph_t g)
{
    if (!g) {
        bp_set_error("graph is NULL");
        return -1;
    }
    if (g->current_num_vertices != g->total_num_vertices) {
        bp_set_error("graph has %u of %u nodes", g->current_num_vertices, g->total_num_vertices);
        return -1;
    }
    unsigned V = g->current_num_vertices, E = g->current_num_edges;
    /* the destination CSR must describe the edges (set_up_dest_nodes_to_edges) */
    for (unsigned v = 0; v < V; ++v) {
        unsigned b = g->dest_nodes_to_edges_node_list[v], e = bp_csr_end(g, g->dest_nodes_to_edges_node_list, v);
        if (b > e || e > E) {
            bp_set_error("destination CSR is not set up (call set_up_dest_nodes_to_edges)");
            return -1;
        }
        for (unsigned k = b; k < e; ++k) {
            unsigned edge = g->dest_nodes_to_edges_edge_list[k];
            if (edge >= E || g->edges_dest_index[edge] != v) {
                bp_set_error("destination CSR is not set up (call set_up_dest_nodes_to_edges)");
                return -1;
            }
        }
    }
    if (V > 0 && bp_csr_end(g, g->dest_nodes_to_edges_node_list, V - 1) != E) {
        bp_set_error("destination CSR does not cover all %u edges", E);
        return -1;
    }
    for (unsigned e = 0; e < E; ++e) {
        unsigned s = g->edges_src_index[e], t = g->edges_dest_index[e];
        if (s >= V || t >= V || g->edges_x_dim[e] != g->node_num_vars[s] || g->edges_y_dim[e] != g->node_num_vars[t]) {
            bp_set_error("edge %u is inconsistent with its nodes", e);
            return -1;
        }
    }
    for (unsigned v = 0; v < V; ++v)
        if (g->node_num_vars[v] > MAX_STATES) {
            bp_set_error("node %u has %u states (limit %d)", v, g->node_num_vars[v], MAX_STATES);
            return -1;
        }
    return 0;
}

/* the handle's pinned staging buffer, at least n floats */
static float *stage_buf(bp_dev *d, size_t n)
{
    if (n > d->h_stage_n) {
        if (d->h_stage)
            (void)hipHostFree(d->h_stage);
        d->h_stage = nullptr;
        d->h_stage_n = 0;
        if (hipHostMalloc((void **)&d->h_stage, std::max<size_t>(n, 1) * sizeof(float), hipHostMallocDefault) !=
            hipSuccess) {
            d->h_stage = nullptr;
            bp_set_error("pinned staging buffer of %zu floats: allocation failed", n);
            return nullptr;
        }
        d->h_stage_n = n;
    }
    return d->h_stage;
}

/* Host rows (MAX_STATES stride, first n valid) -> device rows (SP stride,
 * zero padded): host row h goes to device row to[h], walked in host order so
 * the large host array streams. */
static void pack_rows(const float *src, const unsigned *valid, size_t count, unsigned SP, float *out,
                      const unsigned *to)
{
    par_for(count, [=](size_t a, size_t b) {
        for (size_t h = a; h < b; ++h) {
            float *row = out + (size_t)to[h] * SP;
            unsigned n = valid[h];
            for (unsigned k = 0; k < SP; ++k)
                row[k] = (k < n) ? src[h * MAX_STATES + k] : 0.0f;
        }
    });
}

/* device rows -> host rows (first n of each), in host order */
static void unpack_rows(const float *in, const unsigned *valid, size_t count, unsigned SP, float *dst,
                        const unsigned *from)
{
    par_for(count, [=](size_t a, size_t b) {
        for (size_t h = a; h < b; ++h) {
            const float *row = in + (size_t)from[h] * SP;
            unsigned n = valid[h];
            for (unsigned k = 0; k < n; ++k)
                dst[h * MAX_STATES + k] = row[k];
        }
    });
}

/* The shared table pool.  tables holds the per-slot tables [s][S*S]; the
 * pool holds each table once, in the order of the first slot that reads
 * it, and tab[s] is slot s's pool index, with bit 31 set where s is the
 * later slot of a pair (partner[s] < s) whose table is exactly the
 * transpose of its partner's, which it then reads transposed.  The earlier
 * slot of a pair never reuses, so every slot's choice is independent. */
static void build_table_pool(const std::vector<unsigned> &partner, const float *tables, unsigned E, unsigned S,
                             std::vector<unsigned> &tab, std::vector<float> &pool)
{
    const size_t T = (size_t)S * S;
    std::vector<unsigned char> reuse(E, 0);
    par_for(E, [&](size_t a, size_t b) {
        for (size_t s = a; s < b; ++s) {
            unsigned f = partner[s];
            if (f == ~0u || f >= s)
                continue;
            const float *A = tables + s * T, *Bt = tables + (size_t)f * T;
            bool t = true;
            for (unsigned i = 0; i < S && t; ++i)
                for (unsigned j = 0; j < S && t; ++j)
                    t = memcmp(&A[i * S + j], &Bt[j * S + i], 4) == 0;
            reuse[s] = t;
        }
    });
    tab.resize(E);
    unsigned n = 0;
    for (unsigned s = 0; s < E; ++s)
        if (!reuse[s])
            tab[s] = n++;
    for (unsigned s = 0; s < E; ++s)
        if (reuse[s])
            tab[s] = tab[partner[s]] | 0x80000000u;
    pool.resize((size_t)n * T);
    par_for(E, [&](size_t a, size_t b) {
        for (size_t s = a; s < b; ++s)
            if (!reuse[s])
                memcpy(pool.data() + (size_t)tab[s] * T, tables + s * T, T * sizeof(float));
    });
}

static int upload_state(bp_dev *d, Graph_t g, bool tables)
{
    size_t VS = (size_t)d->V * d->SP, ES = (size_t)d->E * d->SP;
    float *tmp = stage_buf(d, std::max(std::max(VS, ES), tables ? d->jfloats : (size_t)0));
    if (!tmp)
        return -1;
    /* node states, in device order */
    std::vector<unsigned> node_n(d->V);
    for (unsigned v = 0; v < d->V; ++v)
        node_n[v] = g->node_num_vars[v];
    pack_rows(g->node_states, node_n.data(), d->V, d->SP, tmp, d->perm.data());
    HIP_TRY(hipMemcpyAsync(d->states, tmp, VS * sizeof(float), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemcpyAsync(d->states0, tmp, VS * sizeof(float), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    /* messages in slot order: msg[0] = previous, msg[1] = current */
    pack_rows(*g->previous_edge_messages, d->edge_x.data(), d->E, d->SP, tmp, d->edge_slot.data());
    HIP_TRY(hipMemcpyAsync(d->msg[0], tmp, ES * sizeof(float), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    pack_rows(*g->current_edge_messages, d->edge_x.data(), d->E, d->SP, tmp, d->edge_slot.data());
    HIP_TRY(hipMemcpyAsync(d->msg[1], tmp, ES * sizeof(float), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    d->cur_is_a = (g->current_edge_messages == &g->edges_messages);
    if (tables) {
        unsigned S = d->S;
        /* mixed: x rows at toff[s], row stride mixed_row_stride(S), zeros
         * beyond y; uniform: [s][S*S], zeros beyond x rows and y columns */
        const unsigned RS = mixed_row_stride(S);
        if (d->mixed)
            memset(tmp + d->jfloats - (size_t)S * RS, 0, (size_t)S * RS * sizeof(float));
        par_for(d->E, [&](size_t a, size_t b) {
            for (size_t e = a; e < b; ++e) {
                unsigned s = d->edge_slot[e];
                unsigned x = g->edges_x_dim[e], y = g->edges_y_dim[e];
                const float *J = g->edges_joint_probabilities + e * MAX_STATES * MAX_STATES;
                if (d->mixed) {
                    float *T = tmp + d->toff[s];
                    for (unsigned i = 0; i < x; ++i)
                        for (unsigned j = 0; j < RS; ++j)
                            T[i * RS + j] = (j < y) ? J[i * MAX_STATES + j] : 0.0f;
                } else {
                    float *T = tmp + (size_t)s * S * S;
                    for (unsigned i = 0; i < S; ++i)
                        for (unsigned j = 0; j < S; ++j)
                            T[i * S + j] = (i < x && j < y) ? J[i * MAX_STATES + j] : 0.0f;
                }
            }
        });
        HIP_TRY(hipMemcpyAsync(d->J, tmp, d->jfloats * sizeof(float), hipMemcpyHostToDevice, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
        if (d->shared) {
            std::vector<unsigned> tab;
            std::vector<float> pool;
            build_table_pool(d->partner, tmp, d->E, S, tab, pool);
            HIP_TRY(hipMemcpy(d->Jp, pool.data(), pool.size() * sizeof(float), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(d->slot_pool, tab.data(), (size_t)d->E * 4, hipMemcpyHostToDevice));
        }
    }
    d->latest = 0;
    d->bcur = 0;
    d->buf_valid = false;
    d->swaps = 0;
    return 0;
}

static double algorithmic_bytes(Graph_t g)
{
    /* distinct tables (a transposed reverse pair counts once), per-edge and
     * per-node traffic of one iteration */
    double bytes = 0.0;
    unsigned E = g->current_num_edges, V = g->current_num_vertices;
    /* a reverse edge emitted by the parsers directly follows its forward
     * edge: rev[e] says edge e + 1 is e's reverse with the transposed table
     * (tested on all host threads: the tables are 400 bytes apart) */
    std::vector<unsigned char> rev(E, 0), paired(E, 0);
    par_for(E ? E - 1 : 0, [&](size_t a, size_t b) {
        for (size_t e = a; e < b; ++e) {
            size_t f = e + 1;
            if (g->edges_src_index[f] != g->edges_dest_index[e] || g->edges_dest_index[f] != g->edges_src_index[e])
                continue;
            unsigned x = g->edges_x_dim[e], y = g->edges_y_dim[e];
            const float *A = g->edges_joint_probabilities + e * 100;
            const float *B = g->edges_joint_probabilities + f * 100;
            bool t = true;
            for (unsigned i = 0; i < x && t; ++i)
                for (unsigned j = 0; j < y && t; ++j)
                    t = (memcmp(&A[i * 10 + j], &B[j * 10 + i], 4) == 0);
            rev[e] = t;
        }
    });
    for (unsigned e = 0; e + 1 < E; ++e)
        if (!paired[e] && rev[e])
            paired[e + 1] = 1;
    for (unsigned e = 0; e < E; ++e) {
        double x = g->edges_x_dim[e], y = g->edges_y_dim[e];
        if (!paired[e])
            bytes += 4.0 * x * y;
        bytes += 4.0 * (2.0 * x + y + 1.0);
    }
    for (unsigned v = 0; v < V; ++v)
        bytes += 12.0 * g->node_num_vars[v] + 4.0;
    return bytes;
}

/* Transposed pairs (a reverse edge whose table is its forward edge's
 * transpose, xml-expression.c:402-406) can share one stored table: the
 * later slot of the pair reads the earlier one's, transposed.  That only
 * pays where the partner's table is still in the XCD's L2 when the later
 * slot reads it, i.e. where partners sit close in slot order in the same
 * XCD's chunk range (a lattice in Cuthill-McKee order; not a random DAG,
 * whose partners are scattered and would each pull a 128-byte line for
 * one table).  Sharing is switched on when at least 90% of the paired
 * slots have their partner in the same XCD range within 65,536 slots. */
static void plan_table_sharing(bp_dev *d, Graph_t g, const std::vector<uint2> &chunks)
{
    unsigned E = d->E;
    d->partner.assign(E, ~0u);
    /* rev[e]: the lowest-numbered edge dest(e) -> src(e), found among the
     * in-edges of src(e) (all host threads) */
    std::vector<unsigned> rev(E, ~0u);
    par_for(E, [&](size_t a, size_t b) {
        for (size_t e = a; e < b; ++e) {
            unsigned from = g->edges_src_index[e], to = g->edges_dest_index[e];
            if (from == to)
                continue;
            unsigned best = ~0u;
            for (unsigned k = g->dest_nodes_to_edges_node_list[from];
                 k < bp_csr_end(g, g->dest_nodes_to_edges_node_list, from); ++k) {
                unsigned f = g->dest_nodes_to_edges_edge_list[k];
                if (g->edges_src_index[f] == to && f < best)
                    best = f;
            }
            rev[e] = best;
        }
    });
    unsigned paired = 0, local = 0;
    unsigned nch = (unsigned)chunks.size() - 1;
    unsigned long long W = (unsigned long long)d->G * WAVES_PER_BLOCK;
    unsigned bound[9];
    for (unsigned x = 0; x <= 8; ++x)
        bound[x] = chunks[(unsigned)((x * W / 8) * nch / W)].y;
    auto xcd = [&](unsigned s) {
        unsigned x = 0;
        while (x < 7 && s >= bound[x + 1])
            ++x;
        return x;
    };
    for (unsigned e = 0; e < E; ++e) {
        if (rev[e] == ~0u)
            continue;
        unsigned s = d->edge_slot[e], f = d->edge_slot[rev[e]];
        if (d->partner[s] != ~0u || d->partner[f] != ~0u)
            continue;
        d->partner[s] = f;
        d->partner[f] = s;
        paired += 2;
        if (xcd(s) == xcd(f) && (s > f ? s - f : f - s) <= 65536u)
            local += 2;
    }
    d->shared = paired > 0 && (unsigned long long)local * 10 >= (unsigned long long)paired * 9;
    if (!d->shared)
        d->partner.clear();
}

static int upload_impl(Graph_t g, int device, bp_dev **out, bool plan)
{
    bp_clear_error();
    *out = nullptr;
    if (check_graph(g))
        return -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        bp_set_error("no HIP device available: the belief-propagation kernels need an MI355X (gfx950)");
        return -1;
    }
    bp_dev *d = new bp_dev();
    if (device < 0)
        HIP_TRY(hipGetDevice(&d->device));
    else
        d->device = device;
    HIP_TRY(hipSetDevice(d->device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, d->device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        bp_set_error("device %d is %s; libbp_amd.so is built for gfx950 only", d->device, prop.gcnArchName);
        delete d;
        return -1;
    }
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
        bp_set_error("hipStreamCreate failed");
        delete d;
        return -1;
    }

    unsigned V = g->current_num_vertices, E = g->current_num_edges;
    d->V = V;
    d->E = E;
    unsigned S_all = 1;
    for (unsigned v = 0; v < V; ++v)
        S_all = std::max(S_all, g->node_num_vars[v]);
    const unsigned cap = WAVE * ((S_all <= 4 && !(S_all >= 4 && BP_PLAIN_L > 1)) ? (unsigned)BP_SLOTS_SMALL : 1u);
    if (g_node_order && V > 1) {
        bp_cm_order(g, d->inv);
        if (g_node_order == 1)
            xcd_group(g, d->inv, cap, stride_of(S_all));
    } else {
        d->inv.resize(V);
        for (unsigned v = 0; v < V; ++v)
            d->inv[v] = v;
    }
    d->perm.assign(V, 0);
    for (unsigned i = 0; i < V; ++i)
        d->perm[d->inv[i]] = i;
    unsigned S = 1, smin = MAX_STATES;
    d->node_n.resize(V);
    for (unsigned i = 0; i < V; ++i)
        d->node_n[i] = g->node_num_vars[d->inv[i]];
    for (unsigned v = 0; v < V; ++v) {
        S = std::max(S, d->node_n[v]);
        smin = std::min(smin, d->node_n[v]);
    }
    if (V == 0)
        smin = S;
    d->S = S;
    d->SP = stride_of(S);
    d->mixed = (smin != S);
    d->edge_x.assign(g->edges_x_dim, g->edges_x_dim + E);

    /* slots: device node after device node, each node's in-edges in the
     * reference's destination-list order */
    std::vector<unsigned> in_off(V + 1);
    d->slot_edge.clear();
    d->slot_edge.reserve(E);
    for (unsigned i = 0; i < V; ++i) {
        unsigned h = d->inv[i];
        in_off[i] = (unsigned)d->slot_edge.size();
        for (unsigned k = g->dest_nodes_to_edges_node_list[h]; k < bp_csr_end(g, g->dest_nodes_to_edges_node_list, h);
             ++k)
            d->slot_edge.push_back(g->dest_nodes_to_edges_edge_list[k]);
    }
    in_off[V] = E;
    d->in_off_h = in_off;
    d->edge_slot.assign(E, 0);
    for (unsigned s = 0; s < E; ++s)
        d->edge_slot[d->slot_edge[s]] = s;
    std::vector<unsigned> slot_src(E);
    std::vector<unsigned char> slot_xy(E);
    for (unsigned s = 0; s < E; ++s) {
        unsigned e = d->slot_edge[s];
        slot_src[s] = d->perm[g->edges_src_index[e]];
        slot_xy[s] = (unsigned char)(g->edges_x_dim[e] | (g->edges_y_dim[e] << 4));
    }
    std::vector<uint2> chunks;
    bp_build_chunks(in_off, V, cap, chunks);
    d->nchunks = (unsigned)chunks.size() - 1;
    /* one chunk per wave; a multiple of 8 blocks so each XCD gets one range */
    unsigned want = (d->nchunks + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    unsigned G = std::max(8u, (want + 7u) & ~7u);
    if (BP_PERSIST > 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            G = std::min(G, std::max(8u, ((unsigned)cus * BP_PERSIST + 7u) & ~7u));
    }
    d->G = G;
    d->NP = G * WAVES_PER_BLOCK;
    if (plan && !d->mixed && S <= 4 && BP_SHARE_TABLES)
        plan_table_sharing(d, g, chunks);
    /* delta partials of up to part_cap iterations are reduced per batch */
    d->part_cap = (unsigned)std::min<size_t>(256, std::max<size_t>(1, (64u << 20) / (8ull * d->NP)));
    d->alg_bytes = algorithmic_bytes(g);

    size_t SP = d->SP;
    /* mixed state counts: each slot's table is its own x rows, each row
     * padded to mixed_row_stride(S) floats (16-byte aligned rows, so a row
     * is whole vector loads), behind a per-slot offset; the kernels read
     * every row up to S unconditionally, so the array carries S rows of
     * slack past the last table */
    if (d->mixed) {
        const size_t RS = mixed_row_stride(S);
        d->toff.resize(E);
        size_t tot = 0;
        for (unsigned s = 0; s < E; ++s) {
            unsigned e = d->slot_edge[s];
            d->toff[s] = (unsigned)tot;
            tot += (size_t)g->edges_x_dim[e] * RS;
        }
        tot += (size_t)S * RS;
        if (tot >= 0xffffffffull) {
            bp_set_error("upload: more than 2^32 table entries");
            bp_free(d);
            return -1;
        }
        d->jfloats = tot;
        if (dmalloc(&d->slot_toff, E) ||
            (E && hipMemcpy(d->slot_toff, d->toff.data(), (size_t)E * 4, hipMemcpyHostToDevice) != hipSuccess)) {
            bp_set_error("upload of the table offsets failed");
            bp_free(d);
            return -1;
        }
    } else {
        d->jfloats = (size_t)E * S * S;
    }
    if (d->shared && (dmalloc(&d->Jp, (size_t)E * S * S) || dmalloc(&d->slot_pool, E))) {
        bp_free(d);
        return -1;
    }
    if (dmalloc(&d->in_off, V + 1) || dmalloc(&d->slot_src, E) || dmalloc(&d->slot_xy, E) ||
        dmalloc(&d->J, d->jfloats) || dmalloc(&d->nvars, V) || dmalloc(&d->states, V * SP) ||
        dmalloc(&d->buf[0], V * SP) || dmalloc(&d->buf[1], V * SP) || dmalloc(&d->msg[0], (size_t)E * SP) ||
        dmalloc(&d->msg[1], (size_t)E * SP) || dmalloc(&d->states0, V * SP) ||
        dmalloc(&d->chunk, chunks.size()) || dmalloc(&d->partials, (size_t)d->part_cap * d->NP) ||
        dmalloc(&d->deltas, std::max(d->part_cap, 16u)) || dmalloc(&d->last, 1) || dmalloc(&d->stop, 1)) {
        bp_free(d);
        return -1;
    }
    std::vector<unsigned char> nv(V);
    for (unsigned v = 0; v < V; ++v)
        nv[v] = (unsigned char)d->node_n[v];
    bool ok = hipMemcpy(d->in_off, in_off.data(), (V + 1) * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d->slot_src, slot_src.data(), (size_t)E * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d->slot_xy, slot_xy.data(), E, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d->nvars, nv.data(), V, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d->chunk, chunks.data(), chunks.size() * sizeof(uint2), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemset(d->buf[0], 0, V * SP * 4) == hipSuccess && hipMemset(d->buf[1], 0, V * SP * 4) == hipSuccess &&
              hipMemset(d->stop, 0, 4) == hipSuccess;
    if (!ok) {
        bp_set_error("upload of the graph structure failed");
        bp_free(d);
        return -1;
    }
    if (upload_state(d, g, true)) {
        bp_free(d);
        return -1;
    }
    /* the tree schedule needs levels; without them bp_tree() reports it */
    if (plan && g->num_levels > 0 && bp_tree_plan_build(d, g) != 0)
        bp_clear_error();
    *out = d;
    return 0;
}

extern "C" int bp_upload(Graph_t g, int device, bp_dev **out) { return upload_impl(g, device, out, true); }

/* an upload without the tree schedule, for the adjoint and batched handles */
int bp_upload_graph(Graph_t g, int device, bp_dev **out) { return upload_impl(g, device, out, false); }

extern "C" int bp_upload_state(bp_dev *d, Graph_t g)
{
    bp_clear_error();
    if (!d || !g || g->current_num_vertices != d->V || g->current_num_edges != d->E) {
        bp_set_error("bp_upload_state: graph does not match the handle");
        return -1;
    }
    HIP_TRY(hipSetDevice(d->device));
    return upload_state(d, g, true);
}

/* ------------------------------------------------------------------------ */
/* launches                                                                  */

LoopyArgs bp_make_args(bp_dev *d)
{
    LoopyArgs a;
    a.in_off = d->in_off;
    a.slot_src = d->slot_src;
    a.slot_xy = d->slot_xy;
    a.J = d->J;
    a.nvars = d->nvars;
    a.states = d->states;
    a.states_out = d->states;
    a.buf_in = d->buf[d->bcur];
    a.buf_out = d->buf[d->bcur ^ 1];
    a.msg_in = d->msg[d->latest];
    a.msg_out = d->msg[d->latest ^ 1];
    a.chunk = d->chunk;
    a.chunk_base = nullptr;
    a.order = nullptr;
    a.nchunks = d->nchunks;
    a.partials = d->partials;
    a.stop = nullptr;
    a.slot_tab = d->mixed ? d->slot_toff : d->shared ? d->slot_pool : nullptr;
    if (d->shared)
        a.J = d->Jp;
    a.tabs = nullptr;
    a.chunk_partials = nullptr;
    a.queue = nullptr;
    a.qblock = 1;
    return a;
}

/* ev (nullable) = {start, stop}: the launch is bracketed by events taken at
 * the kernel's own begin and end (hipExtLaunchKernelGGL), so their elapsed
 * time is the kernel's duration alone, without the gap between launches. */
template <int S> struct PlainLanes {
    static constexpr int value = (S >= 4) ? BP_PLAIN_L : 1;
};

template <int S, bool MIXED, int MODE, bool POOLED>
static void launch_k(const LoopyArgs &a, unsigned G, hipStream_t st, const hipEvent_t *ev)
{
    dim3 grid(G), block(BLOCK);
    if constexpr (PlainLanes<S>::value > 1) {
        constexpr int L = PlainLanes<S>::value;
        if (ev)
            hipExtLaunchKernelGGL((k_loopy<S, MIXED, MODE, 1, false, L, POOLED>), grid, block, 0, st, ev[0], ev[1], 0,
                                  a);
        else
            hipLaunchKernelGGL((k_loopy<S, MIXED, MODE, 1, false, L, POOLED>), grid, block, 0, st, a);
    } else {
        constexpr int R = SlotsPerLane<S>::value;
        if (ev)
            hipExtLaunchKernelGGL((k_loopy<S, MIXED, MODE, R, false, 1, POOLED>), grid, block, 0, st, ev[0], ev[1], 0,
                                  a);
        else
            hipLaunchKernelGGL((k_loopy<S, MIXED, MODE, R, false, 1, POOLED>), grid, block, 0, st, a);
    }
}

template <int S, bool MIXED, bool POOLED = false>
static void launch_s(int mode, const LoopyArgs &a, unsigned G, hipStream_t st, const hipEvent_t *ev)
{
    if (mode == ITER)
        launch_k<S, MIXED, ITER, POOLED>(a, G, st, ev);
    else if (mode == INIT)
        launch_k<S, MIXED, INIT, POOLED>(a, G, st, ev);
    else if (mode == EDGE)
        launch_k<S, MIXED, EDGE, POOLED>(a, G, st, ev);
    else if (mode == FWD)
        launch_k<S, MIXED, FWD, POOLED>(a, G, st, ev);
    else if (mode == FINIT)
        launch_k<S, MIXED, FINIT, POOLED>(a, G, st, ev);
    else
        launch_k<S, MIXED, PREP, POOLED>(a, G, st, ev);
}

template <bool MIXED>
static void launch_m(unsigned S, int mode, const LoopyArgs &a, unsigned G, hipStream_t st, const hipEvent_t *ev)
{
    switch (S) {
    case 1: launch_s<1, MIXED>(mode, a, G, st, ev); break;
    case 2: launch_s<2, MIXED>(mode, a, G, st, ev); break;
    case 3: launch_s<3, MIXED>(mode, a, G, st, ev); break;
    case 4: launch_s<4, MIXED>(mode, a, G, st, ev); break;
    case 5: launch_s<5, MIXED>(mode, a, G, st, ev); break;
    case 6: launch_s<6, MIXED>(mode, a, G, st, ev); break;
    case 7: launch_s<7, MIXED>(mode, a, G, st, ev); break;
    case 8: launch_s<8, MIXED>(mode, a, G, st, ev); break;
    case 9: launch_s<9, MIXED>(mode, a, G, st, ev); break;
    default: launch_s<10, MIXED>(mode, a, G, st, ev); break;
    }
}

/* uniform state counts up to 4 whose transposed pairs share stored tables */
static void launch_pooled(unsigned S, int mode, const LoopyArgs &a, unsigned G, hipStream_t st, const hipEvent_t *ev)
{
    switch (S) {
    case 1: launch_s<1, false, true>(mode, a, G, st, ev); break;
    case 2: launch_s<2, false, true>(mode, a, G, st, ev); break;
    case 3: launch_s<3, false, true>(mode, a, G, st, ev); break;
    default: launch_s<4, false, true>(mode, a, G, st, ev); break;
    }
}

static void launch_on(bp_dev *d, int mode, const LoopyArgs &a, hipStream_t st, const hipEvent_t *ev = nullptr)
{
    if (d->V == 0)
        return;
    if (d->mixed)
        launch_m<true>(d->S, mode, a, d->G, st, ev);
    else if (d->shared && a.slot_tab)
        launch_pooled(d->S, mode, a, d->G, st, ev);
    else
        launch_m<false>(d->S, mode, a, d->G, st, ev);
}

static void launch(bp_dev *d, int mode, const LoopyArgs &a, const hipEvent_t *ev = nullptr)
{
    launch_on(d, mode, a, d->stream, ev);
}

void bp_launch_loopy(bp_dev *d, int mode, const LoopyArgs &a, hipStream_t st, const hipEvent_t *ev)
{
    launch_on(d, mode, a, st, ev);
}

/* states (.) prod(previous messages) into buf[bcur] */
static int ensure_buffer(bp_dev *d)
{
    if (d->buf_valid)
        return 0;
    LoopyArgs a = bp_make_args(d);
    a.buf_out = d->buf[d->bcur];
    launch(d, PREP, a);
    HIP_TRY(hipGetLastError());
    d->buf_valid = true;
    return 0;
}

extern "C" int bp_loopy_init(bp_dev *d)
{
    bp_clear_error();
    if (!d) {
        bp_set_error("bp_loopy_init: NULL handle");
        return -1;
    }
    HIP_TRY(hipSetDevice(d->device));
    LoopyArgs a = bp_make_args(d);
    a.msg_out = d->msg[d->latest]; /* the previous-message array, graph.c:734 */
    a.buf_out = d->buf[d->bcur];
    launch(d, INIT, a);
    HIP_TRY(hipGetLastError());
    d->buf_valid = true;
    return 0;
}

/* Records n iterations (+ the batch's delta reduction) on the stream. */
static void enqueue_iterations(bp_dev *d, unsigned n, int latest, int bcur)
{
    for (unsigned i = 0; i < n; ++i) {
        LoopyArgs a = bp_make_args(d);
        a.msg_in = d->msg[latest];
        a.msg_out = d->msg[latest ^ 1];
        a.buf_in = d->buf[bcur];
        a.buf_out = d->buf[bcur ^ 1];
        a.partials = d->partials + (size_t)i * d->NP;
        launch(d, ITER, a);
        latest ^= 1;
        bcur ^= 1;
    }
    hipLaunchKernelGGL(k_finish, dim3(n), dim3(BLOCK), 0, d->stream, d->partials, d->NP, n, d->deltas, d->last,
                       d->stop, 0.0f, 0, 0u);
}

extern "C" int bp_loopy_iterate(bp_dev *d, unsigned int n, double *deltas)
{
    bp_clear_error();
    if (!d) {
        bp_set_error("bp_loopy_iterate: NULL handle");
        return -1;
    }
    HIP_TRY(hipSetDevice(d->device));
    if (ensure_buffer(d))
        return -1;
    unsigned done = 0;
    while (done < n) {
        unsigned batch = std::min(n - done, d->part_cap);
        unsigned long long key = ((unsigned long long)batch << 2) | ((unsigned)d->latest << 1) | (unsigned)d->bcur;
        auto it = d->graphs.find(key);
        if (it == d->graphs.end()) {
            hipGraph_t graph;
            HIP_TRY(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
            enqueue_iterations(d, batch, d->latest, d->bcur);
            HIP_TRY(hipStreamEndCapture(d->stream, &graph));
            hipGraphExec_t exec;
            hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                bp_set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e));
                return -1;
            }
            it = d->graphs.emplace(key, exec).first;
        }
        HIP_TRY(hipGraphLaunch(it->second, d->stream));
        if (batch & 1u) {
            d->latest ^= 1;
            d->bcur ^= 1;
        }
        d->swaps += batch;
        if (deltas) {
            HIP_TRY(hipMemcpyAsync(deltas + done, d->deltas, batch * sizeof(double), hipMemcpyDeviceToHost, d->stream));
            HIP_TRY(hipStreamSynchronize(d->stream));
        }
        done += batch;
    }
    return 0;
}

/* ---- the allocation-free iteration entry point ---------------------------
 * bp_loopy_iterate_noalloc: the kernels of bp_loopy_iterate launched one by
 * one (no HIP graph is created), with the per-wave FP64 delta partials of
 * the running iteration and the per-iteration deltas in `scratch`, a fixed
 * array of the library, instead of the handle's buffers.  Each GPU holds its
 * own copy of the array; its device address is looked up once per device
 * (hipGetSymbolAddress, with that device current) and only that address is
 * passed to kernels.  Calls that overlap on one GPU share the array and
 * race.  Nothing in it needs clearing: every wave writes its partial each
 * iteration and k_finish writes each delta before it is read. */
static constexpr size_t NOALLOC_SCRATCH_DOUBLES = (4u << 20) / sizeof(double); /* 4 MiB */
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

static constexpr int NOALLOC_MAX_DEVICES = 64;
static double *noalloc_ptr[NOALLOC_MAX_DEVICES];
static std::mutex noalloc_mu;

/* The device address of `scratch` on `device` (which must be current). */
static int noalloc_scratch(int device, double **out)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) {
        bp_set_error("bp_loopy_iterate_noalloc: device %d is beyond the %d supported", device, NOALLOC_MAX_DEVICES);
        return -1;
    }
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            bp_set_error("bp_loopy_iterate_noalloc: hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                         hipGetErrorString(e));
            return -1;
        }
        noalloc_ptr[device] = (double *)p;
    }
    *out = noalloc_ptr[device];
    return 0;
}

extern "C" int bp_loopy_iterate_noalloc(bp_dev *d, unsigned int n, double *deltas)
{
    bp_clear_error();
    if (!d) {
        bp_set_error("bp_loopy_iterate_noalloc: NULL handle");
        return -1;
    }
    if ((size_t)d->NP + n > NOALLOC_SCRATCH_DOUBLES) {
        bp_set_error("bp_loopy_iterate_noalloc: %u delta partials per iteration + %u deltas exceed the fixed "
                     "scratch of %zu doubles (4 MiB); use bp_loopy_iterate", d->NP, n, NOALLOC_SCRATCH_DOUBLES);
        return -1;
    }
    HIP_TRY(hipSetDevice(d->device));
    double *sc = nullptr;
    if (noalloc_scratch(d->device, &sc))
        return -1;
    double *partials = sc, *ds = sc + d->NP;
    if (ensure_buffer(d))
        return -1;
    for (unsigned i = 0; i < n; ++i) {
        LoopyArgs a = bp_make_args(d);
        a.partials = partials;
        launch(d, ITER, a);
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(BLOCK), 0, d->stream, partials, d->NP, 1u, ds + i, nullptr,
                           nullptr, 0.0f, 0, 0u);
        d->latest ^= 1;
        d->bcur ^= 1;
        d->swaps += 1;
    }
    HIP_TRY(hipGetLastError());
    if (deltas && n) {
        HIP_TRY(hipMemcpyAsync(deltas, ds, n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
    }
    return 0;
}

/* Buffer roles while a run is being enqueued.  ITER reads msg[latest] and
 * writes msg[latest^1], then the host's pointers swap (graph.c:1179-1182):
 * latest flips and a swap is owed.  EDGE copies current into previous and
 * writes current without a swap (graph.c:1186-1227): it reads msg[latest^1],
 * writes msg[latest] and relabels; its beliefs ping-pong between states and
 * buf[0] because every node reads its sources' beliefs of the same step. */
struct Roles {
    int latest, bcur;
    bool scratch; /* EDGE: the current beliefs are in buf[0] */
};

static LoopyArgs step_args(bp_dev *d, int mode, Roles &r)
{
    LoopyArgs a = bp_make_args(d);
    if (mode == ITER) {
        a.msg_in = d->msg[r.latest];
        a.msg_out = d->msg[r.latest ^ 1];
        a.buf_in = d->buf[r.bcur];
        a.buf_out = d->buf[r.bcur ^ 1];
        r.bcur ^= 1;
    } else {
        a.msg_in = d->msg[r.latest ^ 1];
        a.msg_out = d->msg[r.latest];
        a.states = r.scratch ? d->buf[0] : d->states;
        a.states_out = r.scratch ? d->states : d->buf[0];
        r.scratch = !r.scratch;
    }
    r.latest ^= 1;
    return a;
}

static int commit_roles(bp_dev *d, int mode, const Roles &r, unsigned ran)
{
    d->latest = r.latest;
    d->bcur = r.bcur;
    if (mode == ITER) {
        d->swaps += ran;
        return 0;
    }
    d->buf_valid = false;
    if (r.scratch) {
        size_t n = (size_t)d->V * d->SP;
        if (n)
            hipLaunchKernelGGL(k_copy_floats, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d->stream, d->buf[0],
                               d->states, n);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

/* Runs until the stop test of graph.c:1313 fires on an FP64 delta after an
 * iteration, or max_iterations; the test runs on the device, later
 * launches of the batch return at once.  *result: the index of the
 * converged iteration, or max_iterations (graph.c:1278-1324). */
static int run_until(bp_dev *d, int mode, float convergence, unsigned max_iterations, unsigned *result,
                     double *tail /* [2]: the last two deltas, nullable */)
{
    if (mode == ITER && ensure_buffer(d))
        return -1;
    hipLaunchKernelGGL(k_until_reset, dim3(1), dim3(64), 0, d->stream, d->last, d->stop);
    HIP_TRY(hipGetLastError());
    const unsigned batch = 16;
    unsigned done = 0, stop = 0;
    double hist[2] = {-1.0, 0.0};
    std::vector<double> ds(batch);
    Roles r = {d->latest, d->bcur, false};
    while (done < max_iterations) {
        unsigned n = std::min(batch, max_iterations - done);
        Roles before = r;
        for (unsigned i = 0; i < n; ++i) {
            LoopyArgs a = step_args(d, mode, r);
            a.partials = d->partials;
            a.stop = d->stop;
            launch(d, mode, a);
            hipLaunchKernelGGL(k_finish, dim3(1), dim3(BLOCK), 0, d->stream, d->partials, d->NP, 1u, d->deltas,
                               d->last, d->stop, convergence, 1, i);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&stop, d->stop, sizeof(unsigned), hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipMemcpyAsync(ds.data(), d->deltas, n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
        unsigned ran = stop ? stop : n; /* launches after the stop returned at once */
        r = before;
        for (unsigned i = 0; i < ran; ++i)
            (void)step_args(d, mode, r);
        for (unsigned i = 0; i < ran; ++i) {
            hist[0] = hist[1];
            hist[1] = ds[i];
        }
        if (commit_roles(d, mode, r, ran))
            return -1;
        done += ran;
        if (stop)
            break;
    }
    hipLaunchKernelGGL(k_until_reset, dim3(1), dim3(64), 0, d->stream, d->last, d->stop);
    HIP_TRY(hipStreamSynchronize(d->stream));
    *result = stop ? done - 1 : max_iterations;
    if (tail) {
        tail[0] = hist[0];
        tail[1] = hist[1];
    }
    return 0;
}

/* The OpenACC drivers' schedule (graph.c:1326-1420,1438-1520): the stop test
 * after every BATCH_SIZE iterations on the last delta; *result counts the
 * iterations of the batches before the converged one. */
static int run_batched(bp_dev *d, int mode, float convergence, unsigned max_iterations, unsigned *result)
{
    if (mode == ITER && ensure_buffer(d))
        return -1;
    double previous = -1.0, delta = 0.0;
    unsigned num_iter = 0, i;
    Roles r = {d->latest, d->bcur, false};
    unsigned ran = 0;
    for (i = 0; i < max_iterations; i += BATCH_SIZE) {
        for (unsigned j = 0; j < BATCH_SIZE; ++j) {
            LoopyArgs a = step_args(d, mode, r);
            a.partials = d->partials;
            launch(d, mode, a);
            hipLaunchKernelGGL(k_finish, dim3(1), dim3(BLOCK), 0, d->stream, d->partials, d->NP, 1u, d->deltas + j,
                               d->last, d->stop, 0.0f, 0, 0u);
        }
        ran += BATCH_SIZE;
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&delta, d->deltas + BATCH_SIZE - 1, sizeof(double), hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
        double change = delta - previous;
        if (change < 0.0)
            change = -change;
        if (delta < (double)convergence || change < (double)convergence)
            break;
        previous = delta;
        num_iter += BATCH_SIZE;
    }
    if (commit_roles(d, mode, r, ran))
        return -1;
    HIP_TRY(hipStreamSynchronize(d->stream));
    if (i >= max_iterations)
        printf("No Convergence: previous: %f vs current: %f\n", previous, delta);
    *result = num_iter;
    return 0;
}

/* n iterations of `mode` with no stop test (deltas still reduced). */
static int run_fixed(bp_dev *d, int mode, unsigned n)
{
    if (mode == ITER && ensure_buffer(d))
        return -1;
    Roles r = {d->latest, d->bcur, false};
    for (unsigned i = 0; i < n; ++i) {
        LoopyArgs a = step_args(d, mode, r);
        a.partials = d->partials;
        launch(d, mode, a);
    }
    HIP_TRY(hipGetLastError());
    return commit_roles(d, mode, r, n);
}

extern "C" int bp_loopy_until(bp_dev *d, float convergence, unsigned int max_iterations, unsigned int *result)
{
    bp_clear_error();
    if (!d || !result) {
        bp_set_error("bp_loopy_until: NULL argument");
        return -1;
    }
    HIP_TRY(hipSetDevice(d->device));
    return run_until(d, ITER, convergence, max_iterations, result, nullptr);
}

extern "C" int bp_edge_iterate(bp_dev *d, unsigned int n)
{
    bp_clear_error();
    if (!d) {
        bp_set_error("bp_edge_iterate: NULL handle");
        return -1;
    }
    HIP_TRY(hipSetDevice(d->device));
    return run_fixed(d, EDGE, n);
}

extern "C" int bp_edge_until(bp_dev *d, float convergence, unsigned int max_iterations, unsigned int *result)
{
    bp_clear_error();
    if (!d || !result) {
        bp_set_error("bp_edge_until: NULL argument");
        return -1;
    }
    HIP_TRY(hipSetDevice(d->device));
    return run_until(d, EDGE, convergence, max_iterations, result, nullptr);
}

extern "C" int bp_restore_states(bp_dev *d)
{
    bp_clear_error();
    if (!d) {
        bp_set_error("bp_restore_states: NULL handle");
        return -1;
    }
    HIP_TRY(hipSetDevice(d->device));
    size_t n = (size_t)d->V * d->SP;
    if (n) {
        hipLaunchKernelGGL(k_copy_floats, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d->stream, d->states0,
                           d->states, n);
        HIP_TRY(hipGetLastError());
    }
    d->buf_valid = false;
    return 0;
}

extern "C" int bp_sync(bp_dev *d)
{
    bp_clear_error();
    if (!d) {
        bp_set_error("bp_sync: NULL handle");
        return -1;
    }
    HIP_TRY(hipSetDevice(d->device));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return 0;
}

extern "C" int bp_profile_loopy(bp_dev *d, unsigned int n, float *ms)
{
    bp_clear_error();
    if (!d || !ms) {
        bp_set_error("bp_profile_loopy: NULL argument");
        return -1;
    }
    HIP_TRY(hipSetDevice(d->device));
    if (ensure_buffer(d))
        return -1;
    std::vector<hipEvent_t> ev(2 * (size_t)n);
    for (auto &e : ev)
        HIP_TRY(hipEventCreate(&e));
    unsigned done = 0;
    while (done < n) {
        unsigned batch = std::min(n - done, d->part_cap);
        for (unsigned i = 0; i < batch; ++i) {
            LoopyArgs a = bp_make_args(d);
            a.partials = d->partials + (size_t)i * d->NP;
            launch(d, ITER, a, &ev[2 * (size_t)(done + i)]);
            d->latest ^= 1;
            d->bcur ^= 1;
            d->swaps++;
        }
        hipLaunchKernelGGL(k_finish, dim3(batch), dim3(BLOCK), 0, d->stream, d->partials, d->NP, batch, d->deltas,
                           d->last, d->stop, 0.0f, 0, 0u);
        HIP_TRY(hipStreamSynchronize(d->stream));
        for (unsigned i = 0; i < batch; ++i)
            HIP_TRY(hipEventElapsedTime(&ms[done + i], ev[2 * (size_t)(done + i)], ev[2 * (size_t)(done + i) + 1]));
        done += batch;
    }
    for (auto &e : ev)
        (void)hipEventDestroy(e);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* download                                                                  */

extern "C" int bp_download_states(bp_dev *d, float *out)
{
    bp_clear_error();
    HIP_TRY(hipSetDevice(d->device));
    std::vector<float> st((size_t)d->V * d->SP);
    HIP_TRY(hipMemcpyAsync(st.data(), d->states, st.size() * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    for (unsigned i = 0; i < d->V; ++i)
        std::memcpy(out + (size_t)d->inv[i] * d->SP, st.data() + (size_t)i * d->SP, d->SP * sizeof(float));
    return 0;
}

static int download_msgs(bp_dev *d, const float *dev, float *host)
{
    size_t ES = (size_t)d->E * d->SP;
    float *tmp = stage_buf(d, ES);
    if (!tmp)
        return -1;
    HIP_TRY(hipMemcpyAsync(tmp, dev, ES * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    unpack_rows(tmp, d->edge_x.data(), d->E, d->SP, host, d->edge_slot.data());
    return 0;
}

extern "C" int bp_download(bp_dev *d, Graph_t g)
{
    bp_clear_error();
    if (!d || !g || g->current_num_vertices != d->V || g->current_num_edges != d->E) {
        bp_set_error("bp_download: graph does not match the handle");
        return -1;
    }
    HIP_TRY(hipSetDevice(d->device));
    size_t VS = (size_t)d->V * d->SP;
    float *st = stage_buf(d, VS);
    if (!st)
        return -1;
    HIP_TRY(hipMemcpyAsync(st, d->states, VS * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    {
        std::vector<unsigned> node_n(d->V);
        for (unsigned v = 0; v < d->V; ++v)
            node_n[v] = g->node_num_vars[v];
        unpack_rows(st, node_n.data(), d->V, d->SP, g->node_states, d->perm.data());
    }
    /* one pointer swap per iteration run on the device (graph.c:1179-1182) */
    if (d->swaps & 1u) {
        float **t = g->previous_edge_messages;
        g->previous_edge_messages = g->current_edge_messages;
        g->current_edge_messages = t;
    }
    d->swaps = 0;
    d->cur_is_a = (g->current_edge_messages == &g->edges_messages);
    /* the device buffers keep their roles (cached HIP graphs hold their
     * addresses); msg[latest] is the newest, i.e. the host's previous array */
    if (download_msgs(d, d->msg[d->latest], *g->previous_edge_messages) ||
        download_msgs(d, d->msg[d->latest ^ 1], *g->current_edge_messages))
        return -1;
    return 0;
}

extern "C" int bp_dev_info(const bp_dev *d, uint64_t *info)
{
    if (!d || !info)
        return -1;
    info[0] = d->V;
    info[1] = d->E;
    info[2] = d->SP;
    info[3] = d->S;
    info[4] = d->nchunks;
    info[5] = d->mixed ? 1 : 0;
    info[6] = bp_tree_plan_waves(d->tree);
    info[7] = d->shared ? 1 : 0;
    return 0;
}

extern "C" double bp_loopy_algorithmic_bytes(const bp_dev *d) { return d ? d->alg_bytes : 0.0; }

/* ------------------------------------------------------------------------ */
/* drop-in entry points (include/graph.h)                                    */

/* The device mirror of a Graph_t, rebuilt when its structure changed;
 * states, tables and messages are re-read on every call because callers
 * may write the public arrays between calls. */
static bp_dev *dropin_device_status(Graph_t g)
{
    struct bp_graph_impl *impl = bp_impl(g);
    bp_dev *d = (bp_dev *)impl->dev_cache;
    if (d && impl->dev_cache_generation == impl->generation)
        return bp_upload_state(d, g) == 0 ? d : nullptr;
    bp_dev_cache_release(impl);
    if (upload_impl(g, -1, &d, false) != 0)
        return nullptr;
    impl->dev_cache = d;
    impl->dev_cache_generation = impl->generation;
    return d;
}

extern "C" bp_dev *bp_dropin_device(Graph_t g, const char *who)
{
    bp_dev *d = dropin_device_status(g);
    if (!d)
        bp_fail("%s: %s", who, bp_last_error());
    return d;
}

extern "C" void bp_dev_cache_release(struct bp_graph_impl *impl)
{
    if (impl && impl->dev_cache) {
        bp_free((bp_dev *)impl->dev_cache);
        impl->dev_cache = nullptr;
    }
}

extern "C" void init_previous_edge(Graph_t g)
{
    bp_dev *d = bp_dropin_device(g, "init_previous_edge");
    if (!d)
        return;
    if (bp_loopy_init(d) || bp_download(d, g))
        bp_fail("init_previous_edge: %s", bp_last_error());
}

extern "C" void loopy_propagate_one_iteration(Graph_t g)
{
    bp_dev *d = bp_dropin_device(g, "loopy_propagate_one_iteration");
    if (!d)
        return;
    if (bp_loopy_iterate(d, 1, nullptr) || bp_download(d, g))
        bp_fail("loopy_propagate_one_iteration: %s", bp_last_error());
}

/* the drop-in convergence drivers: run on the device mirror, download */
static unsigned int dropin_run(Graph_t g, const char *who, int mode, int batched, float convergence,
                               unsigned int max_iterations)
{
    bp_dev *d = bp_dropin_device(g, who);
    if (!d)
        return 0;
    unsigned int result = 0;
    double tail[2] = {-1.0, 0.0};
    int rc = hipSetDevice(d->device) == hipSuccess ? 0 : -1;
    if (rc == 0)
        rc = batched ? run_batched(d, mode, convergence, max_iterations, &result)
                     : run_until(d, mode, convergence, max_iterations, &result, tail);
    if (rc || bp_download(d, g)) {
        bp_fail("%s: %s", who, bp_last_error());
        return 0;
    }
    if (!batched && result == max_iterations)
        printf("No Convergence: previous: %f vs current: %f\n", tail[0], tail[1]);
    return result;
}

extern "C" unsigned int loopy_propagate_until(Graph_t g, float convergence, unsigned int max_iterations)
{
    return dropin_run(g, "loopy_propagate_until", ITER, 0, convergence, max_iterations);
}

extern "C" unsigned int loopy_propagate_until_edge(Graph_t g, float convergence, unsigned int max_iterations)
{
    return dropin_run(g, "loopy_propagate_until_edge", EDGE, 0, convergence, max_iterations);
}

extern "C" unsigned int loopy_progagate_until_acc(Graph_t g, float convergence, unsigned int max_iterations)
{
    return dropin_run(g, "loopy_progagate_until_acc", ITER, 1, convergence, max_iterations);
}

extern "C" unsigned int loopy_progagate_until_edge_acc(Graph_t g, float convergence, unsigned int max_iterations)
{
    return dropin_run(g, "loopy_progagate_until_edge_acc", EDGE, 1, convergence, max_iterations);
}

extern "C" void loopy_propagate_edge_one_iteration(Graph_t g)
{
    bp_dev *d = bp_dropin_device(g, "loopy_propagate_edge_one_iteration");
    if (!d)
        return;
    if (bp_edge_iterate(d, 1) || bp_download(d, g))
        bp_fail("loopy_propagate_edge_one_iteration: %s", bp_last_error());
}

/* The CUDA driver's contract (belief-propagation.cu:529): the caller has run
 * init_previous_edge; the graph's states, tables and messages are uploaded,
 * iterated to convergence and downloaded in one call.  The device structure
 * is the graph's cached mirror (built on the first call, rebuilt when the
 * structure changes through the API), as for the other drop-in entry
 * points: rebuilding it on every call took about 90% of a call's time on a
 * 1M-arc network. */
static unsigned int hip_until(Graph_t g, int mode, float convergence, unsigned int max_iterations)
{
    bp_clear_error();
    bp_dev *d = dropin_device_status(g);
    if (!d)
        return 0xffffffffu;
    unsigned int result = 0;
    int rc = hipSetDevice(d->device) == hipSuccess ? 0 : -1;
    if (rc == 0)
        rc = run_until(d, mode, convergence, max_iterations, &result, nullptr);
    if (rc == 0)
        rc = bp_download(d, g);
    return rc == 0 ? result : 0xffffffffu;
}

extern "C" unsigned int loopy_propagate_until_hip(Graph_t g, float convergence, unsigned int max_iterations)
{
    return hip_until(g, ITER, convergence, max_iterations);
}

extern "C" unsigned int loopy_propagate_until_hip_edge(Graph_t g, float convergence, unsigned int max_iterations)
{
    return hip_until(g, EDGE, convergence, max_iterations);
}

/* graph.c:426-448 on the GPU: one edge, tables in MAX_STATES stride */
/* send_message's device buffer ([J | b | m]) and stream, one per device,
 * created on first use and kept: a caller looping over edges pays one
 * small copy each way per message, not an allocation and the synchronising
 * hipFree of each call.  The lock makes concurrent calls take turns. */
static constexpr int SEND_MAX_DEVICES = 64;
static std::mutex g_send_mu;
static struct {
    float *buf;
    hipStream_t stream;
} g_send[SEND_MAX_DEVICES];

extern "C" void send_message(float *states, unsigned int offset, unsigned int edge_index,
                             float *edge_joint_probabilities, float *edge_messages, unsigned int *edge_num_src,
                             unsigned int *edge_num_dest)
{
    constexpr size_t NJ = MAX_STATES * MAX_STATES;
    unsigned x = edge_num_src[edge_index], y = edge_num_dest[edge_index];
    if (x > MAX_STATES || y > MAX_STATES) {
        bp_fail("send_message: table %ux%u exceeds MAX_STATES", x, y);
        return;
    }
    float in[NJ + MAX_STATES] = {0}, m[MAX_STATES] = {0};
    for (unsigned i = 0; i < x; ++i)
        for (unsigned j = 0; j < y; ++j)
            in[i * MAX_STATES + j] = edge_joint_probabilities[(size_t)edge_index * NJ + i * MAX_STATES + j];
    for (unsigned j = 0; j < y; ++j)
        in[NJ + j] = states[offset + j];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SEND_MAX_DEVICES) {
        bp_fail("send_message: no usable HIP device");
        return;
    }
    std::lock_guard<std::mutex> lock(g_send_mu);
    auto &st = g_send[dev];
    if (!st.buf) {
        if (hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking) != hipSuccess ||
            hipMalloc((void **)&st.buf, (NJ + 2 * MAX_STATES) * sizeof(float)) != hipSuccess) {
            st.buf = nullptr;
            bp_fail("send_message: no usable HIP device");
            return;
        }
    }
    bool ok = hipMemcpyAsync(st.buf, in, sizeof(in), hipMemcpyHostToDevice, st.stream) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_send_one<MAX_STATES>, dim3(1), dim3(64), 0, st.stream, st.buf, st.buf + NJ, x, y,
                           st.buf + NJ + MAX_STATES);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(m, st.buf + NJ + MAX_STATES, sizeof(m), hipMemcpyDeviceToHost, st.stream) == hipSuccess &&
             hipStreamSynchronize(st.stream) == hipSuccess;
    }
    if (!ok) {
        bp_fail("send_message: the GPU send failed");
        return;
    }
    for (unsigned i = 0; i < x; ++i)
        edge_messages[(size_t)edge_index * MAX_STATES + i] = m[i];
}
