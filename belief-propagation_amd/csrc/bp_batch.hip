/*
 * bp_batch.hip -- many networks per launch.
 *
 * Layout: the batch is the disjoint union of its instances.  Nodes and slots
 * are numbered instance after instance; within an instance, nodes are in
 * the device order (Cuthill-McKee, shared by a run of instances over the
 * same node set) and each node's slots follow the reference's
 * destination-grouped edge order (src/graph/graph.c:285-313) as in
 * bp_upload.  Tables live in one pool deduplicated by content (the same arc
 * under every evidence set), each in the orientation its slots read it, rows
 * padded to four floats.  Chunks never span two instances, so each chunk's
 * FP64 delta partial belongs to one instance and k_batch_finish sums them
 * per instance (deterministic, no atomics).  Chunks run in an order that
 * merges the instances of a network by node range, so a network's evidence
 * sets read their shared tables close together in time.
 *
 * Rows are compact: a node's belief and combine-buffer rows hold its own
 * state count rounded up to an even number of floats, a slot's message rows
 * its x entries the same way, back to back in node (slot) order, instead of
 * every row padded to the batch's largest state count (on 2-8 states, 5
 * on average, 5.4 floats per row instead of 8).  slot_src holds the
 * source's row offset; each chunk carries the row offsets of its first
 * node and first slot, and the kernel finds the others by prefix sums of
 * the widths across the wave.
 *
 * The iteration kernel is k_loopy with packed tables and mixed state counts;
 * its arithmetic per node and slot is the plain path's, so every instance
 * gets the same bits as when it runs alone.
 */
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "bp_dev.h"
#include "../../include/bp_batch.h"

using namespace bpk;

namespace {

__global__ void k_copy_batch(const float *__restrict__ src, float *__restrict__ dst, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = src[i];
}

__global__ void k_zero_counter(unsigned *q)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *q = 0u;
}

/* deltas[i] = sum of instance i's chunk partials (chunks
 * [inst_chunk[i], inst_chunk[i + 1])): one wave per instance, lane l
 * summing chunks l, l + 64, ..., then a fixed-order wave reduction */
__global__ __launch_bounds__(256) void k_batch_finish(const double *__restrict__ partials,
                                                      const unsigned *__restrict__ inst_chunk, unsigned n_inst,
                                                      double *__restrict__ deltas)
{
    unsigned i = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    if (i >= n_inst)
        return;
    unsigned c0 = inst_chunk[i], c1 = inst_chunk[i + 1];
    double t = 0.0;
    for (unsigned c = c0 + lane; c < c1; c += WAVE)
        t += partials[c];
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1)
        t += __shfl_xor(t, o, WAVE);
    if (lane == 0)
        deltas[i] = t;
}

inline unsigned nblocks(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

uint64_t hash_table(unsigned x, unsigned y, const float *T, bool transposed)
{
    /* FNV-1a over (rows, cols, the entries row by row) */
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](uint32_t v) {
        for (int k = 0; k < 4; ++k) {
            h ^= (v >> (8 * k)) & 0xffu;
            h *= 1099511628211ull;
        }
    };
    unsigned rows = transposed ? y : x, cols = transposed ? x : y;
    mix(rows);
    mix(cols);
    for (unsigned i = 0; i < rows; ++i)
        for (unsigned j = 0; j < cols; ++j) {
            float v = transposed ? T[j * MAX_STATES + i] : T[i * MAX_STATES + j];
            uint32_t u;
            std::memcpy(&u, &v, 4);
            mix(u);
        }
    return h;
}

/* Chunks in instance order (chunk c ends where chunk c + 1 begins; a
 * sentinel closes the list) and the order they run in.  The instances of a
 * group -- consecutive instances over the same node set, i.e. one network
 * under several evidence sets -- run merged by starting node, so the
 * group's copies of a node range run back to back and read the same tables
 * while they are cached. */
struct Shape {
    unsigned R = 1;
    std::vector<uint2> chunks;
    std::vector<uint2> base;          /* [nchunks] row offsets of each chunk's first node and first slot */
    uint2 *d_base = nullptr;
    std::vector<unsigned> wide;       /* chunks of one node with more than 64 R in-slots (k_loopy_wide) */
    unsigned *d_wide = nullptr;
    std::vector<unsigned> inst_chunk; /* [instances+1] */
    std::vector<unsigned> order;      /* [nchunks] */
    uint2 *d_chunk = nullptr;
    unsigned *d_inst_chunk = nullptr, *d_order = nullptr;
    unsigned nchunks = 0;
    bool built = false;
};

struct UniqueTable {
    unsigned off, x, y;
};

} // namespace

struct bp_batch {
    /* host staging */
    unsigned S = 1;
    std::vector<unsigned> node_off{0}, slot_off{0};
    std::vector<unsigned char> nvars;
    std::vector<float> st_stage;   /* [V][MAX_STATES] */
    std::vector<unsigned> in_off;  /* [V] global */
    std::vector<unsigned> slot_src, slot_tab, slot_edge;
    std::vector<unsigned char> slot_xy;
    std::vector<unsigned> edge_x;  /* per slot: x of the slot's edge (for downloads) */
    std::vector<float> tabs;
    std::vector<uint64_t> inst_sig; /* [instances] hash of the node set (node count, state counts) */
    std::vector<unsigned> node_inv; /* [V] graph node (within its instance) at each device node */
    std::vector<unsigned> last_inv; /* the device order of the last instance added */
    std::vector<UniqueTable> uniq;
    std::unordered_map<uint64_t, unsigned> lookup; /* hash -> index in uniq */
    /* compact rows: floats of node rows and of message rows before each
     * instance ([instances+1]) */
    std::vector<unsigned long long> node_row_off{0}, msg_row_off{0};
    double alg_bytes = 0.0;
    uint64_t distinct_floats = 0; /* a table and its transpose counted once */
    bool uploaded = false;

    /* device */
    int device = 0;
    hipStream_t stream = nullptr;
    unsigned V = 0, E = 0, SP = 1, n_inst = 0;
    int cus = 256;
    unsigned *d_in_off = nullptr, *d_slot_src = nullptr, *d_slot_tab = nullptr, *d_queue = nullptr;
    unsigned char *d_slot_xy = nullptr, *d_nvars = nullptr;
    float *d_tabs = nullptr, *d_states = nullptr, *d_states0 = nullptr, *d_buf[2] = {nullptr, nullptr};
    float *d_msg[2] = {nullptr, nullptr};
    double *d_partials = nullptr;
    double *d_deltas = nullptr;
    unsigned deltas_cap = 0;
    Shape shapes[2];

    /* configuration and run state */
    unsigned R = 1, L = 1, qblock = 0;
    int latest = 0, bcur = 0;
    unsigned swaps = 0;
    bool ready = false;
    std::map<std::string, hipGraphExec_t> graphs;
};

static void batch_clear_graphs(bp_batch *b)
{
    for (auto &kv : b->graphs)
        (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

extern "C" int bp_batch_create(bp_batch **out)
{
    bp_clear_error();
    if (!out) {
        bp_set_error("bp_batch_create: NULL argument");
        return -1;
    }
    *out = new bp_batch();
    return 0;
}

extern "C" void bp_batch_free(bp_batch *b)
{
    if (!b)
        return;
    if (b->uploaded) {
        (void)hipSetDevice(b->device);
        if (b->stream)
            (void)hipStreamSynchronize(b->stream);
        batch_clear_graphs(b);
        void *ps[] = {b->d_in_off, b->d_slot_src, b->d_slot_tab, b->d_queue, b->d_slot_xy, b->d_nvars, b->d_tabs,
                      b->d_states, b->d_states0, b->d_buf[0], b->d_buf[1], b->d_msg[0], b->d_msg[1],
                      b->d_partials, b->d_deltas, b->shapes[0].d_chunk, b->shapes[0].d_inst_chunk,
                      b->shapes[0].d_order, b->shapes[0].d_base, b->shapes[0].d_wide, b->shapes[1].d_chunk,
                      b->shapes[1].d_inst_chunk, b->shapes[1].d_order, b->shapes[1].d_base, b->shapes[1].d_wide};
        for (void *p : ps)
            if (p)
                (void)hipFree(p);
        if (b->stream)
            (void)hipStreamDestroy(b->stream);
    }
    delete b;
}

/* Row stride of a pooled table with y columns: rows padded with zeros to a
 * multiple of four floats, so every row starts 16-byte aligned. */
static inline unsigned pooled_stride(unsigned y) { return (y + 3u) & ~3u; }

/* index of the pooled table equal to T (x rows, y cols, MAX_STATES-strided),
 * appended when new.  The pool holds each table in the orientation its
 * slots read it; a table whose transpose is already pooled (the reverse of
 * an edge) is stored again, but counted once in the distinct floats. */
static unsigned intern_table(bp_batch *b, unsigned x, unsigned y, const float *T)
{
    auto same = [&](const UniqueTable &u, bool transposed) {
        unsigned rows = transposed ? y : x, cols = transposed ? x : y, st = pooled_stride(cols);
        if (u.x != rows || u.y != cols)
            return false;
        for (unsigned i = 0; i < rows; ++i)
            for (unsigned j = 0; j < cols; ++j) {
                float v = transposed ? T[j * MAX_STATES + i] : T[i * MAX_STATES + j];
                if (std::memcmp(&v, &b->tabs[u.off + i * st + j], 4) != 0)
                    return false;
            }
        return true;
    };
    uint64_t h = hash_table(x, y, T, false);
    auto it = b->lookup.find(h);
    if (it != b->lookup.end() && same(b->uniq[it->second], false))
        return it->second;
    uint64_t ht = hash_table(x, y, T, true);
    auto jt = b->lookup.find(ht);
    bool transpose_pooled = jt != b->lookup.end() && same(b->uniq[jt->second], true);
    unsigned st = pooled_stride(y);
    UniqueTable u{(unsigned)b->tabs.size(), x, y};
    for (unsigned i = 0; i < x; ++i)
        for (unsigned j = 0; j < st; ++j)
            b->tabs.push_back(j < y ? T[i * MAX_STATES + j] : 0.0f);
    b->uniq.push_back(u);
    unsigned idx = (unsigned)b->uniq.size() - 1;
    if (it == b->lookup.end())
        b->lookup.emplace(h, idx);
    if (!transpose_pooled) {
        b->alg_bytes += 4.0 * x * y;
        b->distinct_floats += (uint64_t)x * y;
    }
    return idx;
}

extern "C" int bp_batch_add(bp_batch *b, Graph_t g, unsigned int *index)
{
    bp_clear_error();
    if (!b || !g) {
        bp_set_error("bp_batch_add: NULL argument");
        return -1;
    }
    if (b->uploaded) {
        bp_set_error("bp_batch_add: the batch is already on the device");
        return -1;
    }
    unsigned V = g->current_num_vertices, E = g->current_num_edges;
    if (V != g->total_num_vertices) {
        bp_set_error("bp_batch_add: graph has %u of %u nodes", V, g->total_num_vertices);
        return -1;
    }
    for (unsigned v = 0; v < V; ++v) {
        unsigned bgn = g->dest_nodes_to_edges_node_list[v], end = bp_csr_end(g, g->dest_nodes_to_edges_node_list, v);
        if (bgn > end || end > E) {
            bp_set_error("bp_batch_add: destination CSR is not set up (call set_up_dest_nodes_to_edges)");
            return -1;
        }
        for (unsigned k = bgn; k < end; ++k) {
            unsigned e = g->dest_nodes_to_edges_edge_list[k];
            if (e >= E || g->edges_dest_index[e] != v) {
                bp_set_error("bp_batch_add: destination CSR is not set up (call set_up_dest_nodes_to_edges)");
                return -1;
            }
        }
    }
    if (V > 0 && bp_csr_end(g, g->dest_nodes_to_edges_node_list, V - 1) != E) {
        bp_set_error("bp_batch_add: destination CSR does not cover all %u edges", E);
        return -1;
    }
    uint64_t nv = (uint64_t)b->node_off.back() + V, ne = (uint64_t)b->slot_off.back() + E;
    if (nv >= 0xffffffffull || ne >= 0xffffffffull) {
        bp_set_error("bp_batch_add: the batch exceeds 32-bit node or edge indices");
        return -1;
    }
    for (unsigned v = 0; v < V; ++v)
        if (g->node_num_vars[v] == 0 || g->node_num_vars[v] > MAX_STATES) {
            bp_set_error("bp_batch_add: node %u has %u states", v, g->node_num_vars[v]);
            return -1;
        }
    for (unsigned e = 0; e < E; ++e) {
        unsigned s = g->edges_src_index[e], t = g->edges_dest_index[e];
        if (s >= V || t >= V || g->edges_x_dim[e] != g->node_num_vars[s] || g->edges_y_dim[e] != g->node_num_vars[t]) {
            bp_set_error("bp_batch_add: edge %u is inconsistent with its nodes", e);
            return -1;
        }
    }
    unsigned n0 = b->node_off.back(), s0 = b->slot_off.back();
    uint64_t sig = 1469598103934665603ull;
    auto mix = [&sig](uint32_t u) {
        sig ^= u;
        sig *= 1099511628211ull;
    };
    mix(V);
    for (unsigned v = 0; v < V; ++v)
        mix(g->node_num_vars[v]);
    /* device node order: Cuthill-McKee (or the graph's order, per
     * bp_set_node_order), computed once per run of instances over the same
     * node set so that their node ranges keep matching */
    bool same_set = !b->inst_sig.empty() && b->inst_sig.back() == sig && b->last_inv.size() == V;
    if (!same_set) {
        if (bp_node_order_mode() && V > 1) {
            bp_cm_order(g, b->last_inv);
        } else {
            b->last_inv.resize(V);
            for (unsigned v = 0; v < V; ++v)
                b->last_inv[v] = v;
        }
    }
    const std::vector<unsigned> &inv = b->last_inv;
    std::vector<unsigned> perm(V);
    for (unsigned i = 0; i < V; ++i)
        perm[inv[i]] = i;
    /* compact rows: each device node's row offset, then every slot's
     * message widths */
    const unsigned long long r0 = b->node_row_off.back(), m0 = b->msg_row_off.back();
    std::vector<unsigned long long> row(V);
    unsigned long long rt = r0, mt = m0;
    for (unsigned i = 0; i < V; ++i) {
        row[i] = rt;
        rt += (g->node_num_vars[inv[i]] + 1u) & ~1u;
    }
    for (unsigned e = 0; e < E; ++e)
        mt += (g->edges_x_dim[e] + 1u) & ~1u;
    if (rt >= 0xffffffffull || mt >= 0xffffffffull) {
        bp_set_error("bp_batch_add: the batch exceeds 2^32 floats of rows");
        return -1;
    }
    unsigned slots = 0;
    for (unsigned i = 0; i < V; ++i) {
        unsigned v = inv[i];
        unsigned n = g->node_num_vars[v];
        b->S = std::max(b->S, n);
        b->nvars.push_back((unsigned char)n);
        b->node_inv.push_back(v);
        for (unsigned k = 0; k < MAX_STATES; ++k)
            b->st_stage.push_back(k < n ? g->node_states[(size_t)v * MAX_STATES + k] : 0.0f);
        b->in_off.push_back(s0 + slots);
        b->alg_bytes += 12.0 * n + 4.0;
        for (unsigned k = g->dest_nodes_to_edges_node_list[v]; k < bp_csr_end(g, g->dest_nodes_to_edges_node_list, v);
             ++k, ++slots) {
            unsigned e = g->dest_nodes_to_edges_edge_list[k];
            unsigned x = g->edges_x_dim[e], y = g->edges_y_dim[e];
            if (b->tabs.size() + (size_t)x * pooled_stride(y) >= 0xffffffffull) {
                bp_set_error("bp_batch_add: the table pool exceeds 2^32 floats");
                return -1;
            }
            unsigned u = intern_table(b, x, y, g->edges_joint_probabilities + (size_t)e * MAX_STATES * MAX_STATES);
            b->slot_src.push_back((unsigned)row[perm[g->edges_src_index[e]]]);
            b->slot_tab.push_back(b->uniq[u].off);
            b->slot_xy.push_back((unsigned char)(x | (y << 4)));
            b->slot_edge.push_back(e);
            b->edge_x.push_back(x);
            b->alg_bytes += 4.0 * (2.0 * x + y + 1.0);
        }
    }
    b->node_off.push_back(n0 + V);
    b->slot_off.push_back(s0 + E);
    b->node_row_off.push_back(rt);
    b->msg_row_off.push_back(mt);
    b->inst_sig.push_back(sig);
    if (index)
        *index = (unsigned)b->node_off.size() - 2;
    return 0;
}

/* chunks per instance, never across instances (cap = 64 x slots per lane);
 * then the order they run in: each group of up to MAX_GROUP consecutive
 * instances with equal node-set signatures merged by (first node within
 * the instance, instance).  The order only decides which chunks run
 * together: every order gives the same results. */
constexpr unsigned MAX_GROUP = 16;

/* Most slots (and nodes) in a chunk of one slot per lane; a measuring
 * switch (with two lanes per slot, a chunk of 32 slots takes one pass) */
#ifndef BP_BATCH_CHUNK_CAP
#define BP_BATCH_CHUNK_CAP 128
#endif

static void build_shape(bp_batch *b, Shape &sh, unsigned R)
{
    sh.R = R;
    sh.chunks.clear();
    sh.inst_chunk.assign(1, 0);
    unsigned cap = std::min<unsigned>(WAVE * R, BP_BATCH_CHUNK_CAP);
    unsigned ninst = (unsigned)b->node_off.size() - 1;
    for (unsigned i = 0; i < ninst; ++i) {
        unsigned v = b->node_off[i], vend = b->node_off[i + 1];
        while (v < vend) {
            unsigned s0 = b->in_off[v];
            sh.chunks.push_back(make_uint2(v, s0));
            unsigned w = v;
            while (w < vend && w - v < cap && (w + 1 < b->V ? b->in_off[w + 1] : b->E) - s0 <= cap)
                ++w;
            if (w == v)
                w = v + 1;
            v = w;
        }
        sh.inst_chunk.push_back((unsigned)sh.chunks.size());
    }
    sh.nchunks = (unsigned)sh.chunks.size();
    sh.chunks.push_back(make_uint2(b->V, b->E));
    /* row offsets of each chunk's first node and first slot: prefix sums
     * of the even-rounded widths, walked along with the chunks */
    sh.base.resize(sh.nchunks);
    {
        unsigned long long nr = 0, mr = 0;
        unsigned v = 0, s = 0;
        for (unsigned c = 0; c < sh.nchunks; ++c) {
            for (; v < sh.chunks[c].x; ++v)
                nr += (b->nvars[v] + 1u) & ~1u;
            for (; s < sh.chunks[c].y; ++s)
                mr += (b->edge_x[s] + 1u) & ~1u;
            sh.base[c] = make_uint2((unsigned)nr, (unsigned)mr);
        }
    }
    sh.wide.clear();
    for (unsigned c = 0; c < sh.nchunks; ++c)
        if (sh.chunks[c + 1].y - sh.chunks[c].y > WAVE * R) /* beyond what the main kernel takes */
            sh.wide.push_back(c);
    sh.order.clear();
    sh.order.reserve(sh.nchunks);
    std::vector<std::pair<uint64_t, unsigned>> keys;
    for (unsigned i = 0; i < ninst;) {
        unsigned j = i + 1;
        while (j < ninst && j - i < MAX_GROUP && b->inst_sig[j] == b->inst_sig[i])
            ++j;
        keys.clear();
        for (unsigned k = i; k < j; ++k)
            for (unsigned c = sh.inst_chunk[k]; c < sh.inst_chunk[k + 1]; ++c)
                keys.emplace_back(((uint64_t)(sh.chunks[c].x - b->node_off[k]) << 32) | (k - i), c);
        std::sort(keys.begin(), keys.end());
        for (const auto &kv : keys)
            sh.order.push_back(kv.second);
        i = j;
    }
    sh.built = true;
}

template <typename T> static int up(T **dst, const T *src, size_t n)
{
    if (dmalloc(dst, n))
        return -1;
    if (n)
        HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int bp_batch_upload(bp_batch *b, int device)
{
    bp_clear_error();
    if (!b) {
        bp_set_error("bp_batch_upload: NULL argument");
        return -1;
    }
    if (b->uploaded) {
        bp_set_error("bp_batch_upload: already uploaded");
        return -1;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        bp_set_error("no HIP device available: the belief-propagation kernels need an MI355X (gfx950)");
        return -1;
    }
    if (device < 0)
        HIP_TRY(hipGetDevice(&device));
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        bp_set_error("device %d is %s; libbp_amd.so is built for gfx950 only", device, prop.gcnArchName);
        return -1;
    }
    b->device = device;
    b->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    b->V = b->node_off.back();
    b->E = b->slot_off.back();
    b->n_inst = (unsigned)b->node_off.size() - 1;
    b->SP = stride_of(b->S);
    build_shape(b, b->shapes[0], 1);
    if (b->S <= 6)
        build_shape(b, b->shapes[1], 2);
    b->R = (b->S <= 4) ? 2 : 1;
    b->qblock = 0;
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->uploaded = true; /* from here on bp_batch_free releases device memory */

    /* compact rows; floats of slack past the last row, which the kernels'
     * unpredicated row loads (up to 8 floats, or S) may read */
    const size_t slack = std::max(b->S, 8u) + 1;
    size_t VS = b->node_row_off.back() + slack, ES = b->msg_row_off.back() + slack;
    std::vector<float> st(VS, 0.0f);
    {
        size_t r = 0;
        for (size_t v = 0; v < b->V; ++v) {
            for (unsigned k = 0; k < b->nvars[v]; ++k)
                st[r + k] = b->st_stage[v * MAX_STATES + k];
            r += (b->nvars[v] + 1u) & ~1u;
        }
    }
    /* zeros past the pool's last table: a lane reads every row of its pass
     * (row indices below 16 = lanes per slot x rows per lane, at the
     * table's own stride of at most 12 floats) as whole four-float vectors
     * (at most 12 floats), whatever the table's row count */
    b->tabs.resize(b->tabs.size() + (size_t)16 * pooled_stride(MAX_STATES), 0.0f);
    std::vector<unsigned> in_off(b->in_off);
    in_off.push_back(b->E);
    unsigned maxchunks = std::max(b->shapes[0].nchunks, b->shapes[1].nchunks);
    if (up(&b->d_in_off, in_off.data(), in_off.size()) || up(&b->d_slot_src, b->slot_src.data(), b->E) ||
        up(&b->d_slot_tab, b->slot_tab.data(), b->E) || up(&b->d_slot_xy, b->slot_xy.data(), b->E) ||
        up(&b->d_nvars, b->nvars.data(), b->V) || up(&b->d_tabs, b->tabs.data(), b->tabs.size()) ||
        up(&b->d_states0, st.data(), VS) || dmalloc(&b->d_states, VS) || dmalloc(&b->d_buf[0], VS) ||
        dmalloc(&b->d_buf[1], VS) || dmalloc(&b->d_msg[0], ES) || dmalloc(&b->d_msg[1], ES) ||
        dmalloc(&b->d_partials, std::max(maxchunks, 1u)) || dmalloc(&b->d_queue, 1))
        return -1;
    for (int k = 0; k < 2; ++k) {
        Shape &sh = b->shapes[k];
        if (!sh.built)
            continue;
        if (up(&sh.d_chunk, sh.chunks.data(), sh.chunks.size()) ||
            up(&sh.d_inst_chunk, sh.inst_chunk.data(), sh.inst_chunk.size()) ||
            up(&sh.d_order, sh.order.data(), sh.order.size()) || up(&sh.d_base, sh.base.data(), sh.base.size()) ||
            up(&sh.d_wide, sh.wide.data(), sh.wide.size()))
            return -1;
    }
    HIP_TRY(hipMemcpy(b->d_states, b->d_states0, VS * sizeof(float), hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemset(b->d_buf[0], 0, VS * sizeof(float)));
    HIP_TRY(hipMemset(b->d_buf[1], 0, VS * sizeof(float)));
    HIP_TRY(hipMemset(b->d_msg[0], 0, ES * sizeof(float)));
    HIP_TRY(hipMemset(b->d_msg[1], 0, ES * sizeof(float)));
    /* host staging no longer needed (downloads keep slot_edge, edge_x, offsets, nvars) */
    std::vector<float>().swap(b->st_stage);
    std::vector<unsigned>().swap(b->slot_src);
    std::vector<unsigned>().swap(b->slot_tab);
    std::vector<unsigned>().swap(b->last_inv);
    std::vector<unsigned char>().swap(b->slot_xy);
    std::vector<float>().swap(b->tabs);
    b->lookup.clear();
    return 0;
}

static bool valid_config(const bp_batch *b, unsigned R, unsigned L, unsigned qblock)
{
    (void)qblock;
    if (L == 1) /* two slots per lane would spill registers beyond six states */
        return (R == 1) || (R == 2 && b->S <= 6);
    /* several lanes per slot: one slot per lane-group, four or more states */
    return R == 1 && b->S >= 4 && (L == 2 || L == 4 || L == 8);
}

extern "C" int bp_batch_set_config(bp_batch *b, unsigned int slots_per_lane, unsigned int lanes_per_slot,
                                   unsigned int queue_block)
{
    bp_clear_error();
    if (!b || !b->uploaded) {
        bp_set_error("bp_batch_set_config: batch not uploaded");
        return -1;
    }
    if (!valid_config(b, slots_per_lane, lanes_per_slot, queue_block)) {
        bp_set_error("bp_batch_set_config: %u slots per lane with %u lanes per slot is not available for %u states",
                     slots_per_lane, lanes_per_slot, b->S);
        return -1;
    }
    b->R = slots_per_lane;
    b->L = lanes_per_slot;
    b->qblock = queue_block;
    return 0;
}

extern "C" int bp_batch_get_config(const bp_batch *b, unsigned int *slots_per_lane, unsigned int *lanes_per_slot,
                                   unsigned int *queue_block)
{
    if (!b || !slots_per_lane || !lanes_per_slot || !queue_block) {
        bp_set_error("bp_batch_get_config: NULL argument");
        return -1;
    }
    *slots_per_lane = b->R;
    *lanes_per_slot = b->L;
    *queue_block = b->qblock;
    return 0;
}

static const Shape &cur_shape(const bp_batch *b) { return b->shapes[b->R == 2 ? 1 : 0]; }

static LoopyArgs batch_args(bp_batch *b)
{
    const Shape &sh = cur_shape(b);
    LoopyArgs a;
    a.in_off = b->d_in_off;
    a.slot_src = b->d_slot_src;
    a.slot_xy = b->d_slot_xy;
    a.J = nullptr;
    a.nvars = b->d_nvars;
    a.states = b->d_states;
    a.states_out = b->d_states;
    a.buf_in = b->d_buf[b->bcur];
    a.buf_out = b->d_buf[b->bcur ^ 1];
    a.msg_in = b->d_msg[b->latest];
    a.msg_out = b->d_msg[b->latest ^ 1];
    a.chunk = sh.d_chunk;
    a.chunk_base = sh.d_base;
    a.order = sh.d_order;
    a.nchunks = sh.nchunks;
    a.partials = nullptr;
    a.stop = nullptr;
    a.slot_tab = b->d_slot_tab;
    a.tabs = b->d_tabs;
    a.chunk_partials = b->d_partials;
    a.queue = b->qblock ? b->d_queue : nullptr;
    a.qblock = b->qblock ? b->qblock : 1;
    return a;
}

static unsigned batch_grid(const bp_batch *b)
{
    const Shape &sh = cur_shape(b);
    unsigned want = (sh.nchunks + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    unsigned G = std::max(8u, (want + 7u) & ~7u);
    if (b->qblock) /* persistent: eight workgroups (32 waves) per CU */
        G = std::min(G, std::max(8u, (unsigned)b->cus * 8u));
    return G;
}

template <int S, int R, int L>
static void launch_batch_srl(int mode, const LoopyArgs &a, unsigned G, hipStream_t st, const hipEvent_t *ev,
                             const Shape &sh)
{
    dim3 grid(G), block(BLOCK);
    unsigned nw = (unsigned)sh.wide.size(), gw = (nw + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    if (mode == ITER) {
        if (ev)
            hipExtLaunchKernelGGL((k_loopy<S, true, ITER, R, true, L>), grid, block, 0, st, ev[0], ev[1], 0, a);
        else
            hipLaunchKernelGGL((k_loopy<S, true, ITER, R, true, L>), grid, block, 0, st, a);
        if (nw)
            hipLaunchKernelGGL((k_loopy_wide<S, true, ITER>), dim3(gw), block, 0, st, a, sh.d_wide, nw);
    } else {
        hipLaunchKernelGGL((k_loopy<S, true, INIT, R, true, L>), grid, block, 0, st, a);
        if (nw)
            hipLaunchKernelGGL((k_loopy_wide<S, true, INIT>), dim3(gw), block, 0, st, a, sh.d_wide, nw);
    }
}

/* 1: batches of 5 to 7 states run the two-lane shape on the 8-state kernel
 * (a compile-time switch so that the comparison can be repeated) */
#ifndef BP_BATCH_ROUND_STATES
#define BP_BATCH_ROUND_STATES 1
#endif

template <int S>
static void launch_batch_s(unsigned R, unsigned L, int mode, const LoopyArgs &a, unsigned G, hipStream_t st,
                           const hipEvent_t *ev, const Shape &sh)
{
    if constexpr (S <= 6) {
        if (R == 2) {
            launch_batch_srl<S, 2, 1>(mode, a, G, st, ev, sh);
            return;
        }
    }
    if constexpr (S >= 4) {
        /* two lanes per slot: 5 to 7 states run the 8-state kernel (every
         * read is masked by the slot's and node's own counts, so the results
         * are the same).  Compiled for 5-7 states that shape needs 115-159
         * registers, for 8 states 80; on mixed 2-7 state batches its
         * iteration is 1.33 ms against 2.38, and the best shape 8% faster.
         * With 4 or 8 lanes the own-count kernels are the faster ones. */
        constexpr int KS = (BP_BATCH_ROUND_STATES && S >= 5 && S <= 7) ? 8 : S;
        if (L == 2) {
            launch_batch_srl<KS, 1, 2>(mode, a, G, st, ev, sh);
            return;
        }
        if (L == 4) {
            launch_batch_srl<S, 1, 4>(mode, a, G, st, ev, sh);
            return;
        }
        if (L == 8) {
            launch_batch_srl<S, 1, 8>(mode, a, G, st, ev, sh);
            return;
        }
    }
    launch_batch_srl<S, 1, 1>(mode, a, G, st, ev, sh);
}

static void launch_batch(bp_batch *b, int mode, const LoopyArgs &a, hipStream_t st, const hipEvent_t *ev = nullptr)
{
    if (b->V == 0)
        return;
    if (b->qblock)
        hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, b->d_queue);
    unsigned G = batch_grid(b);
    const Shape &sh = cur_shape(b);
    switch (b->S) {
    case 1: launch_batch_s<1>(b->R, b->L, mode, a, G, st, ev, sh); break;
    case 2: launch_batch_s<2>(b->R, b->L, mode, a, G, st, ev, sh); break;
    case 3: launch_batch_s<3>(b->R, b->L, mode, a, G, st, ev, sh); break;
    case 4: launch_batch_s<4>(b->R, b->L, mode, a, G, st, ev, sh); break;
    case 5: launch_batch_s<5>(b->R, b->L, mode, a, G, st, ev, sh); break;
    case 6: launch_batch_s<6>(b->R, b->L, mode, a, G, st, ev, sh); break;
    case 7: launch_batch_s<7>(b->R, b->L, mode, a, G, st, ev, sh); break;
    case 8: launch_batch_s<8>(b->R, b->L, mode, a, G, st, ev, sh); break;
    case 9: launch_batch_s<9>(b->R, b->L, mode, a, G, st, ev, sh); break;
    default: launch_batch_s<10>(b->R, b->L, mode, a, G, st, ev, sh); break;
    }
}

extern "C" int bp_batch_init(bp_batch *b)
{
    bp_clear_error();
    if (!b || !b->uploaded) {
        bp_set_error("bp_batch_init: batch not uploaded");
        return -1;
    }
    HIP_TRY(hipSetDevice(b->device));
    size_t VS = b->node_row_off.back();
    if (VS)
        hipLaunchKernelGGL(k_copy_batch, dim3(nblocks(VS)), dim3(256), 0, b->stream, b->d_states0, b->d_states, VS);
    b->latest = 0;
    b->bcur = 0;
    b->swaps = 0;
    LoopyArgs a = batch_args(b);
    a.msg_out = b->d_msg[b->latest];
    a.buf_out = b->d_buf[b->bcur];
    launch_batch(b, INIT, a, b->stream);
    HIP_TRY(hipGetLastError());
    b->ready = true;
    return 0;
}

static int ensure_deltas(bp_batch *b, unsigned n)
{
    size_t need = (size_t)n * b->n_inst;
    if (need <= b->deltas_cap)
        return 0;
    if (b->d_deltas)
        HIP_TRY(hipFree(b->d_deltas));
    b->d_deltas = nullptr;
    batch_clear_graphs(b); /* captured graphs hold the old address */
    if (dmalloc(&b->d_deltas, need))
        return -1;
    b->deltas_cap = (unsigned)need;
    return 0;
}

extern "C" int bp_batch_iterate(bp_batch *b, unsigned int n, double *deltas)
{
    bp_clear_error();
    if (!b || !b->uploaded || !b->ready) {
        bp_set_error("bp_batch_iterate: call bp_batch_init first");
        return -1;
    }
    if (n == 0)
        return 0;
    HIP_TRY(hipSetDevice(b->device));
    if (ensure_deltas(b, n))
        return -1;
    char key[96];
    snprintf(key, sizeof(key), "%u/%u/%u/%u/%d/%d", n, b->R, b->L, b->qblock, b->latest, b->bcur);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int lat = b->latest, bc = b->bcur;
        const Shape &sh = cur_shape(b);
        for (unsigned i = 0; i < n; ++i) {
            LoopyArgs a = batch_args(b);
            a.buf_in = b->d_buf[bc];
            a.buf_out = b->d_buf[bc ^ 1];
            a.msg_in = b->d_msg[lat];
            a.msg_out = b->d_msg[lat ^ 1];
            launch_batch(b, ITER, a, b->stream);
            hipLaunchKernelGGL(k_batch_finish, dim3(nblocks(b->n_inst, WAVES_PER_BLOCK)), dim3(256), 0, b->stream,
                               b->d_partials, sh.d_inst_chunk, b->n_inst, b->d_deltas + (size_t)i * b->n_inst);
            lat ^= 1;
            bc ^= 1;
        }
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamEndCapture(b->stream, &graph));
        hipGraphExec_t exec = nullptr;
        hipError_t err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (err != hipSuccess) {
            bp_set_error("hipGraphInstantiate failed: %s", hipGetErrorString(err));
            return -1;
        }
        it = b->graphs.emplace(key, exec).first;
    }
    HIP_TRY(hipGraphLaunch(it->second, b->stream));
    if (n & 1u) {
        b->latest ^= 1;
        b->bcur ^= 1;
    }
    b->swaps += n;
    if (deltas) {
        HIP_TRY(hipMemcpyAsync(deltas, b->d_deltas, (size_t)n * b->n_inst * sizeof(double), hipMemcpyDeviceToHost,
                               b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
    }
    return 0;
}

extern "C" int bp_batch_sync(bp_batch *b)
{
    bp_clear_error();
    if (!b || !b->uploaded) {
        bp_set_error("bp_batch_sync: batch not uploaded");
        return -1;
    }
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

extern "C" int bp_batch_download(bp_batch *b, unsigned int index, Graph_t g)
{
    bp_clear_error();
    if (!b || !g || !b->uploaded) {
        bp_set_error("bp_batch_download: NULL argument or batch not uploaded");
        return -1;
    }
    if (index >= b->n_inst) {
        bp_set_error("bp_batch_download: instance %u of %u", index, b->n_inst);
        return -1;
    }
    unsigned n0 = b->node_off[index], V = b->node_off[index + 1] - n0;
    unsigned s0 = b->slot_off[index], E = b->slot_off[index + 1] - s0;
    if (g->current_num_vertices != V || g->current_num_edges != E) {
        bp_set_error("bp_batch_download: graph does not match instance %u", index);
        return -1;
    }
    HIP_TRY(hipSetDevice(b->device));
    size_t r0 = b->node_row_off[index], rn = b->node_row_off[index + 1] - r0;
    size_t q0 = b->msg_row_off[index], qn = b->msg_row_off[index + 1] - q0;
    std::vector<float> st(rn), m0(qn), m1(qn);
    if (rn)
        HIP_TRY(hipMemcpyAsync(st.data(), b->d_states + r0, rn * 4, hipMemcpyDeviceToHost, b->stream));
    if (qn) {
        HIP_TRY(hipMemcpyAsync(m0.data(), b->d_msg[b->latest] + q0, qn * 4, hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipMemcpyAsync(m1.data(), b->d_msg[b->latest ^ 1] + q0, qn * 4, hipMemcpyDeviceToHost, b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (size_t i = 0, r = 0; i < V; ++i) {
        unsigned v = b->node_inv[n0 + i], n = b->nvars[n0 + i];
        for (unsigned k = 0; k < n; ++k)
            g->node_states[(size_t)v * MAX_STATES + k] = st[r + k];
        r += (n + 1u) & ~1u;
    }
    if (b->swaps & 1u) {
        float **t = g->previous_edge_messages;
        g->previous_edge_messages = g->current_edge_messages;
        g->current_edge_messages = t;
    }
    float *prev = *g->previous_edge_messages, *cur = *g->current_edge_messages;
    for (size_t k = 0, r = 0; k < E; ++k) {
        unsigned e = b->slot_edge[s0 + k], x = b->edge_x[s0 + k];
        for (unsigned i = 0; i < x; ++i) {
            prev[(size_t)e * MAX_STATES + i] = m0[r + i];
            cur[(size_t)e * MAX_STATES + i] = m1[r + i];
        }
        r += (x + 1u) & ~1u;
    }
    return 0;
}

extern "C" int bp_batch_info(const bp_batch *b, uint64_t *info)
{
    if (!b || !info) {
        bp_set_error("bp_batch_info: NULL argument");
        return -1;
    }
    info[0] = b->node_off.size() - 1;
    info[1] = b->node_off.back();
    info[2] = b->slot_off.back();
    info[3] = b->S;
    info[4] = b->distinct_floats;
    info[5] = b->uploaded ? cur_shape(b).nchunks : 0;
    info[6] = (uint64_t)b->alg_bytes;
    return 0;
}

extern "C" int bp_batch_profile(bp_batch *b, unsigned int n, float *ms)
{
    bp_clear_error();
    if (!b || !ms || !b->uploaded || !b->ready || n == 0) {
        bp_set_error("bp_batch_profile: bad arguments or batch not initialised");
        return -1;
    }
    HIP_TRY(hipSetDevice(b->device));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float total = 0.0f;
    for (unsigned i = 0; i < n; ++i) {
        LoopyArgs a = batch_args(b);
        hipEvent_t ev[2] = {e0, e1};
        launch_batch(b, ITER, a, b->stream, ev);
        HIP_TRY(hipStreamSynchronize(b->stream));
        float t = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        total += t;
        b->latest ^= 1;
        b->bcur ^= 1;
        b->swaps++;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = total / n;
    return 0;
}

/* ---- autotuner ------------------------------------------------------------- */

static std::string batch_signature(const bp_batch *b)
{
    hipDeviceProp_t prop;
    std::string dev = "gpu";
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess)
        dev = prop.gcnArchName;
    for (char &c : dev)
        if (c == '"' || c == ':')
            c = '_';
    char buf[256];
    snprintf(buf, sizeof(buf), "%s/cu%d/inst%u/V%u/E%u/S%u", dev.c_str(), b->cus, b->n_inst, b->V, b->E, b->S);
    return buf;
}

static bool cache_lookup(const char *path, const std::string &sig, unsigned *R, unsigned *L, unsigned *q)
{
    if (!path)
        return false;
    FILE *f = fopen(path, "r");
    if (!f)
        return false;
    std::string text;
    char chunk[4096];
    size_t n;
    while ((n = fread(chunk, 1, sizeof(chunk), f)) > 0)
        text.append(chunk, n);
    fclose(f);
    std::string key = "\"" + sig + "\"";
    size_t pos = text.find(key);
    if (pos == std::string::npos)
        return false;
    size_t r = text.find("\"slots_per_lane\":", pos), l = text.find("\"lanes_per_slot\":", pos),
           qq = text.find("\"queue_block\":", pos);
    if (r == std::string::npos || l == std::string::npos || qq == std::string::npos)
        return false;
    *R = (unsigned)strtoul(text.c_str() + r + 17, nullptr, 10);
    *L = (unsigned)strtoul(text.c_str() + l + 17, nullptr, 10);
    *q = (unsigned)strtoul(text.c_str() + qq + 14, nullptr, 10);
    return true;
}

static void cache_store(const char *path, const std::string &sig, unsigned R, unsigned L, unsigned q, double ms)
{
    if (!path)
        return;
    std::string text;
    if (FILE *f = fopen(path, "r")) {
        char chunk[4096];
        size_t n;
        while ((n = fread(chunk, 1, sizeof(chunk), f)) > 0)
            text.append(chunk, n);
        fclose(f);
    }
    size_t end = text.rfind('}');
    char entry[512];
    snprintf(entry, sizeof(entry),
             "  \"%s\": {\"slots_per_lane\": %u, \"lanes_per_slot\": %u, \"queue_block\": %u, "
             "\"ms_per_iteration\": %.6f}",
             sig.c_str(), R, L, q, ms);
    std::string out;
    if (end == std::string::npos || text.find('"') == std::string::npos)
        out = std::string("{\n") + entry + "\n}\n";
    else
        out = text.substr(0, end) + ",\n" + entry + "\n}\n";
    /* strip a blank line left between the previous entry and the comma */
    size_t bad;
    while ((bad = out.find("\n,\n")) != std::string::npos)
        out.replace(bad, 3, ",\n");
    if (FILE *f = fopen(path, "w")) {
        fputs(out.c_str(), f);
        fclose(f);
    }
}

extern "C" int bp_batch_tune(bp_batch *b, unsigned int iterations, const char *cache_path)
{
    bp_clear_error();
    if (!b || !b->uploaded) {
        bp_set_error("bp_batch_tune: batch not uploaded");
        return -1;
    }
    HIP_TRY(hipSetDevice(b->device));
    std::string sig = batch_signature(b);
    unsigned R = 0, L = 0, q = 0;
    if (cache_lookup(cache_path, sig, &R, &L, &q) && valid_config(b, R, L, q)) {
        b->R = R;
        b->L = L;
        b->qblock = q;
        return 0;
    }
    const unsigned Rs[2] = {1, 2};
    const unsigned Ls[4] = {1, 2, 4, 8};
    const unsigned Qs[4] = {0, 1, 4, 16};
    double best = 1e300;
    unsigned bestR = b->R, bestL = b->L, bestQ = b->qblock;
    unsigned n = iterations ? iterations : 4;
    for (unsigned R2 : Rs)
        for (unsigned L2 : Ls)
            for (unsigned q2 : Qs) {
                if (!valid_config(b, R2, L2, q2))
                    continue;
                b->R = R2;
                b->L = L2;
                b->qblock = q2;
                if (bp_batch_init(b))
                    return -1;
                float ms = 0.0f;
                if (bp_batch_profile(b, 1, &ms) || bp_batch_profile(b, n, &ms))
                    return -1;
                if (ms < best) {
                    best = ms;
                    bestR = R2;
                    bestL = L2;
                    bestQ = q2;
                }
            }
    b->R = bestR;
    b->L = bestL;
    b->qblock = bestQ;
    b->ready = false; /* the timing runs changed the state: init again */
    cache_store(cache_path, sig, bestR, bestL, bestQ, best);
    return 0;
}
