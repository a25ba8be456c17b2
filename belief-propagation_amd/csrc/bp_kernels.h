/*
 * bp_kernels.h -- gfx950 kernels of the loopy BP hot path (included by
 * bp_device.hip only).
 *
 * Device layout (built by bp_upload from a Graph_t):
 *   slot s = position in the destination-grouped edge order of the
 *            reference (dest_nodes_to_edges_edge_list, graph.c:285-313), so
 *            every node's in-edges are contiguous and in the order the
 *            reference multiplies them;
 *   in_off[V+1]      slot range of each node
 *   slot_src[E]      source node of the slot's edge
 *   J[E][S*S]        the edge's table, J[i][j], i = source state (mixed
 *                    state counts: x rows of mixed_row_stride(S) floats at
 *                    J + slot_tab[s];
 *                    batches: a pooled table at tabs + slot_tab[s]), j = destination state
 *   msg[2][E][SP]    messages, double-buffered (old / new)
 *   states[V][SP], buf[2][V][SP]  beliefs and the combine buffer
 *                    states (.) prod(in-messages), double-buffered
 * S = largest state count, SP = row stride; rows are zero beyond a node's
 * state count, which is what gives mixed-state networks their defined
 * zero-extension semantics.
 *
 * One loopy iteration (src/graph/graph.c:1121-1183) is ONE launch of
 * k_loopy<ITER>: the reference's send phase is restated as a pull over each
 * node's in-edges, which needs the source's combine buffer of this iteration
 * -- that buffer only depends on the source's own previous state and
 * messages, so the previous launch leaves it in buf[next].  Per wave:
 *   1. one lane per in-slot of a chunk of nodes (<= 64 R slots and nodes, R
 *      slots per lane):
 *      gather buf[src], m = N(J . buf[src]), write m, |m - old| into an FP64
 *      partial (the delta of graph.c:1300-1309);
 *   2. the messages are staged in LDS; one lane per node multiplies them in
 *      slot order: t = prod m (graph.c:985-1012), states = N(states * t), and
 *      buf[next] = states (.) prod m (graph.c:851-882 of the next iteration).
 * Every float operation is the reference's, in its order (sums from 0.0 left
 * to right, products in dest-CSR order, NaN factors skipped, sum <= 0 -> 1,
 * correctly rounded division); the library is built with
 * -ffp-contract=off so no FMA changes a rounding.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

/* Compile-time tuning (measured choices are the defaults). */
#ifndef BP_NT_LOADS
#define BP_NT_LOADS 0
#endif
#ifndef BP_NT_STORES
#define BP_NT_STORES 0
#endif
#ifndef BP_SLOTS_SMALL
#define BP_SLOTS_SMALL 2
#endif
#ifndef BP_XCD_REMAP
#define BP_XCD_REMAP 1
#endif
/* uniform-state handles may share one stored table per transposed pair
 * (bp_upload decides per graph); 0 keeps one table per slot */
#ifndef BP_SHARE_TABLES
#define BP_SHARE_TABLES 1
#endif
/* > 0: the iteration grid is capped at this many workgroups per CU and each
 * wave walks its share of chunks in a loop (0: one chunk per wave). */
#ifndef BP_PERSIST
#define BP_PERSIST 0
#endif

namespace bpk {

#ifndef BP_BLOCK
#define BP_BLOCK 256
#endif
/* Lanes per slot of the plain path (bp_upload handles) for four or more
 * states: 1, 2, 4 or 8 (one for fewer states). */
#ifndef BP_PLAIN_L
#define BP_PLAIN_L 1
#endif
/* Lanes-per-slot path: 1 writes each new message as one whole row from its
 * LDS copy, one lane per slot, after the passes; 0 lets every lane store
 * the entries it formed. */
#ifndef BP_ROWS_MSG_FROM_LDS
#define BP_ROWS_MSG_FROM_LDS 1
#endif
/* Lanes-per-slot path: 1 loads each node's belief after the passes instead
 * of before them (8 fewer registers held through the passes at S = 8);
 * always so for a kernel given a register budget below. */
#ifndef BP_ROWS_LATE_STATES
#define BP_ROWS_LATE_STATES 0
#endif
/* Minimum waves per SIMD requested of the compiler for k_loopy (1: its own
 * choice).  Measuring switches: BP_MIN_WAVES for every k_loopy,
 * BP_BATCH_MIN_WAVES for the batch kernels (PACKED) alone. */
#ifndef BP_MIN_WAVES
#define BP_MIN_WAVES 0
#endif
#ifndef BP_BATCH_MIN_WAVES
#define BP_BATCH_MIN_WAVES 0
#endif

constexpr int WAVE = 64;
constexpr int BLOCK = BP_BLOCK;
constexpr int WAVES_PER_BLOCK = BLOCK / WAVE;

/* ITER  one node-centric loopy iteration (graph.c:1121-1183)
 * INIT  init_previous_edge (graph.c:725-751) + the first combine buffer
 * PREP  the combine buffer from given states and messages
 * EDGE  one edge-centric iteration (graph.c:1186-1227, defects fixed):
 *       messages from the sources' beliefs, beliefs multiplied by every
 *       in-message without NaN skip, divided by their sum when it is > 0
 * FWD   the beliefs and combine buffer of ITER without storing messages or
 *       the delta, reading states and writing states_out (the unrolled
 *       forward of the adjoint, which keeps every iterate it needs)
 * FINIT the combine buffer of INIT without storing messages */
enum Mode { ITER = 0, INIT = 1, PREP = 2, EDGE = 3, FWD = 4, FINIT = 5 };

struct LoopyArgs {
    const unsigned *in_off;      /* [V+1] */
    const unsigned *slot_src;    /* [E] */
    const unsigned char *slot_xy;/* [E] x | y << 4 (mixed networks only) */
    const float *J;              /* [E][S*S] */
    const unsigned char *nvars;  /* [V] (mixed networks only) */
    float *states;               /* [V][SP] (EDGE, FWD: the beliefs before the iteration) */
    float *states_out;           /* [V][SP] EDGE, FWD: the beliefs after it */
    const float *buf_in;         /* [V][SP] gather source (ITER) */
    float *buf_out;              /* [V][SP] */
    const float *msg_in;         /* [E][SP] old messages (ITER) / messages (PREP) */
    float *msg_out;              /* [E][SP] new messages (ITER, INIT) */
    const uint2 *chunk;          /* [nchunks+1] (first node, first slot) of each chunk */
    const uint2 *chunk_base;     /* PACKED: [nchunks] row offsets (floats) of the chunk's first node and first slot */
    const unsigned *order;       /* non-null: [nchunks] the p-th chunk taken is chunk order[p] */
    unsigned nchunks;
    double *partials;            /* [G * waves per block] this launch's per-wave delta partials (ITER) */
    const unsigned *stop;        /* non-null: skip the launch once *stop != 0 */
    /* pooled tables (batched networks, PACKED): slot s's table starts at
     * tabs + slot_tab[s], row-major with the slot's x rows of y entries,
     * each row padded with zeros to a multiple of four floats (16-byte
     * aligned rows).  Otherwise, with mixed state counts, slot s's table is
     * J + slot_tab[s], x rows of y entries, each row padded to
     * mixed_row_stride(S) floats (uniform state counts: J[s][S*S], or with
     * a shared pool slot_tab[s] = pool entry | transposed << 31). */
    const unsigned *slot_tab;
    const float *tabs;
    double *chunk_partials;      /* non-null: one delta partial per chunk instead of per block */
    unsigned *queue;             /* non-null: chunks are taken from this counter in blocks of qblock */
    unsigned qblock;
};

/* A wave-uniform read of data the launch does not write (chunk records,
 * chunk order): through the constant address space, so the compiler reads
 * it with one scalar load into SGPRs instead of a vector load per field. */
__device__ __forceinline__ unsigned uniform_load(const unsigned *p)
{
    return *(const __attribute__((address_space(4))) unsigned *)p;
}

__device__ __forceinline__ uint2 uniform_load(const uint2 *p)
{
    unsigned long long v = *(const __attribute__((address_space(4))) unsigned long long *)p;
    return make_uint2((unsigned)v, (unsigned)(v >> 32));
}

template <int S> struct Stride {
    static constexpr int value = (S <= 2) ? S : (S <= 4) ? 4 : (S <= 8) ? 8 : 10;
};

/* Mixed state counts (plain handles): a slot's table is its x rows, each
 * padded to this many floats (S = the handle's largest state count), so
 * every row starts 16-byte aligned and is read as whole vector loads. */
__host__ __device__ constexpr unsigned mixed_row_stride(unsigned S) { return (S + 3u) & ~3u; }

/* The batch kernel for up to 8 states with 2 lanes per slot needs 90
 * registers, which leaves room for 5 waves per SIMD; held to 80 (6 waves,
 * no spills once the belief is loaded after the passes) its iteration on
 * mixed 2-8 state batches is 7% shorter.  Other shapes keep the compiler's
 * choice: a budget of 6 waves makes several of them spill. */
template <bool PACKED, int S, int L> struct MinWaves {
    static constexpr int value = (PACKED && BP_BATCH_MIN_WAVES > 0) ? BP_BATCH_MIN_WAVES
                                 : (BP_MIN_WAVES > 0)                ? BP_MIN_WAVES
                                 : (PACKED && S == 8 && L == 2)      ? 6
                                                                     : 1;
};

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));

/* Streaming loads and stores: data read or written once per launch (slot
 * arrays, tables, messages, a node's own belief) are marked non-temporal so
 * they do not displace the combine buffers, which every XCD gathers at
 * random, from L2. */
template <int N> __device__ __forceinline__ void load_row(const float *__restrict__ p, float *out);

template <int N>
__device__ __forceinline__ void load_row_nt(const float *__restrict__ p, float *out)
{
    if constexpr (!BP_NT_LOADS) {
        load_row<N>(p, out);
    } else if constexpr (N % 4 == 0) {
#pragma unroll
        for (int i = 0; i < N / 4; ++i) {
            v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(p) + i);
            out[4 * i] = v.x;
            out[4 * i + 1] = v.y;
            out[4 * i + 2] = v.z;
            out[4 * i + 3] = v.w;
        }
    } else if constexpr (N % 2 == 0) {
#pragma unroll
        for (int i = 0; i < N / 2; ++i) {
            v2f v = __builtin_nontemporal_load(reinterpret_cast<const v2f *>(p) + i);
            out[2 * i] = v.x;
            out[2 * i + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i)
            out[i] = __builtin_nontemporal_load(p + i);
    }
}

template <int N>
__device__ __forceinline__ void store_row_nt(float *__restrict__ p, const float *v)
{
    if constexpr (!BP_NT_STORES) {
#pragma unroll
        for (int i = 0; i < N; ++i)
            p[i] = v[i];
    } else if constexpr (N % 4 == 0) {
#pragma unroll
        for (int i = 0; i < N / 4; ++i)
            __builtin_nontemporal_store(v4f{v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]},
                                        reinterpret_cast<v4f *>(p) + i);
    } else if constexpr (N % 2 == 0) {
#pragma unroll
        for (int i = 0; i < N / 2; ++i)
            __builtin_nontemporal_store(v2f{v[2 * i], v[2 * i + 1]}, reinterpret_cast<v2f *>(p) + i);
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i)
            __builtin_nontemporal_store(v[i], p + i);
    }
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Logical block index that gives each XCD a contiguous range of chunks
 * (workgroups are dealt to the 8 XCDs round-robin). */
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned G)
{
    if (!BP_XCD_REMAP || (G & 7u) != 0u)
        return b;
    return (b & 7u) * (G >> 3) + (b >> 3);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, WAVE);
    return v;
}

/* m = N(J . b): m[i] = sum_j J[i][j] * b[j] for i < x, j < y (graph.c:896-909).
 * Entries past x are +0, the zero extension a node's products read past
 * the message's own entries, also when the sum is NaN (an infinite table
 * entry): 0 / NaN would be NaN, which the products skip instead of
 * multiplying by 0. */
template <int S, bool MIXED>
__device__ __forceinline__ void send(const float *__restrict__ Jrow, const float *b, unsigned x, unsigned y, float *m)
{
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i < S; ++i) {
        float acc = 0.0f;
        if (!MIXED || (unsigned)i < x) {
#pragma unroll
            for (int j = 0; j < S; ++j)
                if (!MIXED || (unsigned)j < y)
                    acc = acc + Jrow[i * S + j] * b[j];
            sum = sum + acc;
        }
        m[i] = acc;
    }
    if (sum <= 0.0f)
        sum = 1.0f;
#pragma unroll
    for (int i = 0; i < S; ++i)
        m[i] = (!MIXED || (unsigned)i < x) ? m[i] / sum : 0.0f;
}

/* send() for a table whose entries past y are +0 in every row i < x (the
 * zero padding of mixed-state rows read by load_mixed_rows): the source row
 * is zeroed from y on once, every product is then taken, and rows i >= x
 * (whatever lies there) are dropped after their sums.  Each extra product
 * is +0, and adding +0 leaves a sum unchanged (a sum that starts at +0 is
 * never -0), so the results are send()'s bit for bit with one select per
 * entry of the row instead of one per product. */
template <int S>
__device__ __forceinline__ void send_padded(const float *__restrict__ Jrow, const float *b, unsigned x, unsigned y,
                                            float *m)
{
    float bb[S];
#pragma unroll
    for (int j = 0; j < S; ++j)
        bb[j] = (unsigned)j < y ? b[j] : 0.0f;
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i < S; ++i) {
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < S; ++j)
            acc = acc + Jrow[i * S + j] * bb[j];
        m[i] = (unsigned)i < x ? acc : 0.0f;
        sum = sum + m[i];
    }
    if (sum <= 0.0f)
        sum = 1.0f;
#pragma unroll
    for (int i = 0; i < S; ++i)
        m[i] = (unsigned)i < x ? m[i] / sum : 0.0f;
}

/* Row i of a pooled table (rows of y entries padded to a multiple of four)
 * into r[0..S): aligned four-float loads, zero beyond y. */
template <int S>
__device__ __forceinline__ void load_pooled_row(const float *__restrict__ T, unsigned i, unsigned y, float *r)
{
    const float *row = T + (size_t)i * ((y + 3u) & ~3u);
#pragma unroll
    for (int q = 0; q < (S + 3) / 4; ++q) {
        v4f v = {0.0f, 0.0f, 0.0f, 0.0f};
        if ((unsigned)(4 * q) < y)
            v = *reinterpret_cast<const v4f *>(row + 4 * q);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (4 * q + c < S)
                r[4 * q + c] = v[c];
    }
}

/* A mixed-state table (plain handles) at T into J[S][S]: its rows at
 * stride mixed_row_stride(S), every row up to S read unconditionally as
 * aligned four-float vectors (the table array carries slack past its last
 * table).  No branch, no runtime index into registers.  Entries outside x
 * rows and y columns hold whatever lies there: send() reads only i < x,
 * j < y. */
template <int S>
__device__ __forceinline__ void load_mixed_rows(const float *__restrict__ T, float *J)
{
    constexpr int RS = (int)mixed_row_stride(S);
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
        for (int q = 0; q < RS / 4; ++q) {
            const v4f v = *reinterpret_cast<const v4f *>(T + i * RS + 4 * q);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (4 * q + c < S)
                    J[i * S + 4 * q + c] = v[c];
        }
}

/* A uniform-state table from the shared pool: entry slot_tab[s] of J,
 * read transposed when bit 31 is set (the slot is the later one of a pair
 * whose tables are exact transposes). */
template <int S>
__device__ __forceinline__ void load_shared_table(const LoopyArgs &a, unsigned s, float *J)
{
    unsigned t = a.slot_tab[s];
    const float *T = a.J + (size_t)(t & 0x7fffffffu) * S * S;
    const bool tr = (t >> 31) != 0;
    /* each entry from its own address (a select between two entries of a
     * loaded register array compiles to a dynamically indexed read) */
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
        for (int j = 0; j < S; ++j)
            J[i * S + j] = T[tr ? j * S + i : i * S + j];
}

/* A pooled uniform table as stored: S = 2 and 4 in aligned four-float
 * vectors (the pool's tables are S*S floats, 16-byte aligned for those). */
template <int S> struct PoolRaw {
    static constexpr int NV = (S == 2 || S == 4) ? S * S / 4 : 1;
    v4f v[NV];
};

template <int S> __device__ __forceinline__ void load_pool_raw(const float *T, PoolRaw<S> &p)
{
    if constexpr (S == 2 || S == 4) {
#pragma unroll
        for (int q = 0; q < PoolRaw<S>::NV; ++q)
            p.v[q] = reinterpret_cast<const v4f *>(T)[q];
    }
}

/* J = the stored table, or its transpose when tr: selects between
 * register values, after the loads have arrived */
template <int S> __device__ __forceinline__ void pool_table(const PoolRaw<S> &p, bool tr, float *J)
{
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const int f = i * S + j, g = j * S + i;
            float as = p.v[f / 4][f % 4], at = p.v[g / 4][g % 4];
            J[f] = tr ? at : as;
        }
}

/* The table of slot s as J[i*S + j] (zero beyond x rows and y columns). */
template <int S, bool MIXED, bool PACKED, bool POOLED>
__device__ __forceinline__ void load_table(const LoopyArgs &a, unsigned s, unsigned xy, float *J)
{
    if constexpr (!PACKED && !MIXED) {
        if constexpr (POOLED)
            load_shared_table<S>(a, s, J);
        else
            load_row_nt<S * S>(a.J + (size_t)s * S * S, J);
    } else if constexpr (!PACKED) {
        load_mixed_rows<S>(a.J + a.slot_tab[s], J);
    } else {
        unsigned x = xy & 15u, y = xy >> 4;
        const float *T = a.tabs + a.slot_tab[s];
#pragma unroll
        for (int i = 0; i < S; ++i) {
            float r[S];
            load_pooled_row<S>(T, i, (unsigned)i < x ? y : 0u, r);
#pragma unroll
            for (int j = 0; j < S; ++j)
                J[i * S + j] = r[j];
        }
    }
}

/* m = N(J . b) with the table read row by row from memory instead of held
 * in registers (large state counts); same operations in the same order as
 * send(). */
template <int S, bool MIXED, bool PACKED, bool POOLED>
__device__ __forceinline__ void send_from_memory(const LoopyArgs &a, unsigned s, unsigned xy, const float *b, float *m)
{
    unsigned x = MIXED ? (xy & 15u) : (unsigned)S, y = MIXED ? (xy >> 4) : (unsigned)S;
    if constexpr (POOLED) {
        float Js[S * S];
        load_shared_table<S>(a, s, Js);
        send<S, false>(Js, b, S, S, m);
        return;
    }
    const float *T = PACKED ? a.tabs + a.slot_tab[s] : MIXED ? a.J + a.slot_tab[s] : a.J + (size_t)s * S * S;
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i < S; ++i) {
        float acc = 0.0f;
        if (!MIXED || (unsigned)i < x) {
            float row[S];
            if constexpr (PACKED) {
                load_pooled_row<S>(T, i, y, row);
            } else if constexpr (MIXED) {
                constexpr int RS = (int)mixed_row_stride(S);
#pragma unroll
                for (int j = 0; j < S; ++j)
                    row[j] = (unsigned)j < y ? T[(size_t)i * RS + j] : 0.0f;
            } else {
#pragma unroll
                for (int j = 0; j < S; ++j)
                    row[j] = T[(size_t)i * S + j];
            }
#pragma unroll
            for (int j = 0; j < S; ++j)
                if (!MIXED || (unsigned)j < y)
                    acc = acc + row[j] * b[j];
            sum = sum + acc;
        }
        m[i] = acc;
    }
    if (sum <= 0.0f)
        sum = 1.0f;
#pragma unroll
    for (int i = 0; i < S; ++i)
        m[i] = (!MIXED || (unsigned)i < x) ? m[i] / sum : 0.0f;
}

/* ---- compact rows (PACKED: the batch layout) ------------------------------
 * A node's belief and combine-buffer rows hold its n entries rounded up to
 * an even count (zero-padded), and a slot's message rows its x entries the
 * same way, back to back in node (slot) order: rows start 8-byte aligned
 * and no row carries the padding to the largest state count.  A chunk's
 * first node and first slot rows are in chunk_base; the others follow by
 * wave-wide prefix sums of the widths.  slot_src holds the row offset of
 * the slot's source instead of its node number. */
__device__ __forceinline__ unsigned even_width(unsigned n) { return (n + 1u) & ~1u; }

/* exclusive prefix sum of v across the 64 lanes */
__device__ __forceinline__ unsigned wave_excl_scan(unsigned v, unsigned lane)
{
    unsigned x = v;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        unsigned y = __shfl_up(x, off, WAVE);
        if (lane >= (unsigned)off)
            x += y;
    }
    return x - v;
}

/* the sum of v across the 64 lanes, given its exclusive prefix sum */
__device__ __forceinline__ unsigned wave_total(unsigned excl, unsigned v) { return __shfl(excl + v, WAVE - 1, WAVE); }

/* The first S floats at p (a compact row and whatever follows it: the
 * buffers carry S floats of slack past their last row), 8-byte loads,
 * unpredicated.  out[k] for k >= the row's count is not the row's: callers
 * that read those entries use load_crow_z. */
template <int S>
__device__ __forceinline__ void load_crow(const float *__restrict__ p, float *out)
{
#pragma unroll
    for (int q = 0; q < (S + 1) / 2; ++q) {
        float2 v = *reinterpret_cast<const float2 *>(p + 2 * q);
        out[2 * q] = v.x;
        if (2 * q + 1 < S)
            out[2 * q + 1] = v.y;
    }
}

/* out[k] = row[k] for k < n, 0 beyond (a compact row of n entries read
 * zero-extended, as a gathered source row must be) */
template <int S>
__device__ __forceinline__ void load_crow_z(const float *__restrict__ p, unsigned n, float *out)
{
    load_crow<S>(p, out);
#pragma unroll
    for (int k = 0; k < S; ++k)
        out[k] = ((unsigned)k < n) ? out[k] : 0.0f;
}

/* the compact row of n entries v[0..n), zero-padded to an even count.
 * Each pair's second entry is selected before, and outside, the branch
 * that guards its store, and the empty asm keeps it there: with the select
 * inside the branch, the optimiser merges the two tests on n into a switch
 * (n = 0: no store, n = 1: second entry 0, else v[1]), and this compiler
 * (ROCm 7.2, gfx950) can lower that divergent switch so that the else
 * case stores 0 as well; measured on a one-wave test of the node phase. */
template <int S>
__device__ __forceinline__ void store_crow(float *__restrict__ p, unsigned n, const float *v)
{
#pragma unroll
    for (int q = 0; q < (S + 1) / 2; ++q) {
        float hi = (2 * q + 1 < S && (unsigned)(2 * q + 1) < n) ? v[2 * q + 1] : 0.0f;
        asm volatile("" : "+v"(hi));
        if ((unsigned)(2 * q) < n)
            *reinterpret_cast<float2 *>(p + 2 * q) = make_float2(v[2 * q], hi);
    }
}

/* The message of slot s into m[0..S) (zero beyond x), without side effects.
 * mrow: the slot's message row offset (PACKED). */
template <int S, bool MIXED, int MODE, bool PACKED, bool POOLED>
__device__ __forceinline__ void compute_message(const LoopyArgs &a, unsigned s, float *m, unsigned &x,
                                                unsigned mrow = 0)
{
    constexpr int SP = Stride<S>::value;
    x = S;
    unsigned y = S;
    if (MIXED) {
        unsigned xy = a.slot_xy[s];
        x = xy & 15u;
        y = xy >> 4;
    }
    if (MODE == PREP) {
        if constexpr (PACKED) {
            load_crow_z<S>(a.msg_in + mrow, x, m);
        } else {
#pragma unroll
            for (int i = 0; i < S; ++i)
                m[i] = a.msg_in[(size_t)s * SP + i];
        }
        return;
    }
    unsigned u = a.slot_src[s];
    const float *src = (MODE == INIT || MODE == EDGE || MODE == FINIT) ? a.states : a.buf_in;
    float b[S];
    if constexpr (PACKED) {
#pragma unroll
        for (int j = 0; j < S; ++j)
            b[j] = ((unsigned)j < x) ? src[u + j] : 0.0f; /* the source's row: x = its state count */
    } else {
#pragma unroll
        for (int j = 0; j < S; ++j)
            b[j] = src[(size_t)u * SP + j];
    }
    send_from_memory<S, MIXED, PACKED, POOLED>(a, s, x | (y << 4), b, m);
}

/* Slot phase: the message of slot s, written to msg_out (ITER, INIT, EDGE)
 * with its |new - old| added to dacc (ITER, EDGE). */
template <int S, bool MIXED, int MODE, bool PACKED, bool POOLED>
__device__ __forceinline__ void slot_message(const LoopyArgs &a, unsigned s, float *m, double &dacc,
                                             unsigned mrow = 0)
{
    constexpr int SP = Stride<S>::value;
    unsigned x;
    compute_message<S, MIXED, MODE, PACKED, POOLED>(a, s, m, x, mrow);
    if (MODE == PREP || MODE == FWD || MODE == FINIT)
        return;
    if (MODE == ITER || MODE == EDGE) {
        const float *old = PACKED ? a.msg_in + mrow : a.msg_in + (size_t)s * SP;
#pragma unroll
        for (int i = 0; i < S; ++i) {
            if (!MIXED || (unsigned)i < x) {
                float q = old[i] - m[i];
                if (q != q)
                    q = 0.0f;
                dacc += (double)fabsf(q);
            }
        }
    }
    if constexpr (PACKED) {
        float *dst = a.msg_out + mrow;
#pragma unroll
        for (int i = 0; i < S; ++i)
            if ((unsigned)i < ((x + 1u) & ~1u))
                dst[i] = ((unsigned)i < x) ? m[i] : 0.0f;
    } else {
        float *dst = a.msg_out + (size_t)s * SP;
#pragma unroll
        for (int i = 0; i < SP; ++i)
            dst[i] = (i < S) ? m[i] : 0.0f;
    }
}

/* buf[k] = buf[k] * m[k] unless m[k] is NaN (graph.c:450-459) */
template <int S, bool MIXED>
__device__ __forceinline__ void combine(float *buf, const float *m, unsigned n)
{
#pragma unroll
    for (int k = 0; k < S; ++k) {
        if (!MIXED || (unsigned)k < n) {
            float mk = m[k];
            if (mk == mk)
                buf[k] = buf[k] * mk;
        }
    }
}

/* the product of the edge-centric schedule: every factor, NaN included
 * (graph.c:1034-1043) */
template <int S, bool MIXED>
__device__ __forceinline__ void combine_all(float *buf, const float *m, unsigned n)
{
#pragma unroll
    for (int k = 0; k < S; ++k)
        if (!MIXED || (unsigned)k < n)
            buf[k] = buf[k] * m[k];
}

template <int S, bool MIXED, int MODE>
__device__ __forceinline__ void combine_mode(float *buf, const float *m, unsigned n)
{
    if (MODE == EDGE)
        combine_all<S, MIXED>(buf, m, n);
    else
        combine<S, MIXED>(buf, m, n);
}

/* divided by the sum only when it is positive (graph.c:1046-1061) */
template <int S, bool MIXED>
__device__ __forceinline__ void normalise_positive(float *v, unsigned n)
{
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < S; ++k)
        if (!MIXED || (unsigned)k < n)
            sum = sum + v[k];
    if (sum > 0.0f) {
#pragma unroll
        for (int k = 0; k < S; ++k)
            if (!MIXED || (unsigned)k < n)
                v[k] = v[k] / sum;
    }
}

template <int S, bool MIXED>
__device__ __forceinline__ void normalise(float *v, unsigned n)
{
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < S; ++k)
        if (!MIXED || (unsigned)k < n)
            sum = sum + v[k];
    if (sum <= 0.0f)
        sum = 1.0f;
#pragma unroll
    for (int k = 0; k < S; ++k)
        if (!MIXED || (unsigned)k < n)
            v[k] = v[k] / sum;
}

/* row: the node's row offset in floats (v * SP, or its compact row when
 * PACKED) */
template <int S, bool MIXED, int MODE, bool PACKED = false>
__device__ __forceinline__ void write_node(const LoopyArgs &a, size_t row, const float *st, const float *b, unsigned n)
{
    constexpr int SP = Stride<S>::value;
    if constexpr (PACKED) {
        if (MODE == EDGE) {
            store_crow<S>(a.states_out + row, n, b);
            return;
        }
        if (MODE == FWD)
            store_crow<S>(a.states_out + row, n, st);
        if (MODE == ITER)
            store_crow<S>(a.states + row, n, st);
        store_crow<S>(a.buf_out + row, n, b);
        return;
    }
    if (MODE == EDGE) { /* the new beliefs are the product */
#pragma unroll
        for (int k = 0; k < SP; ++k)
            a.states_out[row + k] = (k < S && (!MIXED || (unsigned)k < n)) ? b[k] : 0.0f;
        return;
    }
    if (MODE == FWD) {
#pragma unroll
        for (int k = 0; k < SP; ++k)
            a.states_out[row + k] = (k < S && (!MIXED || (unsigned)k < n)) ? st[k] : 0.0f;
    }
    /* whole padded rows (zeros beyond the node's state count): a partly
     * written row costs more than the zeros (measured on mixed networks) */
    if (MODE == ITER) {
        float r[SP];
#pragma unroll
        for (int k = 0; k < SP; ++k)
            r[k] = (k < S && (!MIXED || (unsigned)k < n)) ? st[k] : 0.0f;
        store_row_nt<SP>(a.states + row, r);
    }
    float r[SP];
#pragma unroll
    for (int k = 0; k < SP; ++k)
        r[k] = (k < S && (!MIXED || (unsigned)k < n)) ? b[k] : 0.0f;
    store_row_nt<SP>(a.buf_out + row, r);
}

/* Row loads: 8- and 16-byte vector loads where the row size allows. */
template <int N>
__device__ __forceinline__ void load_row(const float *__restrict__ p, float *out)
{
    if constexpr (N % 4 == 0) {
#pragma unroll
        for (int i = 0; i < N / 4; ++i) {
            float4 v = reinterpret_cast<const float4 *>(p)[i];
            out[4 * i] = v.x;
            out[4 * i + 1] = v.y;
            out[4 * i + 2] = v.z;
            out[4 * i + 3] = v.w;
        }
    } else if constexpr (N % 2 == 0) {
#pragma unroll
        for (int i = 0; i < N / 2; ++i) {
            float2 v = reinterpret_cast<const float2 *>(p)[i];
            out[2 * i] = v.x;
            out[2 * i + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i)
            out[i] = p[i];
    }
}

/* A message's NaN entries replaced by 1, once: multiplying by every entry
 * of the result is then the NaN-skipping product of combine() (x * 1 == x
 * for every x). */
template <int S>
__device__ __forceinline__ void nan_to_one(float *m)
{
#pragma unroll
    for (int k = 0; k < S; ++k)
        m[k] = (m[k] == m[k]) ? m[k] : 1.0f;
}

template <int S>
__device__ __forceinline__ void mul_all(float *buf, const float *m)
{
#pragma unroll
    for (int k = 0; k < S; ++k)
        buf[k] = buf[k] * m[k];
}

/* Node phase for node v with in-slots [q0, q1) of the chunk's LDS staging
 * (messages at lds[k * W + q]): the belief update of graph.c:985-1024
 * (ITER, FWD), the combine buffer of graph.c:851-882 for the next
 * iteration, or the edge schedule's product (EDGE).  The products run over
 * all S entries: with mixed state counts, entries past the node's count n
 * hold whatever the products give and are never used (the sums below and
 * write_node take k < n only), which spares a mask per entry. */
template <int S, bool MIXED, int MODE, bool PACKED = false>
__device__ __forceinline__ void node_update(const LoopyArgs &a, size_t row, unsigned q0, unsigned q1, float *st,
                                            unsigned n, const float *lds, int W)
{
    /* the first K in-messages are read from LDS once, all together, and
     * kept for both products (most nodes have no more); the rest are read
     * in slot order as they are combined.  Except for the edge schedule
     * (whose product takes NaN factors too), each message's NaN entries
     * are replaced by 1 once, and both products multiply by every entry. */
    constexpr int K = (S <= 4) ? 4 : 2;
    constexpr bool CLEAN = MODE != EDGE;
    const unsigned cnt = q1 - q0;
    float mk[K][S], m[S], b[S];
#pragma unroll
    for (int j = 0; j < K; ++j)
        if ((unsigned)j < cnt) {
#pragma unroll
            for (int k = 0; k < S; ++k)
                mk[j][k] = lds[k * W + q0 + j];
            if (CLEAN)
                nan_to_one<S>(mk[j]);
        }
    if (MODE == ITER || MODE == FWD) {
        float t[S];
#pragma unroll
        for (int k = 0; k < S; ++k)
            t[k] = 1.0f;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if ((unsigned)j < cnt)
                mul_all<S>(t, mk[j]);
        for (unsigned q = q0 + K; q < q1; ++q) {
#pragma unroll
            for (int k = 0; k < S; ++k)
                m[k] = lds[k * W + q];
            nan_to_one<S>(m);
            mul_all<S>(t, m);
        }
        if (q1 > q0) {
#pragma unroll
            for (int k = 0; k < S; ++k)
                st[k] = st[k] * t[k];
        }
        normalise<S, MIXED>(st, n);
    }
#pragma unroll
    for (int k = 0; k < S; ++k)
        b[k] = st[k];
#pragma unroll
    for (int j = 0; j < K; ++j)
        if ((unsigned)j < cnt) {
            if (CLEAN)
                mul_all<S>(b, mk[j]);
            else
                combine_all<S, false>(b, mk[j], n);
        }
    for (unsigned q = q0 + K; q < q1; ++q) {
#pragma unroll
        for (int k = 0; k < S; ++k)
            m[k] = lds[k * W + q];
        if (CLEAN) {
            nan_to_one<S>(m);
            mul_all<S>(b, m);
        } else {
            combine_all<S, false>(b, m, n);
        }
    }
    if (MODE == EDGE)
        normalise_positive<S, MIXED>(b, n);
    write_node<S, MIXED, MODE, PACKED>(a, row, st, b, n);
}

/* Swap of a value between neighbouring lanes (2m <-> 2m + 1), by DPP. */
__device__ __forceinline__ float pair_swap(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}

/* The node phase of node_update (ITER, INIT; compact rows) with two lanes
 * per node, for a chunk of at most 32 nodes: lanes 2m and 2m + 1 hold
 * node m's entries [0, H) and [H, 2H), H = S / 2, so each lane does half
 * the products and divisions.  The node's fields come from lane m (which
 * loaded them, one lane per node).  The normalising sum keeps the order of
 * the one-lane form: the low lane adds its entries to 0, the high lane
 * continues from that partial, and the total goes back to the low lane.
 * As in node_update, each message's NaN entries are replaced by 1 once for
 * both products.  Entries past the node's count n are never summed or
 * stored. */
template <int S, int MODE>
__device__ __forceinline__ void node_update_pair(const LoopyArgs &a, unsigned nrow, unsigned q0, unsigned q1, unsigned n,
                                                 bool ok, const float *lds, unsigned lane)
{
    constexpr int H = S / 2, K = 4;
    const int src = (int)(lane >> 1);
    const unsigned h = lane & 1u, k0 = h * H;
    const unsigned row = (unsigned)__shfl((int)nrow, src, WAVE) + k0;
    const unsigned a0 = (unsigned)__shfl((int)q0, src, WAVE), a1 = (unsigned)__shfl((int)q1, src, WAVE);
    const unsigned nn = (unsigned)__shfl((int)n, src, WAVE);
    if (!ok)
        return;
    const unsigned cnt = a1 - a0;
    float mk[K][H], m[H], st[H], b[H];
#pragma unroll
    for (int j = 0; j < K; ++j)
        if ((unsigned)j < cnt)
#pragma unroll
            for (int k = 0; k < H; ++k)
                mk[j][k] = lds[(k0 + k) * WAVE + a0 + j];
#pragma unroll
    for (int j = 0; j < K; ++j)
        if ((unsigned)j < cnt)
            nan_to_one<H>(mk[j]);
    load_crow<H>(a.states + row, st);
    if (MODE == ITER) {
        float t[H];
#pragma unroll
        for (int k = 0; k < H; ++k)
            t[k] = 1.0f;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if ((unsigned)j < cnt)
                mul_all<H>(t, mk[j]);
        for (unsigned q = a0 + K; q < a1; ++q) {
#pragma unroll
            for (int k = 0; k < H; ++k)
                m[k] = lds[(k0 + k) * WAVE + q];
            nan_to_one<H>(m);
            mul_all<H>(t, m);
        }
        if (a1 > a0) {
#pragma unroll
            for (int k = 0; k < H; ++k)
                st[k] = st[k] * t[k];
        }
        float p = 0.0f;
#pragma unroll
        for (int k = 0; k < H; ++k)
            if (h == 0 && (unsigned)k < nn)
                p = p + st[k];
        const float lo = pair_swap(p);
        float sum = h ? lo : p;
#pragma unroll
        for (int k = 0; k < H; ++k)
            if (h == 1 && (unsigned)(H + k) < nn)
                sum = sum + st[k];
        const float hi = pair_swap(sum);
        sum = h ? sum : hi;
        if (sum <= 0.0f)
            sum = 1.0f;
#pragma unroll
        for (int k = 0; k < H; ++k)
            if (k0 + k < nn)
                st[k] = st[k] / sum;
    }
#pragma unroll
    for (int k = 0; k < H; ++k)
        b[k] = st[k];
#pragma unroll
    for (int j = 0; j < K; ++j)
        if ((unsigned)j < cnt)
            mul_all<H>(b, mk[j]);
    for (unsigned q = a0 + K; q < a1; ++q) {
#pragma unroll
        for (int k = 0; k < H; ++k)
            m[k] = lds[(k0 + k) * WAVE + q];
        nan_to_one<H>(m);
        mul_all<H>(b, m);
    }
    const unsigned own = nn > k0 ? nn - k0 : 0u;
    if (MODE == ITER)
        store_crow<H>(a.states + row, own, st);
    store_crow<H>(a.buf_out + row, own, b);
}

/* The first S floats of a row of SP: one 16-byte load when a 3-state row
 * sits in a 4-float slot, else load_row. */
template <int S, int SP>
__device__ __forceinline__ void load_rowp(const float *__restrict__ p, float *out)
{
    if constexpr (S == 3 && SP == 4) {
        float4 v = *reinterpret_cast<const float4 *>(p);
        out[0] = v.x;
        out[1] = v.y;
        out[2] = v.z;
    } else {
        load_row<S>(p, out);
    }
}

/* a streamed row (non-temporal when BP_NT_LOADS is set) */
template <int S, int SP>
__device__ __forceinline__ void load_rowp_nt(const float *__restrict__ p, float *out)
{
    if constexpr (BP_NT_LOADS)
        load_row_nt<S>(p, out);
    else
        load_rowp<S, SP>(p, out);
}

/* A chunk of at most 64*R nodes with at most 64*R in-slots: lane l takes
 * slots sb+l+64r and nodes nb+l+64r (r < R).  Every load that does not
 * depend on another one is issued before any is consumed, so a wave keeps
 * R slots' and R nodes' worth of reads in flight across one round trip. */
template <int S, bool MIXED, int MODE, int R, bool PACKED, bool POOLED>
__device__ __forceinline__ void chunk_narrow(const LoopyArgs &a, unsigned nb, unsigned ne, unsigned sb, unsigned se,
                                             float *lds, unsigned lane, double &dacc, uint2 base)
{
    constexpr int SP = Stride<S>::value;
    constexpr int W = WAVE * R;
    /* small tables are loaded up front with the other streams; large ones
     * are read row by row when the message is formed (register budget) */
    constexpr bool HOLD = S * S * R <= 32;
    /* uniform state counts with a shared table pool (S = 2, 4): the slot's
     * pool entry comes with the first loads and its table with the gathers,
     * transposed in registers when the message is formed */
    constexpr bool POOL = POOLED && HOLD && (S == 2 || S == 4);
    /* mixed state counts: likewise the slot's table offset first, its rows
     * (whole aligned vectors, no branch) with the gathers */
    constexpr bool OFFT = MIXED && !PACKED && HOLD;
    unsigned u[R], xy[R], q0[R], q1[R], nn[R], tix[(POOL || OFFT) ? R : 1];
    float J[HOLD ? R : 1][HOLD ? S * S : 1], old[R][S], st[R][S];
    PoolRaw<S> praw[POOL ? R : 1];
    /* Lanes past the chunk's end read a valid neighbour instead (and drop
     * the result), so the loads carry no branches and all of them -- then
     * all the gathers -- are in flight before the first wait. */
    /* (a chunk of nodes without in-edges has no slot to stand in: its slot
     * loads are skipped, uniformly for the wave) */
    const bool any_slots = se > sb;
    const unsigned s_last = se - 1, v_last = ne - 1;
    /* the nodes' loads first, then the slots' in one block: the gathers'
     * addresses are formed from the slots' loads, and the wait for them then
     * comes after every first-round load has been issued */
#pragma unroll
    for (int r = 0; r < R; ++r) {
        unsigned v = nb + r * WAVE + lane;
        if (v >= ne)
            v = v_last;
        q0[r] = (BP_NT_LOADS ? __builtin_nontemporal_load(a.in_off + v) : a.in_off[v]) - sb;
        q1[r] = (BP_NT_LOADS ? __builtin_nontemporal_load(a.in_off + v + 1) : a.in_off[v + 1]) - sb;
        /* INIT and EDGE gather beliefs: keep them cached there */
        if constexpr (!PACKED) {
            if (MODE == ITER || MODE == PREP || MODE == FWD)
                load_rowp_nt<S, SP>(a.states + (size_t)v * SP, st[r]);
            else
                load_rowp<S, SP>(a.states + (size_t)v * SP, st[r]);
        }
        nn[r] = MIXED ? a.nvars[v] : S;
        xy[r] = (unsigned)(S | (S << 4));
    }
    if (any_slots) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            unsigned s = sb + r * WAVE + lane;
            if (s >= se)
                s = s_last;
            if (MODE != PREP) {
                u[r] = BP_NT_LOADS ? __builtin_nontemporal_load(a.slot_src + s) : a.slot_src[s];
                if constexpr (MIXED)
                    xy[r] = a.slot_xy[s];
                if constexpr (POOL || OFFT)
                    tix[r] = a.slot_tab[s];
                else if constexpr (HOLD)
                    load_table<S, MIXED, PACKED, POOLED>(a, s, xy[r], J[r]);
            } else if constexpr (PACKED) {
                xy[r] = a.slot_xy[s];
            }
            if (!PACKED && (MODE == ITER || MODE == EDGE || MODE == PREP))
                load_rowp_nt<S, SP>(a.msg_in + (size_t)s * SP, old[r]);
        }
    }
    /* compact rows: each lane's message and node row offsets */
    unsigned mrow[R], nrow[R];
    if constexpr (PACKED) {
        unsigned mb = base.y, vb = base.x;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            unsigned s = sb + r * WAVE + lane, v = nb + r * WAVE + lane;
            unsigned xs = (any_slots && s < se) ? (xy[r] & 15u) : 0u, nv = (v < ne) ? nn[r] : 0u;
            unsigned w = even_width(xs), ex = wave_excl_scan(w, lane);
            mrow[r] = mb + ex;
            mb += wave_total(ex, w);
            unsigned wn = even_width(nv), exn = wave_excl_scan(wn, lane);
            nrow[r] = vb + exn;
            vb += wave_total(exn, wn);
            if (MODE == ITER || MODE == EDGE || MODE == PREP)
                load_crow<S>(a.msg_in + mrow[r], old[r]);
            load_crow<S>(a.states + nrow[r], st[r]);
        }
    }
    float bsrc[R][S];
    if (MODE != PREP && any_slots) {
        const float *src = (MODE == INIT || MODE == EDGE || MODE == FINIT) ? a.states : a.buf_in;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (PACKED)
                load_crow_z<S>(src + u[r], xy[r] & 15u, bsrc[r]); /* the source's row: x entries */
            else
                load_rowp<S, SP>(src + (size_t)u[r] * SP, bsrc[r]);
        }
        if constexpr (POOL) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                load_pool_raw<S>(a.J + (size_t)(tix[r] & 0x7fffffffu) * S * S, praw[r]);
        }
        if constexpr (OFFT) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                load_mixed_rows<S>(a.J + tix[r], J[r]);
        }
    }
    /* slot phase: gather, send, write, delta, stage in LDS */
#pragma unroll
    for (int r = 0; r < R; ++r) {
        unsigned s = sb + r * WAVE + lane;
        if (s < se) {
            float m[S];
            if (MODE == PREP) {
#pragma unroll
                for (int k = 0; k < S; ++k)
                    m[k] = old[r][k];
            } else {
                unsigned x = xy[r] & 15u, y = xy[r] >> 4;
                if constexpr (POOL)
                    pool_table<S>(praw[r], (tix[r] >> 31) != 0, J[r]);
                if constexpr (OFFT)
                    send_padded<S>(J[r], bsrc[r], x, y, m);
                else if constexpr (HOLD)
                    send<S, MIXED>(J[r], bsrc[r], x, y, m);
                else
                    send_from_memory<S, MIXED, PACKED, POOLED>(a, s, xy[r], bsrc[r], m);
                if (MODE == ITER || MODE == EDGE) {
#pragma unroll
                    for (int i = 0; i < S; ++i) {
                        if (!MIXED || (unsigned)i < x) {
                            float q = old[r][i] - m[i];
                            if (q != q)
                                q = 0.0f;
                            dacc += (double)fabsf(q);
                        }
                    }
                }
                if (MODE == ITER || MODE == INIT || MODE == EDGE) {
                    if constexpr (PACKED) {
                        store_crow<S>(a.msg_out + mrow[r], x, m);
                    } else {
                        float row[SP];
#pragma unroll
                        for (int i = 0; i < SP; ++i)
                            row[i] = (i < S) ? m[i] : 0.0f;
                        store_row_nt<SP>(a.msg_out + (size_t)s * SP, row);
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < S; ++k)
                lds[k * W + r * WAVE + lane] = m[k];
        }
    }
    wave_sync();
    /* node phase: one lane per node multiplies its in-messages in slot order */
#pragma unroll
    for (int r = 0; r < R; ++r) {
        unsigned v = nb + r * WAVE + lane;
        if (v < ne)
            node_update<S, MIXED, MODE, PACKED>(a, PACKED ? (size_t)nrow[r] : (size_t)v * SP, q0[r], q1[r], st[r],
                                                nn[r], lds, W);
    }
}

/* A chunk of at most 64 slots and 64 nodes with L lanes per slot: lane
 * (g, k) = (lane / L, lane % L) forms rows i = k, k + L, ... of the message
 * of slot g of each pass (L passes cover the chunk).  The normaliser is the
 * sum of the rows in row order, gathered by shuffles, so every operation is
 * the one send() performs, in its order.  Used for large state counts, where
 * one lane per slot would read a whole table alone. */
template <int S, bool MIXED, int MODE, int L, bool PACKED, bool POOLED>
__device__ __forceinline__ void chunk_rows(const LoopyArgs &a, unsigned nb, unsigned ne, unsigned sb, unsigned se,
                                           float *lds, float *stage_f, unsigned *stage_u, unsigned lane, double &dacc,
                                           uint2 cbase)
{
    constexpr int SP = Stride<S>::value;
    constexpr int W = WAVE;
    constexpr int G = WAVE / L;
    constexpr int NR = (SP + L - 1) / L; /* rows per lane, pad rows included */
    const unsigned k = lane % L, g = lane / L, base = lane - k;
    unsigned v = nb + lane, q0 = 0, q1 = 0, n = S;
    float st[S];
    /* stage every slot's operands once, one lane per slot (coalesced, and
     * the gathers issued together): the source's row b, the old message,
     * the table reference and the state counts.  The loads that depend on
     * nothing come first; the compact rows' offsets (PACKED) are prefix sums
     * of the state counts they bring, and the row loads follow. */
    const unsigned s = sb + lane;
    const bool slot_ok = s < se;
    unsigned xy = (unsigned)(S | (S << 4)), t = 0u, u = 0u;
    if (slot_ok) {
        if (MIXED)
            xy = a.slot_xy[s];
        if (PACKED || MIXED || POOLED)
            t = a.slot_tab[s];
        if (MODE != PREP)
            u = a.slot_src[s];
    }
    if (v < ne) {
        q0 = a.in_off[v] - sb;
        q1 = a.in_off[v + 1] - sb;
        n = MIXED ? a.nvars[v] : S;
    }
    unsigned mrow = 0, nrow = 0;
    if constexpr (PACKED) {
        mrow = cbase.y + wave_excl_scan(slot_ok ? even_width(xy & 15u) : 0u, lane);
        nrow = cbase.x + wave_excl_scan((v < ne) ? even_width(n) : 0u, lane);
    }
    if (slot_ok) {
        float b[S], old[S];
        if (MODE == PREP) {
            if constexpr (PACKED)
                load_crow_z<S>(a.msg_in + mrow, xy & 15u, b);
            else
                load_rowp<S, SP>(a.msg_in + (size_t)s * SP, b);
        } else {
            const float *src = (MODE == INIT || MODE == EDGE || MODE == FINIT) ? a.states : a.buf_in;
            if constexpr (PACKED)
                /* the source's row: x entries, zero beyond (and, as the
                 * products read only j < y, zero from y on as well, so
                 * run_pass needs no mask of its own) */
                load_crow_z<S>(src + u, min(xy & 15u, xy >> 4), b);
            else
                load_rowp<S, SP>(src + (size_t)u * SP, b);
        }
        if (MODE == ITER || MODE == EDGE) {
            if constexpr (PACKED)
                load_crow<S>(a.msg_in + mrow, old);
            else
                load_rowp<S, SP>(a.msg_in + (size_t)s * SP, old);
        }
#pragma unroll
        for (int j = 0; j < S; ++j) {
            stage_f[j * W + lane] = b[j];
            if (MODE == ITER || MODE == EDGE)
                stage_f[(S + j) * W + lane] = old[j];
        }
        stage_u[lane] = xy;
        stage_u[W + lane] = t;
    }
    auto load_states = [&]() {
        if (v < ne) {
            if constexpr (PACKED) {
                load_crow<S>(a.states + nrow, st);
            } else {
                if (MODE == ITER || MODE == PREP || MODE == FWD)
                    load_rowp_nt<S, SP>(a.states + (size_t)v * SP, st);
                else
                    load_rowp<S, SP>(a.states + (size_t)v * SP, st);
            }
        }
    };
    constexpr bool late_states = BP_ROWS_LATE_STATES || MinWaves<PACKED, S, L>::value > 1;
    if (!late_states)
        load_states();
    wave_sync();
    /* The table entries of pass p that this lane multiplies: rows
     * i = k, k + L, ... and columns j < y (zero elsewhere).  They are all
     * loaded before the first product, so the pass waits on memory once
     * rather than once per row (1.5x on mixed 2-8 state batches; loading
     * the next pass's entries during this pass as well, or the first pass's
     * together with the staged rows, was no faster, the extra registers
     * costing occupancy). */
    auto load_pass = [&](int p, float (&tv)[NR][S]) {
        unsigned q = p * G + g, s = sb + q;
        if constexpr (PACKED && MODE != PREP) {
            /* Every row this lane multiplies, read unconditionally as
             * aligned four-float vectors: row i of the pooled table starts
             * at i * st, st = y rounded up to four, and the pool ends with
             * slack for the rows and vectors past a table's last one.
             * Columns y..st-1 are the pool's zero padding; a vector wholly
             * past st belongs to the next row and is replaced by zeros, so
             * every entry j >= y is +0.  Rows i >= x hold whatever lies
             * there and are never used (run_pass drops them).  A slot past
             * the chunk reads from the pool's start. */
            const bool act = s < se;
            const unsigned xy = stage_u[q], tq = stage_u[W + q]; /* stale past the chunk */
            const unsigned y = act ? (xy >> 4) : 1u;
            const unsigned st = (y + 3u) & ~3u;
            const float *T = a.tabs + (act ? tq : 0u) + k * st;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const float *row = T + (size_t)(r * L) * st;
#pragma unroll
                for (int qv = 0; qv < (S + 3) / 4; ++qv) {
                    v4f w = *reinterpret_cast<const v4f *>(row + 4 * qv);
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        if (4 * qv + c < S)
                            tv[r][4 * qv + c] = (qv == 0 || (unsigned)(4 * qv) < y) ? w[c] : 0.0f;
                }
            }
            return;
        }
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int j = 0; j < S; ++j)
                tv[r][j] = 0.0f;
        if (MODE == PREP || s >= se)
            return;
        unsigned x = S, y = S;
        if (MIXED) {
            unsigned xy = stage_u[q];
            x = xy & 15u;
            y = xy >> 4;
        }
        const float *T = PACKED ? a.tabs + stage_u[W + q]
                         : MIXED ? a.J + stage_u[W + q]
                         : POOLED ? a.J + (size_t)(stage_u[W + q] & 0x7fffffffu) * S * S
                                  : a.J + (size_t)s * S * S;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            unsigned i = k + r * L;
            if (i < (unsigned)S && (!MIXED || i < x)) {
                if constexpr (PACKED) {
                    load_pooled_row<S>(T, i, y, tv[r]);
                } else if constexpr (POOLED) {
                    /* row i of the slot's table: row i of the stored one,
                     * or its column i when the slot reads it transposed */
                    const bool tr = (stage_u[W + q] >> 31) != 0;
#pragma unroll
                    for (int j = 0; j < S; ++j)
                        tv[r][j] = T[tr ? j * S + i : i * S + j];
                } else if constexpr (MIXED) {
                    constexpr int RS = (int)mixed_row_stride(S);
#pragma unroll
                    for (int j = 0; j < S; ++j)
                        tv[r][j] = (unsigned)j < y ? T[(size_t)i * RS + j] : 0.0f;
                } else {
#pragma unroll
                    for (int j = 0; j < S; ++j)
                        tv[r][j] = T[(size_t)i * S + j];
                }
            }
        }
    };
    auto run_pass = [&](int p, const float (&tv)[NR][S]) {
        unsigned q = p * G + g, s = sb + q;
        bool act = s < se;
        unsigned x = S, y = S;
        float acc[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r)
            acc[r] = 0.0f;
        if (act) {
            if (MIXED) {
                unsigned xy = stage_u[q];
                x = xy & 15u;
                y = xy >> 4;
            }
            if (MODE == PREP) {
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    unsigned i = k + r * L;
                    if (i < (unsigned)S)
                        acc[r] = stage_f[i * W + q];
                }
            } else if constexpr (PACKED) {
                /* The staged row holds +0 from y on (staging), and so does
                 * every table entry j >= y (load_pass): each such product
                 * is +0 and adding it leaves the sum unchanged (a sum that
                 * starts at +0 is never -0), so the unmasked sums equal the
                 * sums over j < y bit for bit. */
                float b[S];
#pragma unroll
                for (int j = 0; j < S; ++j)
                    b[j] = stage_f[j * W + q];
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    unsigned i = k + r * L;
                    float sacc = 0.0f;
#pragma unroll
                    for (int j = 0; j < S; ++j)
                        sacc = sacc + tv[r][j] * b[j];
                    acc[r] = i < x ? sacc : 0.0f;
                }
            } else {
                float b[S];
#pragma unroll
                for (int j = 0; j < S; ++j)
                    b[j] = stage_f[j * W + q];
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    unsigned i = k + r * L;
                    float sacc = 0.0f;
                    if (i < (unsigned)S && (!MIXED || i < x)) {
#pragma unroll
                        for (int j = 0; j < S; ++j)
                            if (!MIXED || (unsigned)j < y)
                                sacc = sacc + tv[r][j] * b[j];
                    }
                    acc[r] = sacc;
                }
            }
        }
        float m[NR];
        if (MODE == PREP) {
#pragma unroll
            for (int r = 0; r < NR; ++r)
                m[r] = acc[r];
        } else {
            /* the row sums in row order, from the lanes that hold them */
            float sum = 0.0f;
            if constexpr (PACKED && L == 2) {
                /* rows 2r and 2r + 1 sit in acc[r] of the pair's low and
                 * high lane: the low lane adds them in row order, taking
                 * each odd row from its neighbour by one DPP swap, and the
                 * high lane takes the low lane's sum.  Rows past x hold +0,
                 * which leaves the sum unchanged. */
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const float o = pair_swap(acc[r]);
                    sum = sum + acc[r];
                    sum = sum + o;
                }
                const float lo = pair_swap(sum);
                sum = k ? lo : sum;
            } else {
#pragma unroll
                for (int i = 0; i < S; ++i) {
                    float vi = __shfl(acc[i / L], (int)(base + (unsigned)(i % L)), WAVE);
                    if (!MIXED || (unsigned)i < x)
                        sum = sum + vi;
                }
            }
            if (sum <= 0.0f)
                sum = 1.0f;
#pragma unroll
            for (int r = 0; r < NR; ++r)
                m[r] = (!MIXED || k + (unsigned)(r * L) < x) ? acc[r] / sum : 0.0f;
        }
        if (act) {
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                unsigned i = k + r * L;
                if (i >= (unsigned)SP)
                    continue;
                if (i < (unsigned)S) {
                    if ((MODE == ITER || MODE == EDGE) && (!MIXED || i < x)) {
                        float qd = stage_f[(S + i) * W + q] - m[r];
                        if (qd != qd)
                            qd = 0.0f;
                        dacc += (double)fabsf(qd);
                    }
                    lds[i * W + q] = m[r];
                }
                if ((MODE == ITER || MODE == INIT || MODE == EDGE) && !BP_ROWS_MSG_FROM_LDS && !PACKED)
                    a.msg_out[(size_t)s * SP + i] = (i < (unsigned)S) ? m[r] : 0.0f;
            }
        }
    };
#pragma unroll 1
    for (int p = 0; p < L; ++p) {
        if ((unsigned)(p * G) >= se - sb)
            break; /* the chunk's slots are all done */
        float tv[NR][S];
        load_pass(p, tv);
        if constexpr (PACKED)
            __builtin_amdgcn_sched_barrier(0); /* every table load issued before the first product */
        run_pass(p, tv);
    }
    wave_sync();
    if ((MODE == ITER || MODE == INIT || MODE == EDGE) && (BP_ROWS_MSG_FROM_LDS || PACKED)) {
        /* the new messages, one whole row per slot (zeros beyond the
         * message's entries are in LDS already) */
        unsigned s = sb + lane;
        if (s < se) {
            float row[SP];
#pragma unroll
            for (int i = 0; i < SP; ++i)
                row[i] = (i < S) ? lds[i * W + lane] : 0.0f;
            if constexpr (PACKED)
                store_crow<S>(a.msg_out + mrow, stage_u[lane] & 15u, row);
            else
                store_row_nt<SP>(a.msg_out + (size_t)s * SP, row);
        }
    }
    if constexpr (PACKED && L == 2 && S == 8 && late_states && (MODE == ITER || MODE == INIT)) {
        if (ne - nb <= (unsigned)(W / 2)) {
            node_update_pair<S, MODE>(a, nrow, q0, q1, n, nb + (lane >> 1) < ne, lds, lane);
            return;
        }
    }
    if (late_states)
        load_states();
    if (v < ne)
        node_update<S, MIXED, MODE, PACKED>(a, PACKED ? (size_t)nrow : (size_t)v * SP, q0, q1, st, n, lds, W);
}

/* One node with more than 64 in-slots: its slots are walked in blocks of 64
 * and lane 0 multiplies in slot order.  ITER walks twice -- the product
 * for the belief update, then the combine buffer from the updated state --
 * recomputing the messages on the second walk instead of reading back what
 * other lanes wrote. */
template <int S, bool MIXED, int MODE, bool PACKED, bool POOLED>
__device__ void walk_wide(const LoopyArgs &a, unsigned sb, unsigned se, float *lds, unsigned lane, double &dacc,
                          bool first, float *acc, unsigned n, unsigned mbase)
{
    float m[S];
    for (unsigned base = sb; base < se; base += WAVE) {
        unsigned s = base + lane;
        unsigned mrow = 0;
        if constexpr (PACKED) {
            /* this rare path (a node with more in-slots than a chunk) sums
             * the widths serially: a prefix scan here would raise the whole
             * kernel's register count */
            unsigned end = (se - base < (unsigned)WAVE) ? se : base + WAVE;
            for (unsigned t = base; t < end; ++t) {
                if (t == s)
                    mrow = mbase;
                mbase += even_width(a.slot_xy[t] & 15u);
            }
        }
        if (s < se) {
            if (first) {
                slot_message<S, MIXED, MODE, PACKED, POOLED>(a, s, m, dacc, mrow);
            } else {
                unsigned x;
                compute_message<S, MIXED, MODE, PACKED, POOLED>(a, s, m, x, mrow);
            }
#pragma unroll
            for (int k = 0; k < S; ++k)
                lds[k * WAVE + lane] = m[k];
        }
        wave_sync();
        if (lane == 0) {
            unsigned cnt = (se - base < (unsigned)WAVE) ? se - base : (unsigned)WAVE;
            for (unsigned q = 0; q < cnt; ++q) {
#pragma unroll
                for (int k = 0; k < S; ++k)
                    m[k] = lds[k * WAVE + q];
                combine_mode<S, MIXED, MODE>(acc, m, n);
            }
        }
        wave_sync();
    }
}

template <int S, bool MIXED, int MODE, bool PACKED, bool POOLED>
__device__ void chunk_wide(const LoopyArgs &a, unsigned v, unsigned sb, unsigned se, float *lds, unsigned lane,
                           double &dacc, uint2 cbase)
{
    constexpr int SP = Stride<S>::value;
    unsigned n = MIXED ? a.nvars[v] : S;
    const size_t row = PACKED ? (size_t)cbase.x : (size_t)v * SP;
    float st[S], acc[S];
    if constexpr (PACKED) {
        load_crow<S>(a.states + row, st);
    } else {
#pragma unroll
        for (int k = 0; k < S; ++k)
            st[k] = a.states[row + k];
    }
#pragma unroll
    for (int k = 0; k < S; ++k)
        acc[k] = (MODE == ITER || MODE == FWD) ? 1.0f : st[k];
    walk_wide<S, MIXED, MODE, PACKED, POOLED>(a, sb, se, lds, lane, dacc, true, acc, n, cbase.y);
    if (MODE == ITER || MODE == FWD) {
#pragma unroll
        for (int k = 0; k < S; ++k)
            if (!MIXED || (unsigned)k < n)
                st[k] = st[k] * acc[k];
        normalise<S, MIXED>(st, n);
#pragma unroll
        for (int k = 0; k < S; ++k)
            acc[k] = st[k];
        walk_wide<S, MIXED, MODE, PACKED, POOLED>(a, sb, se, lds, lane, dacc, false, acc, n, cbase.y);
    }
    if (MODE == EDGE)
        normalise_positive<S, MIXED>(acc, n);
    if (lane == 0)
        write_node<S, MIXED, MODE, PACKED>(a, row, st, acc, n);
}

/* Slots per lane of the narrow path: two for small state counts. */
template <int S> struct SlotsPerLane {
    static constexpr int value = (S <= 4) ? BP_SLOTS_SMALL : 1;
};

/* One chunk: the narrow path, or the one-node path for a node with more
 * in-slots than a chunk holds. */
template <int S, bool MIXED, int MODE, int R, bool PACKED, int L, bool POOLED>
__device__ __forceinline__ void run_chunk(const LoopyArgs &a, unsigned c, float *lds, float *stage_f,
                                          unsigned *stage_u, unsigned lane, double &dacc)
{
    uint2 lo = uniform_load(a.chunk + c), hi = uniform_load(a.chunk + c + 1);
    unsigned nb = lo.x, sb = lo.y, ne = hi.x, se = hi.y;
    uint2 cbase = make_uint2(0u, 0u);
    if constexpr (PACKED)
        cbase = uniform_load(a.chunk_base + c);
    if (L > 1 && se - sb <= (unsigned)WAVE)
        chunk_rows<S, MIXED, MODE, L, PACKED, POOLED>(a, nb, ne, sb, se, lds, stage_f, stage_u, lane, dacc, cbase);
    else if (L == 1 && se - sb <= (unsigned)(WAVE * R))
        chunk_narrow<S, MIXED, MODE, R, PACKED, POOLED>(a, nb, ne, sb, se, lds, lane, dacc, cbase);
    else if constexpr (!PACKED)
        chunk_wide<S, MIXED, MODE, PACKED, POOLED>(a, nb, sb, se, lds, lane, dacc, cbase);
    /* (PACKED: a node with more in-slots than a chunk holds runs in
     * k_loopy_wide after this launch, so the rare one-node path does not
     * set this kernel's register budget) */
    wave_sync();
}

/* per-chunk delta partial (batched networks: chunks never span two
 * networks, so per-network deltas are sums over each network's chunks) */
__device__ __forceinline__ void flush_chunk_partial(const LoopyArgs &a, unsigned c, unsigned lane, double &dacc)
{
    double d = wave_sum(dacc);
    if (lane == 0)
        a.chunk_partials[c] = d;
    dacc = 0.0;
}

/* POOLED: a uniform-state handle whose transposed pairs share one stored
 * table (slot_tab[s] = pool entry, bit 31 = read transposed) */
template <int S, bool MIXED, int MODE, int R = SlotsPerLane<S>::value, bool PACKED = false, int L = 1,
          bool POOLED = false>
__global__ __launch_bounds__(BLOCK, (MinWaves<PACKED, S, L>::value)) void k_loopy(LoopyArgs a)
{
    if (a.stop && *a.stop)
        return;
    /* lanes-per-slot path: staged slot operands (rows b and old, state counts, tables) */
    __shared__ float stage_f[WAVES_PER_BLOCK][L > 1 ? 2 * S * WAVE : 1];
    __shared__ unsigned stage_u[WAVES_PER_BLOCK][L > 1 ? 2 * WAVE : 1];
    /* the new messages.  With L > 1 they take the place of the staged rows
     * b (the first S * WAVE floats of stage_f): a slot's message is written
     * by the instructions after those that read the slot's row b, and no
     * other slot reads that row, so 8 KB less LDS per workgroup (S = 8)
     * lets 8 workgroups share a CU instead of 6. */
    __shared__ float lds_own[WAVES_PER_BLOCK][L > 1 ? 1 : S * WAVE * R];
    unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE), lane = threadIdx.x % WAVE;
    float *lds_w = (L > 1) ? stage_f[wave] : lds_own[wave];
    double dacc = 0.0;
    const bool per_chunk = (MODE == ITER || MODE == EDGE) && a.chunk_partials != nullptr;
    if (a.queue) {
        /* work queue: each wave takes blocks of qblock consecutive positions
         * of the chunk order */
        for (;;) {
            unsigned base = 0;
            if (lane == 0)
                base = atomicAdd(a.queue, a.qblock);
            base = __builtin_amdgcn_readfirstlane(base);
            if (base >= a.nchunks)
                break;
            unsigned end = base + a.qblock < a.nchunks ? base + a.qblock : a.nchunks;
            for (unsigned p = base; p < end; ++p) {
                unsigned c = a.order ? uniform_load(a.order + p) : p;
                run_chunk<S, MIXED, MODE, R, PACKED, L, POOLED>(a, c, lds_w, stage_f[wave], stage_u[wave], lane, dacc);
                if (per_chunk)
                    flush_chunk_partial(a, c, lane, dacc);
            }
        }
    } else {
        /* wave w of the grid takes positions [w * nchunks / W,
         * (w + 1) * nchunks / W) of the chunk order, W = waves in the grid;
         * with the XCD remap each XCD's waves cover one contiguous range */
        unsigned W = gridDim.x * WAVES_PER_BLOCK;
        unsigned w = xcd_block(blockIdx.x, gridDim.x) * WAVES_PER_BLOCK + wave;
        unsigned c0 = (unsigned)(((unsigned long long)w * a.nchunks) / W);
        unsigned c1 = (unsigned)(((unsigned long long)(w + 1) * a.nchunks) / W);
        for (unsigned p = c0; p < c1; ++p) {
            unsigned c = a.order ? uniform_load(a.order + p) : p;
            run_chunk<S, MIXED, MODE, R, PACKED, L, POOLED>(a, c, lds_w, stage_f[wave], stage_u[wave], lane, dacc);
            if (per_chunk)
                flush_chunk_partial(a, c, lane, dacc);
        }
    }
    /* one FP64 partial per wave, no workgroup barrier: a wave that is done
     * leaves at once instead of waiting for the slowest of its group */
    if ((MODE == ITER || MODE == EDGE) && !per_chunk) {
        double d = wave_sum(dacc);
        if (lane == 0)
            a.partials[blockIdx.x * WAVES_PER_BLOCK + wave] = d;
    }
}

/* The chunks of a batch (PACKED) that hold one node with more in-slots than
 * a chunk of the current shape takes: one wave per chunk, launched after
 * k_loopy over the same chunks (which leaves their delta partials at 0). */
template <int S, bool MIXED, int MODE>
__global__ __launch_bounds__(BLOCK) void k_loopy_wide(LoopyArgs a, const unsigned *wide, unsigned nwide)
{
    __shared__ float lds[WAVES_PER_BLOCK][S * WAVE];
    unsigned wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    unsigned i = blockIdx.x * WAVES_PER_BLOCK + wave;
    if (i >= nwide)
        return;
    unsigned c = wide[i];
    uint2 lo = a.chunk[c], hi = a.chunk[c + 1];
    double dacc = 0.0;
    chunk_wide<S, MIXED, MODE, true, false>(a, lo.x, lo.y, hi.y, lds[wave], lane, dacc, a.chunk_base[c]);
    wave_sync();
    if (MODE == ITER || MODE == EDGE)
        flush_chunk_partial(a, c, lane, dacc);
}

} // namespace bpk
