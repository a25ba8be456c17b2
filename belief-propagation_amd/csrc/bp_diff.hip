/*
 * bp_diff.hip -- differentiable loopy BP: the parameter -> table map, the
 * unrolled forward with checkpoints, and the hand-written adjoint.
 *
 * Parameters are the TABLEs of a network's CPT definitions.  k_tables
 * rebuilds every slot's pairwise table from them with the XMLBIF reader's
 * loop (src/bnf-xml-parser/xml-expression.c:384-400, restated in
 * bp_model.c): for parent state j and child state k, the sequential f32 sum
 * of the TABLE entries in the blocks of that parent's state, so the tables
 * are bit-identical to the reader's.  k_initial puts root priors into the
 * initial beliefs, as graph_set_node_state does (xml-expression.c:290-317).
 *
 * The forward is the loopy kernel in its FWD/FINIT modes: the node-centric
 * iteration of src/graph/graph.c:1121-1183 from init_previous_edge
 * (graph.c:725-751), writing every iterate to the buffer it is given.
 *
 * Adjoint of one iteration t (k_adjoint, one launch over chunks of its own:
 * consecutive nodes with at most 64 x R in-slots).  In exact arithmetic, with N(x) = x / sum(x):
 *   m_t[s]  = N(z_s), z_s = J_s . buf_{t-1}[u],  u = src(s)
 *   T_t[v]  = prod_{s in in(v)} m_t[s]
 *   st_t[v] = N(a_v), a_v = st_{t-1}[v] (.) T_t[v]   (a_v = st_{t-1}[v] without in-edges)
 *   buf_t[v]= st_t[v] (.) T_t[v]
 * Given dst_t (from the loss or iteration t+1) and dbuf_t[v] =
 * sum_{s: src(s)=v} w_{t+1}[s]:
 *   dst_t   += dbuf_t (.) T_t          dT = dbuf_t (.) st_t
 *   da       = (dst_t - <dst_t, st_t>) / sum(a)   (identity when sum(a) <= 0,
 *              where the forward divides by 1)
 *   dst_{t-1} = da (.) T_t,  dT += da (.) st_{t-1}
 *   dm_s     = dT (.) prod_{s' in in(v), s' != s} m_t[s']
 *   dz_s     = (dm_s - <dm_s, m_s>) / sum(z_s)
 *   dJ_s    += dz_s (x) buf_{t-1}[u],  w_t[s] = J_s^T dz_s
 * Step 0 (init) has m_0 from st_0 and buf_0 = st_0 (.) T_0; its w_0 and
 * dbuf_0 (.) T_0 add to dst_0.  Sums over a node's out-edges are gathers in
 * slot order through a source-grouped slot list, so results do not depend
 * on scheduling (no atomics).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "bp_dev.h"
#include "../../include/bp_diff.h"

using namespace bpk;

/* Slots per lane of the adjoint for four or fewer states (one above). */
#ifndef BP_ADJ_SLOTS_SMALL
#define BP_ADJ_SLOTS_SMALL 2
#endif
/* 1: each slot's w is written at the slot's position in the source-grouped
 * list, so a node reads its out-edges' w contiguously; 0: w in slot order,
 * gathered through the list */
#ifndef BP_ADJ_W_SRC
#define BP_ADJ_W_SRC 1
#endif

namespace {

constexpr unsigned NONE = 0xffffffffu;

struct ParentRec {
    unsigned offset, n_par, slot_f, slot_r;
};

struct AdjArgs {
    const unsigned *in_off, *slot_src, *out_off, *out_slot;
    const unsigned *slot_pos; /* position of each slot in the source-grouped list (out_slot's inverse) */
    const float *J;
    const uint2 *chunk;
    unsigned nchunks;
    const float *st_prev; /* st_{t-1} */
    const float *st_cur;  /* st_t (step 0: st_0) */
    const float *bsrc;    /* buf_{t-1} (step 0: st_0) */
    const float *gst_in;  /* dL/dst_t along paths through later iterates and the loss */
    float *gst_out;       /* dL/dst_{t-1} through st_t (step 0: dL/dst_0 without w_0) */
    const float *w_in;    /* w_{t+1} or NULL */
    float *w_out;         /* w_t */
    float *jbar;          /* [E][S*S] */
    int jbar_first;       /* write instead of accumulate */
};

__global__ void k_copy(const float *__restrict__ src, float *__restrict__ dst, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = src[i];
}

/* One thread per (parent record, child state k, parent state j):
 * J_fwd[j][k] = the sequential f32 sum of TABLE[k * slice + idx] over the
 * configurations idx in which the parent is in state j, in ascending idx
 * order -- the blocks of the reader's loop, in its order -- written to the
 * forward slot and, transposed, to the reverse slot.  Loads are issued eight
 * at a time ahead of the adds, which stay in order; no division per entry. */
__global__ void k_tables(const float *__restrict__ P, const ParentRec *__restrict__ parents,
                         const unsigned *__restrict__ par_def, const unsigned *__restrict__ def_tb,
                         const unsigned *__restrict__ def_slice, float *__restrict__ J, unsigned nrec, unsigned S)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nrec * S * S)
        return;
    unsigned pr = (unsigned)(t / ((size_t)S * S)), kj = (unsigned)(t % ((size_t)S * S));
    unsigned k = kj / S, j = kj % S;
    ParentRec rec = parents[pr];
    if (j >= rec.n_par)
        return;
    unsigned d = par_def[pr];
    unsigned slice = def_slice[d], o = rec.offset, n = rec.n_par;
    const float *T = P + def_tb[d] + (size_t)k * slice + (size_t)j * o;
    unsigned cnt = slice / n; /* entries with the parent in state j */
    float acc = 0.0f;
    /* entry m of the sum is T[(m / o) * n * o + m % o]; (hi, lo) = (m / o,
     * m % o) is carried along instead of divided out */
    unsigned hi = 0, lo = 0;
    for (unsigned m0 = 0; m0 < cnt; m0 += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v[u] = (m0 + u < cnt) ? T[(size_t)hi * n * o + lo] : 0.0f;
            if (++lo == o) {
                lo = 0;
                ++hi;
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (m0 + u < cnt)
                acc += v[u];
    }
    size_t SS = (size_t)S * S;
    J[rec.slot_f * SS + j * S + k] = acc;
    if (rec.slot_r != NONE)
        J[rec.slot_r * SS + k * S + j] = acc;
}

__global__ void k_initial(const float *__restrict__ P, const unsigned *__restrict__ prior_tb,
                          const float *__restrict__ base, float *__restrict__ st0, unsigned V, unsigned S, unsigned SP)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)V * SP)
        return;
    unsigned v = (unsigned)(t / SP), k = (unsigned)(t % SP);
    unsigned tb = prior_tb[v];
    st0[t] = (tb != NONE && k < S) ? P[tb + k] : base[t];
}

/* acc += rows[o] for o in [o0, o1), in order; four rows loaded before any
 * is added */
template <int S, int SP>
__device__ __forceinline__ void sum_rows_contig(const float *__restrict__ rows, unsigned o0, unsigned o1, float *acc)
{
    for (unsigned o = o0; o < o1; o += 4) {
        float x[4][S];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            load_rowp<S, SP>(rows + (size_t)((o + u < o1) ? o + u : o) * SP, x[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (o + u < o1)
#pragma unroll
                for (int k = 0; k < S; ++k)
                    acc[k] = acc[k] + x[u][k];
    }
}

/* the node's sum of its out-edges' w (BP_ADJ_W_SRC: stored in list order) */
template <int S, int SP>
__device__ __forceinline__ void sum_out_w(const float *__restrict__ w, const unsigned *__restrict__ out_slot,
                                          unsigned o0, unsigned o1, float *acc);

/* acc += rows[out_slot[o]] for o in [o0, o1), in order; four index loads,
 * then four row loads, are issued before any is added */
template <int S, int SP>
__device__ __forceinline__ void sum_out_rows(const float *__restrict__ rows, const unsigned *__restrict__ out_slot,
                                             unsigned o0, unsigned o1, float *acc)
{
    for (unsigned o = o0; o < o1; o += 4) {
        unsigned idx[4];
        float x[4][S];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            idx[u] = out_slot[(o + u < o1) ? o + u : o];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            load_rowp<S, SP>(rows + (size_t)idx[u] * SP, x[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (o + u < o1)
#pragma unroll
                for (int k = 0; k < S; ++k)
                    acc[k] = acc[k] + x[u][k];
    }
}

template <int S, int SP>
__device__ __forceinline__ void sum_out_w(const float *__restrict__ w, const unsigned *__restrict__ out_slot,
                                          unsigned o0, unsigned o1, float *acc)
{
    if constexpr (BP_ADJ_W_SRC)
        sum_rows_contig<S, SP>(w, o0, o1, acc);
    else
        sum_out_rows<S, SP>(w, out_slot, o0, o1, acc);
}

template <int S, int STEP0, int R>
__global__ __launch_bounds__(BLOCK) void k_adjoint(AdjArgs a)
{
    constexpr int SP = Stride<S>::value;
    constexpr int W = WAVE * R;
    __shared__ float lm[WAVES_PER_BLOCK][S * W];
    __shared__ float ldt[WAVES_PER_BLOCK][S * W];
    __shared__ unsigned short lnode[WAVES_PER_BLOCK][W];
    __shared__ unsigned short lq[WAVES_PER_BLOCK][2 * W];
    unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE), lane = threadIdx.x % WAVE;
    unsigned Wt = gridDim.x * WAVES_PER_BLOCK;
    unsigned w = xcd_block(blockIdx.x, gridDim.x) * WAVES_PER_BLOCK + wave;
    unsigned c0 = (unsigned)(((unsigned long long)w * a.nchunks) / Wt);
    unsigned c1 = (unsigned)(((unsigned long long)(w + 1) * a.nchunks) / Wt);
    float *LM = lm[wave], *LDT = ldt[wave];
    unsigned short *LN = lnode[wave], *LQ = lq[wave];
    for (unsigned c = c0; c < c1; ++c) {
        uint2 lo = bpk::uniform_load(a.chunk + c), hi = bpk::uniform_load(a.chunk + c + 1);
        unsigned nb = lo.x, sb = lo.y, ne = hi.x, se = hi.y;
        /* slot phase 1: the message of every slot, kept with its operands */
        float Jr[R][S * S], b[R][S], m[R][S], sig[R];
        bool pos[R];
        unsigned u[R];
        /* lanes past the chunk's end load a valid neighbour (results
         * dropped): branch-free loads, all in flight before the gathers */
        const bool any_slots = se > sb; /* uniform: no stand-in slot otherwise */
        const unsigned s_last = se - 1;
        if (any_slots) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                unsigned s = sb + r * WAVE + lane;
                if (s >= se)
                    s = s_last;
                u[r] = a.slot_src[s];
                load_row<S * S>(a.J + (size_t)s * S * S, Jr[r]);
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
                load_rowp<S, SP>(a.bsrc + (size_t)u[r] * SP, b[r]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            unsigned s = sb + r * WAVE + lane;
            if (s < se) {
                float sum = 0.0f;
#pragma unroll
                for (int i = 0; i < S; ++i) {
                    float acc = 0.0f;
#pragma unroll
                    for (int j = 0; j < S; ++j)
                        acc = acc + Jr[r][i * S + j] * b[r][j];
                    m[r][i] = acc;
                    sum = sum + acc;
                }
                pos[r] = sum > 0.0f;
                if (!pos[r])
                    sum = 1.0f;
                sig[r] = sum;
#pragma unroll
                for (int i = 0; i < S; ++i) {
                    m[r][i] = m[r][i] / sum;
                    LM[i * W + r * WAVE + lane] = m[r][i];
                }
            }
        }
        wave_sync();
        /* node phase: dT, dst_{t-1} */
#pragma unroll
        for (int r = 0; r < R; ++r) {
            unsigned v = nb + r * WAVE + lane;
            if (v < ne) {
                unsigned q0 = a.in_off[v] - sb, q1 = a.in_off[v + 1] - sb;
                unsigned o0 = a.out_off[v], o1 = a.out_off[v + 1];
                float st[S], g[S], T[S], gb[S], dT[S];
                load_rowp<S, SP>(a.st_cur + (size_t)v * SP, st);
                load_rowp<S, SP>(a.gst_in + (size_t)v * SP, g);
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    T[k] = 1.0f;
                    gb[k] = 0.0f;
                }
                for (unsigned q = q0; q < q1; ++q) {
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        float mk = LM[k * W + q];
                        if (mk == mk)
                            T[k] = T[k] * mk;
                    }
                }
                if (a.w_in)
                    sum_out_w<S, SP>(a.w_in, a.out_slot, o0, o1, gb);
                float out[S];
                if (STEP0) {
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        dT[k] = gb[k] * st[k];
                        out[k] = g[k] + gb[k] * T[k];
                    }
                } else {
                    float sp[S], ds[S], av[S];
                    load_rowp<S, SP>(a.st_prev + (size_t)v * SP, sp);
                    bool indeg = q1 > q0;
                    float sa = 0.0f, dot = 0.0f;
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        ds[k] = g[k] + gb[k] * T[k];
                        dT[k] = gb[k] * st[k];
                        av[k] = indeg ? sp[k] * T[k] : sp[k];
                        sa = sa + av[k];
                        dot = dot + ds[k] * st[k];
                    }
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        float da = (sa > 0.0f) ? (ds[k] - dot) / sa : ds[k];
                        out[k] = indeg ? da * T[k] : da;
                        if (indeg)
                            dT[k] = dT[k] + da * sp[k];
                    }
                }
#pragma unroll
                for (int k = 0; k < SP; ++k)
                    a.gst_out[(size_t)v * SP + k] = (k < S) ? out[k] : 0.0f;
                unsigned l = r * WAVE + lane;
#pragma unroll
                for (int k = 0; k < S; ++k)
                    LDT[k * W + l] = dT[k];
                LQ[2 * l] = (unsigned short)q0;
                LQ[2 * l + 1] = (unsigned short)q1;
                for (unsigned q = q0; q < q1; ++q)
                    LN[q] = (unsigned short)l;
            }
        }
        wave_sync();
        /* slot phase 2: dm, dz, dJ, w */
#pragma unroll
        for (int r = 0; r < R; ++r) {
            unsigned s = sb + r * WAVE + lane;
            if (s < se) {
                unsigned q = r * WAVE + lane;
                unsigned l = LN[q], q0 = LQ[2 * l], q1 = LQ[2 * l + 1];
                float P[S];
#pragma unroll
                for (int k = 0; k < S; ++k)
                    P[k] = 1.0f;
                for (unsigned p = q0; p < q1; ++p) {
                    if (p == q)
                        continue;
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        float mk = LM[k * W + p];
                        if (mk == mk)
                            P[k] = P[k] * mk;
                    }
                }
                float dm[S], dz[S], dot = 0.0f;
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    dm[k] = (m[r][k] == m[r][k]) ? LDT[k * W + l] * P[k] : 0.0f;
                    dot = dot + dm[k] * m[r][k];
                }
#pragma unroll
                for (int k = 0; k < S; ++k)
                    dz[k] = pos[r] ? (dm[k] - dot) / sig[r] : dm[k];
                float *jb = a.jbar + (size_t)s * S * S;
                float gJ[S * S];
                if (!a.jbar_first)
                    load_row<S * S>(jb, gJ);
#pragma unroll
                for (int i = 0; i < S; ++i)
#pragma unroll
                    for (int j = 0; j < S; ++j)
                        gJ[i * S + j] = a.jbar_first ? dz[i] * b[r][j] : gJ[i * S + j] + dz[i] * b[r][j];
#pragma unroll
                for (int i = 0; i < S * S; ++i)
                    jb[i] = gJ[i];
                float wv[SP];
#pragma unroll
                for (int j = 0; j < SP; ++j) {
                    float acc = 0.0f;
                    if (j < S) {
#pragma unroll
                        for (int i = 0; i < S; ++i)
                            acc = acc + Jr[r][i * S + j] * dz[i];
                    }
                    wv[j] = acc;
                }
                size_t wrow = BP_ADJ_W_SRC ? (size_t)a.slot_pos[s] : (size_t)s;
#pragma unroll
                for (int j = 0; j < SP; ++j)
                    a.w_out[wrow * SP + j] = wv[j];
            }
        }
        wave_sync();
    }
}

/* dL/dst_0 = the step-0 partial + sum of w_0 over the node's out-slots */
template <int S>
__global__ void k_initial_grad(const float *__restrict__ part, const float *__restrict__ w0,
                               const unsigned *__restrict__ out_off, const unsigned *__restrict__ out_slot,
                               float *__restrict__ out, unsigned V)
{
    constexpr int SP = Stride<S>::value;
    unsigned v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V)
        return;
    float acc[S];
    load_rowp<S, SP>(part + (size_t)v * SP, acc);
    sum_out_w<S, SP>(w0, out_slot, out_off[v], out_off[v + 1], acc);
#pragma unroll
    for (int k = 0; k < SP; ++k)
        out[(size_t)v * SP + k] = (k < S) ? acc[k] : 0.0f;
}

/* G[pr][k][j] = dJ of the forward slot at [j][k] + dJ of the reverse slot at
 * [k][j]: everything a parameter entry of the definition gets through
 * parent record pr when that parent is in state j and the child in k */
__global__ void k_parent_grad(const float *__restrict__ jbar, const ParentRec *__restrict__ parents,
                              float *__restrict__ G, unsigned nrec, unsigned S)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t SS = (size_t)S * S;
    if (t >= (size_t)nrec * SS)
        return;
    unsigned pr = (unsigned)(t / SS), kj = (unsigned)(t % SS), k = kj / S, j = kj % S;
    ParentRec rec = parents[pr];
    float g = 0.0f;
    if (j < rec.n_par) {
        g = jbar[(size_t)rec.slot_f * SS + j * S + k];
        if (rec.slot_r != NONE)
            g = g + jbar[(size_t)rec.slot_r * SS + k * S + j];
    }
    G[t] = g;
}

/* dL/dparam: the transpose of k_tables and k_initial, from the per-parent
 * sums G: parameter k * slice + idx of a definition gets, for each parent
 * record in order, G[record][k][j] with j that parent's state in
 * configuration idx.  Every node has S states, so a definition's records,
 * which run from its last GIVEN (offset 1) to its first, read the base-S
 * digits of idx from the lowest up: one division by the constant S per
 * record.  The G entries of up to eight records are loaded before any is
 * added. */
struct DefInfo {
    unsigned tb, slice, pbegin, pend, prior, rbegin, pad1, pad2; /* rbegin: the definition's first softmax row */
};

template <int S>
__global__ void k_param_grad(const float *__restrict__ G, const float *__restrict__ dst0,
                             const unsigned *__restrict__ def_of, const DefInfo *__restrict__ defs,
                             float *__restrict__ out, unsigned num_params, unsigned SP)
{
    constexpr unsigned SS = (unsigned)S * S;
    unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= num_params)
        return;
    DefInfo di = defs[def_of[p]];
    unsigned local = p - di.tb;
    unsigned k = local / di.slice, x = local - k * di.slice;
    const float *Gk = G + k * S;
    float g = 0.0f;
    for (unsigned base = di.pbegin; base < di.pend; base += 8) {
        float f[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            f[u] = 0.0f;
            if (base + u < di.pend) {
                unsigned q = x / (unsigned)S, j = x - q * (unsigned)S;
                x = q;
                f[u] = Gk[(size_t)(base + u) * SS + j];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (base + u < di.pend)
                g = g + f[u];
    }
    if (di.prior != NONE)
        g = g + dst0[(size_t)di.prior * SP + k];
    out[p] = g;
}

/* dL/dtheta for the log-parameters directly: one thread per softmax row
 * (a definition's child states under one parent configuration cc).  Every
 * entry of the row has the same configuration, so its digits are formed
 * once; each entry's parameter gradient is k_param_grad's sum, in its
 * order, and the row then goes through the softmax's transpose
 * dtheta = P (.) (dP - <P, dP>) -- the same operations as k_param_grad
 * followed by k_softmax_back, without the parameter gradient in memory. */
template <int S>
__global__ void k_param_grad_log(const float *__restrict__ G, const float *__restrict__ dst0,
                                 const unsigned *__restrict__ row_def, const DefInfo *__restrict__ defs,
                                 const float *__restrict__ P, float *__restrict__ gtheta, unsigned nrows, unsigned SP)
{
    constexpr unsigned SS = (unsigned)S * S;
    constexpr int U = (S <= 4) ? 8 : 4;
    unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows)
        return;
    DefInfo di = defs[row_def[r]];
    unsigned cc = r - di.rbegin, x = cc;
    float g[S];
#pragma unroll
    for (int k = 0; k < S; ++k)
        g[k] = 0.0f;
    for (unsigned base = di.pbegin; base < di.pend; base += U) {
        float f[U][S];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (base + u < di.pend) {
                unsigned q = x / (unsigned)S, j = x - q * (unsigned)S;
                x = q;
                const float *Gr = G + (size_t)(base + u) * SS + j;
#pragma unroll
                for (int k = 0; k < S; ++k)
                    f[u][k] = Gr[k * S];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (base + u < di.pend)
#pragma unroll
                for (int k = 0; k < S; ++k)
                    g[k] = g[k] + f[u][k];
    }
    if (di.prior != NONE)
#pragma unroll
        for (int k = 0; k < S; ++k)
            g[k] = g[k] + dst0[(size_t)di.prior * SP + k];
    float pv[S], dot = 0.0f;
#pragma unroll
    for (int k = 0; k < S; ++k) {
        pv[k] = P[di.tb + (size_t)k * di.slice + cc];
        dot += pv[k] * g[k];
    }
#pragma unroll
    for (int k = 0; k < S; ++k)
        gtheta[di.tb + (size_t)k * di.slice + cc] = pv[k] * (g[k] - dot);
}

/* P = softmax of theta over each row (a definition's child states under
 * one parent configuration): entries first[r] + k * stride[r], k < S */
__global__ void k_softmax(const float *__restrict__ theta, const unsigned *__restrict__ first,
                          const unsigned *__restrict__ stride, float *__restrict__ P, unsigned nrows, unsigned S)
{
    unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows)
        return;
    unsigned f = first[r], st = stride[r];
    float mx = -INFINITY;
    for (unsigned k = 0; k < S; ++k)
        mx = fmaxf(mx, theta[f + k * st]);
    float e[MAX_STATES], sum = 0.0f;
    for (unsigned k = 0; k < S; ++k) {
        e[k] = expf(theta[f + k * st] - mx);
        sum += e[k];
    }
    for (unsigned k = 0; k < S; ++k)
        P[f + k * st] = e[k] / sum;
}

/* dst[i] = src[idx[i]] for rows of SP floats (between the caller's node
 * numbering and the device's) */
__global__ void k_gather_rows(const float *__restrict__ src, const unsigned *__restrict__ idx,
                              float *__restrict__ dst, unsigned V, unsigned SP)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)V * SP)
        return;
    unsigned i = (unsigned)(t / SP), k = (unsigned)(t % SP);
    dst[t] = src[(size_t)idx[i] * SP + k];
}

/* slot order -> reference edge order */
__global__ void k_slot_to_edge(const float *__restrict__ src, const unsigned *__restrict__ slot_edge,
                               float *__restrict__ dst, unsigned E, unsigned width)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)E * width)
        return;
    unsigned s = (unsigned)(t / width), i = (unsigned)(t % width);
    dst[(size_t)slot_edge[s] * width + i] = src[t];
}

inline unsigned blocks_for(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

} // namespace

struct bp_diff {
    bp_dev *d = nullptr;
    unsigned V = 0, E = 0, S = 0, SP = 0;
    unsigned num_params = 0, num_rows = 0;
    std::vector<uint64_t> row_first;
    std::vector<uint32_t> row_stride, row_len;

    unsigned *par_def = nullptr; /* definition of each parent record */
    unsigned nrec = 0;
    unsigned *prior_tb = nullptr;
    unsigned *out_off = nullptr, *out_slot = nullptr, *slot_pos = nullptr, *slot_edge = nullptr;
    unsigned *def_of = nullptr, *def_tb = nullptr, *def_slice = nullptr, *def_pbegin = nullptr, *def_prior = nullptr;
    DefInfo *defs = nullptr; /* the same per definition, packed for k_param_grad */
    float *G = nullptr;      /* [nrec][S][S] per-parent gradient sums */
    ParentRec *parents = nullptr;

    float *P = nullptr;     /* parameters */
    unsigned *perm_d = nullptr, *inv_d = nullptr; /* node -> device node, device node -> node */
    unsigned *row_first_d = nullptr, *row_stride_d = nullptr; /* softmax rows */
    unsigned *row_def_d = nullptr;                            /* each softmax row's definition */
    float *st0 = nullptr;   /* initial beliefs */
    float *gin = nullptr;   /* dL/dbeliefs (copy) */
    float *gst[2] = {nullptr, nullptr};
    float *w[2] = {nullptr, nullptr};
    float *jbar = nullptr;
    float *dst0p = nullptr, *dst0 = nullptr;
    /* the adjoint's chunks: at most 64 x adj_R slots and nodes each */
    uint2 *adj_chunk = nullptr;
    unsigned adj_R = 1, adj_nchunks = 0, adj_G = 8;

    /* iterates: checkpoints t = 0, c, 2c, ...; a segment's recomputed ones;
     * the forward's scratch; the last beliefs */
    unsigned K = 0, c = 0, nck = 0;
    bool have_forward = false, have_params = false;
    float *ck_st = nullptr, *ck_buf = nullptr;
    float *seg_st = nullptr, *seg_buf = nullptr;
    float *fw_st[2] = {nullptr, nullptr}, *fw_buf[2] = {nullptr, nullptr};
    float *final_st = nullptr;
    std::map<std::pair<unsigned, unsigned>, hipGraphExec_t> fwd_graphs, bwd_graphs;
};

static void diff_release_storage(bp_diff *x)
{
    float *ps[] = {x->ck_st, x->ck_buf, x->seg_st, x->seg_buf, x->fw_st[0], x->fw_st[1], x->fw_buf[0], x->fw_buf[1],
                   x->final_st};
    for (float *p : ps)
        if (p)
            (void)hipFree(p);
    x->ck_st = x->ck_buf = x->seg_st = x->seg_buf = x->final_st = nullptr;
    x->fw_st[0] = x->fw_st[1] = x->fw_buf[0] = x->fw_buf[1] = nullptr;
    for (auto &kv : x->fwd_graphs)
        (void)hipGraphExecDestroy(kv.second);
    for (auto &kv : x->bwd_graphs)
        (void)hipGraphExecDestroy(kv.second);
    x->fwd_graphs.clear();
    x->bwd_graphs.clear();
    x->K = x->c = x->nck = 0;
}

extern "C" void bp_diff_free(bp_diff *x)
{
    if (!x)
        return;
    if (x->d) {
        (void)hipSetDevice(x->d->device);
        (void)hipDeviceSynchronize();
    }
    diff_release_storage(x);
    void *ps[] = {x->adj_chunk, x->defs, x->G, x->perm_d, x->inv_d, x->row_first_d, x->row_stride_d, x->row_def_d, x->par_def, x->prior_tb, x->out_off, x->out_slot, x->slot_pos, x->slot_edge, x->def_of, x->def_tb, x->def_slice,
                  x->def_pbegin, x->def_prior, x->parents, x->P, x->st0, x->gin, x->gst[0], x->gst[1],
                  x->w[0], x->w[1], x->jbar, x->dst0p, x->dst0};
    for (void *p : ps)
        if (p)
            (void)hipFree(p);
    bp_free(x->d);
    delete x;
}

template <typename T> static int upload_vec(T **dst, const std::vector<T> &v)
{
    if (dmalloc(dst, v.size()))
        return -1;
    if (!v.empty())
        HIP_TRY(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

/* checks that need no device: one state count, in-degrees within a wave's
 * chunk, observations after the parameters, complete tables */
static int diff_validate(const bp_bn *bn, unsigned npd, Graph_t g)
{
    unsigned V = g->current_num_vertices;
    unsigned S = V ? g->node_num_vars[0] : 1;
    const unsigned R = S <= 4 ? (unsigned)BP_ADJ_SLOTS_SMALL : 1u;
    for (unsigned v = 0; v < V; ++v)
        if (g->node_num_vars[v] != S) {
            bp_set_error("bp_diff_create: node %u has %u states, node 0 has %u; the adjoint needs one state count", v,
                         g->node_num_vars[v], S);
            return -1;
        }
    for (unsigned v = 0; v < V; ++v) {
        unsigned deg = bp_csr_end(g, g->dest_nodes_to_edges_node_list, v) - g->dest_nodes_to_edges_node_list[v];
        if (deg > WAVE * R) {
            bp_set_error("bp_diff_create: node %u has %u in-edges (the adjoint handles up to %u)", v, deg, WAVE * R);
            return -1;
        }
    }
    if (bn->def_table_begin[npd] > 0xfffffff0ull) {
        bp_set_error("bp_diff_create: too many parameters");
        return -1;
    }
    for (unsigned dd = 0; dd < bn->num_defs; ++dd) {
        uint64_t g0 = bn->def_given_begin[dd], g1 = bn->def_given_begin[dd + 1];
        uint64_t t0 = bn->def_table_begin[dd], t1 = bn->def_table_begin[dd + 1];
        if (dd >= npd) {
            if (g1 != g0) {
                bp_set_error("bp_diff_create: definition %u after the parameters has GIVENs", dd);
                return -1;
            }
            continue;
        }
        uint64_t expect = bn->num_states[bn->def_for[dd]];
        for (uint64_t i = g0; i < g1; ++i)
            expect *= bn->num_states[bn->def_given[i]];
        if (t1 - t0 != expect) {
            bp_set_error("bp_diff_create: definition %u has %llu TABLE entries, %llu expected", dd,
                         (unsigned long long)(t1 - t0), (unsigned long long)expect);
            return -1;
        }
    }
    return 0;
}

static int diff_build(bp_diff *x, const bp_bn *bn, unsigned npd, Graph_t g)
{
    bp_dev *d = x->d;
    unsigned V = d->V, E = d->E, S = d->S;
    x->num_params = (unsigned)bn->def_table_begin[npd];

    /* replay the reader's edge emission: definitions in order, parents last
     * GIVEN first, forward edge then reverse edge unless the parent is
     * observed (bp_graph_add_cpt_edges) */
    std::vector<unsigned> def_of(x->num_params), def_tb(npd), def_slice(npd), def_pbegin(npd + 1, 0), def_prior(npd, NONE);
    std::vector<ParentRec> prs;
    std::vector<unsigned> par_def;
    unsigned e = 0;
    for (unsigned dd = 0; dd < bn->num_defs; ++dd) {
        uint64_t g0 = bn->def_given_begin[dd], g1 = bn->def_given_begin[dd + 1];
        uint64_t t0 = bn->def_table_begin[dd], t1 = bn->def_table_begin[dd + 1];
        unsigned child = bn->def_for[dd];
        unsigned nch = bn->num_states[child];
        if (dd >= npd)
            continue;
        def_tb[dd] = (unsigned)t0;
        unsigned slice = (unsigned)((t1 - t0) / nch);
        def_slice[dd] = slice;
        for (uint64_t p = t0; p < t1; ++p)
            def_of[p] = dd;
        unsigned offset = 1;
        for (uint64_t pi = g1; pi-- > g0;) {
            unsigned parent = bn->def_given[pi];
            unsigned npar = bn->num_states[parent];
            ParentRec pr{offset, npar, NONE, NONE};
            if (e >= E || g->edges_src_index[e] != parent || g->edges_dest_index[e] != child) {
                bp_set_error("bp_diff_create: edge %u does not follow the reader's order", e);
                return -1;
            }
            pr.slot_f = d->edge_slot[e];
            ++e;
            if (g->observed_nodes[parent] != 1) {
                if (e >= E || g->edges_src_index[e] != child || g->edges_dest_index[e] != parent) {
                    bp_set_error("bp_diff_create: edge %u does not follow the reader's order", e);
                    return -1;
                }
                pr.slot_r = d->edge_slot[e];
                ++e;
            }
            prs.push_back(pr);
            par_def.push_back(dd);
            offset *= npar;
        }
        def_pbegin[dd + 1] = (unsigned)prs.size();
    }
    if (e != E) {
        bp_set_error("bp_diff_create: %u of %u edges come from the definitions", e, E);
        return -1;
    }
    /* initial beliefs: the last GIVEN-less definition of a node sets them */
    std::vector<unsigned> state_def(V, NONE), prior_tb(V, NONE);
    for (unsigned dd = 0; dd < bn->num_defs; ++dd)
        if (bn->def_given_begin[dd + 1] == bn->def_given_begin[dd])
            state_def[bn->def_for[dd]] = dd;
    /* per device node (the handle numbers nodes in its own order) */
    for (unsigned v = 0; v < V; ++v)
        if (state_def[v] != NONE && state_def[v] < npd) {
            prior_tb[d->perm[v]] = def_tb[state_def[v]];
            def_prior[state_def[v]] = d->perm[v];
        }
    std::vector<DefInfo> definfo(npd);
    for (unsigned dd = 0; dd < npd; ++dd)
        definfo[dd] = DefInfo{def_tb[dd], def_slice[dd], def_pbegin[dd], def_pbegin[dd + 1], def_prior[dd], 0, 0, 0};
    for (unsigned dd = 0, rows = 0; dd < npd; ++dd) {
        definfo[dd].rbegin = rows;
        rows += def_slice[dd];
    }
    /* softmax rows */
    x->row_first.clear();
    x->row_stride.clear();
    x->row_len.clear();
    std::vector<unsigned> row_def;
    for (unsigned dd = 0; dd < npd; ++dd) {
        unsigned nch = bn->num_states[bn->def_for[dd]];
        for (unsigned cc = 0; cc < def_slice[dd]; ++cc) {
            x->row_first.push_back((uint64_t)def_tb[dd] + cc);
            x->row_stride.push_back(def_slice[dd]);
            x->row_len.push_back(nch);
            row_def.push_back(dd);
        }
    }
    x->num_rows = (unsigned)x->row_first.size();
    {
        std::vector<unsigned> rf(x->row_first.begin(), x->row_first.end());
        if (upload_vec(&x->row_first_d, rf) || upload_vec(&x->row_stride_d, x->row_stride) ||
            upload_vec(&x->row_def_d, row_def))
            return -1;
    }
    x->nrec = (unsigned)prs.size();
    /* source-grouped slots */
    std::vector<unsigned> out_off(V + 1, 0), out_slot(E), slot_src(E);
    for (unsigned s = 0; s < E; ++s) {
        slot_src[s] = d->perm[g->edges_src_index[d->slot_edge[s]]];
        out_off[slot_src[s] + 1]++;
    }
    for (unsigned v = 0; v < V; ++v)
        out_off[v + 1] += out_off[v];
    {
        std::vector<unsigned> fill(out_off.begin(), out_off.end() - 1);
        for (unsigned s = 0; s < E; ++s)
            out_slot[fill[slot_src[s]]++] = s;
    }
    std::vector<unsigned> slot_pos(E);
    for (unsigned o = 0; o < E; ++o)
        slot_pos[out_slot[o]] = o;
    {
        std::vector<unsigned> in_off(V + 1, 0);
        for (unsigned s = 0; s < E; ++s)
            in_off[d->perm[g->edges_dest_index[d->slot_edge[s]]] + 1]++;
        for (unsigned v = 0; v < V; ++v)
            in_off[v + 1] += in_off[v];
        x->adj_R = (S <= 4) ? (unsigned)BP_ADJ_SLOTS_SMALL : 1u;
        std::vector<uint2> ch;
        bp_build_chunks(in_off, V, WAVE * x->adj_R, ch);
        x->adj_nchunks = (unsigned)ch.size() - 1;
        unsigned want = (x->adj_nchunks + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
        x->adj_G = std::max(8u, (want + 7u) & ~7u);
        if (upload_vec(&x->adj_chunk, ch))
            return -1;
    }
    size_t VS = (size_t)V * d->SP, ES = (size_t)E * d->SP;
    if (upload_vec(&x->perm_d, d->perm) || upload_vec(&x->inv_d, d->inv) || upload_vec(&x->prior_tb, prior_tb) ||
        upload_vec(&x->out_off, out_off) ||
        upload_vec(&x->out_slot, out_slot) || upload_vec(&x->slot_pos, slot_pos) || upload_vec(&x->slot_edge, d->slot_edge) ||
        upload_vec(&x->def_of, def_of) || upload_vec(&x->def_tb, def_tb) || upload_vec(&x->def_slice, def_slice) ||
        upload_vec(&x->def_pbegin, def_pbegin) || upload_vec(&x->def_prior, def_prior) ||
        upload_vec(&x->defs, definfo) || dmalloc(&x->G, std::max<size_t>((size_t)prs.size() * S * S, 1)) ||
        upload_vec(&x->parents, prs) || upload_vec(&x->par_def, par_def) || dmalloc(&x->P, x->num_params) ||
        dmalloc(&x->st0, VS) || dmalloc(&x->gin, VS) || dmalloc(&x->gst[0], VS) || dmalloc(&x->gst[1], VS) ||
        dmalloc(&x->w[0], ES) || dmalloc(&x->w[1], ES) || dmalloc(&x->jbar, (size_t)E * S * S) ||
        dmalloc(&x->dst0p, VS) || dmalloc(&x->dst0, VS))
        return -1;
    return 0;
}

extern "C" int bp_diff_create(const bp_bn *bn, unsigned int num_param_defs, int device, bp_diff **out)
{
    bp_clear_error();
    if (!bn || !out) {
        bp_set_error("bp_diff_create: NULL argument");
        return -1;
    }
    *out = nullptr;
    if (num_param_defs > bn->num_defs) {
        bp_set_error("bp_diff_create: %u parameter definitions of %u", num_param_defs, bn->num_defs);
        return -1;
    }
    int abort_prev = bp_set_abort_on_error(0);
    Graph_t g = bp_bn_to_graph(bn);
    if (g) {
        set_up_src_nodes_to_edges(g);
        set_up_dest_nodes_to_edges(g);
    }
    bp_set_abort_on_error(abort_prev);
    if (!g || bp_has_error()) {
        if (g)
            graph_destroy(g);
        return -1;
    }
    if (diff_validate(bn, num_param_defs, g)) {
        graph_destroy(g);
        return -1;
    }
    bp_diff *x = new bp_diff();
    int rc = bp_upload_graph(g, device, &x->d);
    if (rc == 0) {
        x->V = x->d->V;
        x->E = x->d->E;
        x->S = x->d->S;
        x->SP = x->d->SP;
        rc = diff_build(x, bn, num_param_defs, g);
    }
    graph_destroy(g);
    if (rc) {
        bp_diff_free(x);
        return -1;
    }
    *out = x;
    return 0;
}

extern "C" int bp_diff_info(const bp_diff *x, uint64_t *info)
{
    if (!x || !info) {
        bp_set_error("bp_diff_info: NULL argument");
        return -1;
    }
    info[0] = x->V;
    info[1] = x->E;
    info[2] = x->S;
    info[3] = x->SP;
    info[4] = x->num_params;
    info[5] = x->num_rows;
    return 0;
}

extern "C" int bp_diff_rows(const bp_diff *x, uint64_t *row_first, uint32_t *row_stride, uint32_t *row_len)
{
    if (!x || !row_first || !row_stride || !row_len) {
        bp_set_error("bp_diff_rows: NULL argument");
        return -1;
    }
    std::copy(x->row_first.begin(), x->row_first.end(), row_first);
    std::copy(x->row_stride.begin(), x->row_stride.end(), row_stride);
    std::copy(x->row_len.begin(), x->row_len.end(), row_len);
    return 0;
}

extern "C" int bp_diff_set_params(bp_diff *x, const float *params, void *stream)
{
    bp_clear_error();
    if (!x || !params) {
        bp_set_error("bp_diff_set_params: NULL argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(x->d->device));
    if (x->num_params)
        hipLaunchKernelGGL(k_copy, dim3(blocks_for(x->num_params)), dim3(256), 0, st, params, x->P,
                           (size_t)x->num_params);
    size_t nT = (size_t)x->nrec * x->S * x->S;
    if (nT)
        hipLaunchKernelGGL(k_tables, dim3(blocks_for(nT)), dim3(256), 0, st, x->P, x->parents, x->par_def, x->def_tb,
                           x->def_slice, x->d->J, x->nrec, x->S);
    size_t VS = (size_t)x->V * x->SP;
    if (VS)
        hipLaunchKernelGGL(k_initial, dim3(blocks_for(VS)), dim3(256), 0, st, x->P, x->prior_tb, x->d->states0,
                           x->st0, x->V, x->S, x->SP);
    HIP_TRY(hipGetLastError());
    x->have_params = true;
    x->have_forward = false;
    return 0;
}

extern "C" int bp_diff_set_log_params(bp_diff *x, const float *theta, void *stream)
{
    bp_clear_error();
    if (!x || !theta) {
        bp_set_error("bp_diff_set_log_params: NULL argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(x->d->device));
    if (x->num_rows)
        hipLaunchKernelGGL(k_softmax, dim3(blocks_for(x->num_rows)), dim3(256), 0, st, theta, x->row_first_d,
                           x->row_stride_d, x->P, x->num_rows, x->S);
    size_t nT = (size_t)x->nrec * x->S * x->S;
    if (nT)
        hipLaunchKernelGGL(k_tables, dim3(blocks_for(nT)), dim3(256), 0, st, x->P, x->parents, x->par_def, x->def_tb,
                           x->def_slice, x->d->J, x->nrec, x->S);
    size_t VS = (size_t)x->V * x->SP;
    if (VS)
        hipLaunchKernelGGL(k_initial, dim3(blocks_for(VS)), dim3(256), 0, st, x->P, x->prior_tb, x->d->states0,
                           x->st0, x->V, x->S, x->SP);
    HIP_TRY(hipGetLastError());
    x->have_params = true;
    x->have_forward = false;
    return 0;
}

static int diff_storage(bp_diff *x, unsigned K, unsigned c)
{
    if (x->K == K && x->c == c && x->final_st)
        return 0;
    diff_release_storage(x);
    size_t VS = (size_t)x->V * x->SP;
    unsigned nck = K / c + 1;
    if (dmalloc(&x->ck_st, nck * VS) || dmalloc(&x->ck_buf, nck * VS) || dmalloc(&x->seg_st, (size_t)c * VS) ||
        dmalloc(&x->seg_buf, (size_t)c * VS) || dmalloc(&x->fw_st[0], VS) || dmalloc(&x->fw_st[1], VS) ||
        dmalloc(&x->fw_buf[0], VS) || dmalloc(&x->fw_buf[1], VS) || dmalloc(&x->final_st, VS)) {
        diff_release_storage(x);
        return -1;
    }
    x->K = K;
    x->c = c;
    x->nck = nck;
    return 0;
}

/* one FWD (or FINIT) launch: (st_in, buf_in) -> (st_out, buf_out) */
static void fwd_step(bp_diff *x, int mode, const float *st_in, const float *buf_in, float *st_out, float *buf_out,
                     hipStream_t st)
{
    LoopyArgs a = bp_make_args(x->d);
    a.states = const_cast<float *>(st_in);
    a.states_out = st_out;
    a.buf_in = buf_in;
    a.buf_out = buf_out;
    a.msg_in = nullptr;
    a.msg_out = nullptr;
    a.partials = nullptr;
    a.stop = nullptr;
    bp_launch_loopy(x->d, mode, a, st);
}

static int capture_begin(bp_diff *x)
{
    HIP_TRY(hipStreamBeginCapture(x->d->stream, hipStreamCaptureModeThreadLocal));
    return 0;
}

static int capture_end(bp_diff *x, hipGraphExec_t *exec)
{
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamEndCapture(x->d->stream, &graph));
    hipError_t err = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (err != hipSuccess) {
        bp_set_error("hipGraphInstantiate failed: %s", hipGetErrorString(err));
        return -1;
    }
    return 0;
}

static void enqueue_forward(bp_diff *x, hipStream_t st)
{
    size_t VS = (size_t)x->V * x->SP;
    unsigned K = x->K, c = x->c;
    fwd_step(x, FINIT, x->st0, nullptr, nullptr, x->ck_buf, st);
    const float *sp = x->st0, *bp = x->ck_buf;
    for (unsigned t = 1; t <= K; ++t) {
        float *so, *bo;
        if (t % c == 0) {
            so = x->ck_st + (size_t)(t / c) * VS;
            bo = x->ck_buf + (size_t)(t / c) * VS;
        } else {
            so = x->fw_st[t & 1];
            bo = x->fw_buf[t & 1];
        }
        fwd_step(x, FWD, sp, bp, so, bo, st);
        sp = so;
        bp = bo;
    }
    hipLaunchKernelGGL(k_copy, dim3(blocks_for(VS)), dim3(256), 0, st, sp, x->final_st, VS);
}

extern "C" int bp_diff_forward(bp_diff *x, unsigned int K, unsigned int c, float *beliefs, void *stream)
{
    bp_clear_error();
    if (!x || !beliefs || c == 0) {
        bp_set_error("bp_diff_forward: NULL argument or zero checkpoint interval");
        return -1;
    }
    if (!x->have_params) {
        bp_set_error("bp_diff_forward: bp_diff_set_params has not been called");
        return -1;
    }
    HIP_TRY(hipSetDevice(x->d->device));
    if (diff_storage(x, K, c))
        return -1;
    /* the ck_st[0] slot is the initial beliefs themselves */
    size_t VS = (size_t)x->V * x->SP;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_copy, dim3(blocks_for(VS)), dim3(256), 0, st, x->st0, x->ck_st, VS);
    auto key = std::make_pair(K, c);
    auto it = x->fwd_graphs.find(key);
    if (it == x->fwd_graphs.end()) {
        hipGraphExec_t exec = nullptr;
        if (capture_begin(x))
            return -1;
        enqueue_forward(x, x->d->stream);
        if (capture_end(x, &exec))
            return -1;
        it = x->fwd_graphs.emplace(key, exec).first;
    }
    HIP_TRY(hipGraphLaunch(it->second, st));
    hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for(VS)), dim3(256), 0, st, x->final_st, x->perm_d, beliefs, x->V,
                       x->SP);
    HIP_TRY(hipGetLastError());
    x->have_forward = true;
    return 0;
}

static void adj_step(bp_diff *x, int step0, const float *st_prev, const float *st_cur, const float *bsrc,
                     const float *gst_in, float *gst_out, const float *w_in, float *w_out, int first, hipStream_t st)
{
    bp_dev *d = x->d;
    AdjArgs a;
    a.in_off = d->in_off;
    a.slot_src = d->slot_src;
    a.out_off = x->out_off;
    a.out_slot = x->out_slot;
    a.slot_pos = x->slot_pos;
    a.J = d->J;
    a.chunk = x->adj_chunk;
    a.nchunks = x->adj_nchunks;
    a.st_prev = st_prev;
    a.st_cur = st_cur;
    a.bsrc = bsrc;
    a.gst_in = gst_in;
    a.gst_out = gst_out;
    a.w_in = w_in;
    a.w_out = w_out;
    a.jbar = x->jbar;
    a.jbar_first = first;
    dim3 grid(x->adj_G), block(BLOCK);
    constexpr int RS = BP_ADJ_SLOTS_SMALL;
#define BP_ADJ_CASE(SS)                                                                   \
    case SS:                                                                              \
        if (step0)                                                                        \
            hipLaunchKernelGGL((k_adjoint<SS, 1, (SS <= 4 ? RS : 1)>), grid, block, 0, st, a); \
        else                                                                              \
            hipLaunchKernelGGL((k_adjoint<SS, 0, (SS <= 4 ? RS : 1)>), grid, block, 0, st, a); \
        break;
    switch (d->S) {
        BP_ADJ_CASE(1)
        BP_ADJ_CASE(2)
        BP_ADJ_CASE(3)
        BP_ADJ_CASE(4)
        BP_ADJ_CASE(5)
        BP_ADJ_CASE(6)
        BP_ADJ_CASE(7)
        BP_ADJ_CASE(8)
        BP_ADJ_CASE(9)
        BP_ADJ_CASE(10)
    }
#undef BP_ADJ_CASE
}

static void enqueue_backward(bp_diff *x, hipStream_t st)
{
    size_t VS = (size_t)x->V * x->SP;
    unsigned K = x->K, c = x->c;
    unsigned nseg = (K + c - 1) / c;
    const float *gin = x->gin; /* dL/dst_K */
    const float *w_in = nullptr;
    int first = 1;
    std::vector<const float *> stp(c + 1), bufp(c + 1);
    for (unsigned k = nseg; k-- > 0;) {
        unsigned t0 = k * c, t1 = std::min(t0 + c, K);
        stp[0] = x->ck_st + (size_t)k * VS;
        bufp[0] = x->ck_buf + (size_t)k * VS;
        for (unsigned t = t0 + 1; t < t1; ++t) {
            float *so = x->seg_st + (size_t)(t - t0) * VS, *bo = x->seg_buf + (size_t)(t - t0) * VS;
            fwd_step(x, FWD, stp[t - 1 - t0], bufp[t - 1 - t0], so, bo, st);
            stp[t - t0] = so;
            bufp[t - t0] = bo;
        }
        stp[t1 - t0] = (t1 == K) ? x->final_st : x->ck_st + (size_t)(k + 1) * VS;
        for (unsigned t = t1; t > t0; --t) {
            float *gout = x->gst[(t - 1) & 1];
            float *wout = x->w[t & 1];
            adj_step(x, 0, stp[t - 1 - t0], stp[t - t0], bufp[t - 1 - t0], gin, gout, w_in, wout, first, st);
            gin = gout;
            w_in = wout;
            first = 0;
        }
    }
    /* step 0: init */
    adj_step(x, 1, nullptr, x->st0, x->st0, gin, x->dst0p, w_in, x->w[0] == w_in ? x->w[1] : x->w[0], first, st);
    const float *w0 = (x->w[0] == w_in) ? x->w[1] : x->w[0];
    dim3 gv(blocks_for(x->V)), b256(256);
#define BP_IG_CASE(SS)                                                                                     \
    case SS:                                                                                               \
        hipLaunchKernelGGL(k_initial_grad<SS>, gv, b256, 0, st, x->dst0p, w0, x->out_off, x->out_slot,     \
                           x->dst0, x->V);                                                                 \
        break;
    switch (x->S) {
        BP_IG_CASE(1)
        BP_IG_CASE(2)
        BP_IG_CASE(3)
        BP_IG_CASE(4)
        BP_IG_CASE(5)
        BP_IG_CASE(6)
        BP_IG_CASE(7)
        BP_IG_CASE(8)
        BP_IG_CASE(9)
        BP_IG_CASE(10)
    }
#undef BP_IG_CASE
    size_t nG = (size_t)x->nrec * x->S * x->S;
    if (nG)
        hipLaunchKernelGGL(k_parent_grad, dim3(blocks_for(nG)), dim3(256), 0, st, x->jbar, x->parents, x->G, x->nrec,
                           x->S);
}

/* The gradient of the parameters (log = 0: dL/dparams into out) or of the
 * log-parameters (log = 1: dL/dtheta, the softmax's transpose folded in),
 * from the per-parent sums of the last backward. */
static void launch_param_grad(bp_diff *x, int log, float *out, hipStream_t st)
{
    if (log && x->num_rows) {
        switch (x->S) {
#define BP_PGL_CASE(SS)                                                                                          \
    case SS:                                                                                                     \
        hipLaunchKernelGGL(k_param_grad_log<SS>, dim3(blocks_for(x->num_rows)), dim3(256), 0, st, x->G, x->dst0,  \
                           x->row_def_d, x->defs, x->P, out, x->num_rows, x->SP);                                \
        break;
            BP_PGL_CASE(1)
            BP_PGL_CASE(2)
            BP_PGL_CASE(3)
            BP_PGL_CASE(4)
            BP_PGL_CASE(5)
            BP_PGL_CASE(6)
            BP_PGL_CASE(7)
            BP_PGL_CASE(8)
            BP_PGL_CASE(9)
            BP_PGL_CASE(10)
#undef BP_PGL_CASE
        }
    } else if (!log && x->num_params) {
        switch (x->S) {
#define BP_PG_CASE(SS)                                                                                        \
    case SS:                                                                                                  \
        hipLaunchKernelGGL(k_param_grad<SS>, dim3(blocks_for(x->num_params)), dim3(256), 0, st, x->G, x->dst0, \
                           x->def_of, x->defs, out, x->num_params, x->SP);                                     \
        break;
            BP_PG_CASE(1)
            BP_PG_CASE(2)
            BP_PG_CASE(3)
            BP_PG_CASE(4)
            BP_PG_CASE(5)
            BP_PG_CASE(6)
            BP_PG_CASE(7)
            BP_PG_CASE(8)
            BP_PG_CASE(9)
            BP_PG_CASE(10)
#undef BP_PG_CASE
        }
    }
}

extern "C" int bp_diff_backward_ex(bp_diff *x, const float *grad_beliefs, float *grad_params, float *grad_tables,
                                   float *grad_initial, void *stream)
{
    bp_clear_error();
    if (!x || !grad_beliefs) {
        bp_set_error("bp_diff_backward: NULL argument");
        return -1;
    }
    if (!x->have_forward) {
        bp_set_error("bp_diff_backward: no forward since the parameters were set");
        return -1;
    }
    HIP_TRY(hipSetDevice(x->d->device));
    hipStream_t st = (hipStream_t)stream;
    size_t VS = (size_t)x->V * x->SP;
    hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for(VS)), dim3(256), 0, st, grad_beliefs, x->inv_d, x->gin, x->V,
                       x->SP);
    auto key = std::make_pair(x->K, x->c);
    auto it = x->bwd_graphs.find(key);
    if (it == x->bwd_graphs.end()) {
        hipGraphExec_t exec = nullptr;
        if (capture_begin(x))
            return -1;
        enqueue_backward(x, x->d->stream);
        if (capture_end(x, &exec))
            return -1;
        it = x->bwd_graphs.emplace(key, exec).first;
    }
    HIP_TRY(hipGraphLaunch(it->second, st));
    if (grad_params)
        launch_param_grad(x, 0, grad_params, st);
    size_t SS = (size_t)x->S * x->S;
    if (grad_tables && x->E)
        hipLaunchKernelGGL(k_slot_to_edge, dim3(blocks_for((size_t)x->E * SS)), dim3(256), 0, st, x->jbar,
                           x->slot_edge, grad_tables, x->E, (unsigned)SS);
    if (grad_initial)
        hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for(VS)), dim3(256), 0, st, x->dst0, x->perm_d, grad_initial,
                           x->V, x->SP);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int bp_diff_backward(bp_diff *x, const float *grad_beliefs, float *grad_params, void *stream)
{
    return bp_diff_backward_ex(x, grad_beliefs, grad_params, nullptr, nullptr, stream);
}

extern "C" int bp_diff_backward_log(bp_diff *x, const float *grad_beliefs, float *grad_theta, void *stream)
{
    if (!grad_theta) {
        bp_set_error("bp_diff_backward_log: NULL argument");
        return -1;
    }
    if (bp_diff_backward_ex(x, grad_beliefs, nullptr, nullptr, nullptr, stream))
        return -1;
    launch_param_grad(x, 1, grad_theta, (hipStream_t)stream);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int bp_diff_tables(bp_diff *x, float *out)
{
    bp_clear_error();
    if (!x || !out) {
        bp_set_error("bp_diff_tables: NULL argument");
        return -1;
    }
    HIP_TRY(hipSetDevice(x->d->device));
    HIP_TRY(hipDeviceSynchronize());
    size_t SS = (size_t)x->S * x->S;
    std::vector<float> tmp((size_t)x->E * SS);
    if (!tmp.empty())
        HIP_TRY(hipMemcpy(tmp.data(), x->d->J, tmp.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (unsigned s = 0; s < x->E; ++s)
        std::memcpy(out + (size_t)x->d->slot_edge[s] * SS, tmp.data() + (size_t)s * SS, SS * sizeof(float));
    return 0;
}
