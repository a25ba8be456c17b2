/*
 * bp_oracle_core.h -- TEST INFRASTRUCTURE (the oracle), never part of the
 * product.  Included twice by bp_oracle.c, with R = float and R = double.
 *
 * A plain, sequential restatement of the reference's CPU belief propagation
 * (src/graph/graph.c), operation for operation in the same order so that the
 * float instantiation reproduces the reference bit for bit:
 *   N(v)        s = 0; s = s + v[i] left to right; s <= 0 -> 1; v[i] = v[i]/s
 *                                                          (graph.c:440-447)
 *   send(b, e)  m[i] = sum_j J[i][j]*b[j] (j ascending, from 0), then N(m)
 *                                                          (graph.c:885-910)
 *   combine     buf[k] = buf[k]*m[k] unless m[k] is NaN     (graph.c:450-459)
 * Reads beyond a node's own state count -- which the reference leaves
 * uninitialised when neighbouring nodes have different state counts
 * (graph.c:856 vs :899, :103-104 vs :455) -- are defined here as 0.0
 * (zero-extension).
 *
 * ORC_PLANT_PAD_ONE, ORC_PLANT_STALE and ORC_PLANT_NORM_STRIDE build
 * deliberately wrong variants (out-of-range reads as 1.0; out-of-range reads
 * of whatever the padded row holds; normalisation over the whole
 * MAX_STATES-wide row).  liboracle.so never sets them: the variants exist so
 * the oracle's tests can show that they reject each of these mistakes.
 */

#ifndef ORC_OUT_OF_RANGE
#if defined(ORC_PLANT_PAD_ONE)
#define ORC_OUT_OF_RANGE(row, i) 1.0
#elif defined(ORC_PLANT_STALE)
#define ORC_OUT_OF_RANGE(row, i) (row)[i]
#else
#define ORC_OUT_OF_RANGE(row, i) 0.0
#endif
#endif

/* FN(x) names the instantiation; R is the real type. */

typedef struct FN(view) {
    unsigned int V, E;
    const unsigned int *num_vars;
    const unsigned int *src, *dst, *xdim, *ydim;
    const unsigned int *src_nodes, *src_edges; /* reference src CSR (no sentinel) */
    const unsigned int *dst_nodes, *dst_edges; /* reference dest CSR */
    const unsigned int *levels;                /* levels_to_nodes */
    unsigned int num_levels;
    char *visited;
    R *states; /* [V][MAX_STATES] */
    R *J;      /* [E][MAX_STATES][MAX_STATES] */
    R **cur;   /* current messages [E][MAX_STATES] */
    R **prev;  /* previous messages */
    R *tree;   /* tree BP messages (the reference's edges_messages) */
} FN(view);

static inline unsigned int FN(end)(const FN(view) *g, const unsigned int *nodes, unsigned int v)
{
    return (v + 1 >= g->V) ? g->E : nodes[v + 1];
}

static void FN(normalise)(R *v, unsigned int n)
{
#ifdef ORC_PLANT_NORM_STRIDE
    n = MAX_STATES;
#endif
    R s = (R)0.0;
    for (unsigned int i = 0; i < n; ++i)
        s = s + v[i];
    if (s <= (R)0.0)
        s = (R)1.0;
    for (unsigned int i = 0; i < n; ++i)
        v[i] = v[i] / s;
}

/* b has nb valid entries; entries up to y_e beyond them read as 0 */
static void FN(send)(const FN(view) *g, const R *b, unsigned int nb, unsigned int e, R *out)
{
    unsigned int x = g->xdim[e], y = g->ydim[e];
    const R *J = g->J + (size_t)e * MAX_STATES * MAX_STATES;
    R s = (R)0.0;
    for (unsigned int i = 0; i < x; ++i) {
        R acc = (R)0.0;
        for (unsigned int j = 0; j < y; ++j) {
            R bj = (j < nb) ? b[j] : (R)ORC_OUT_OF_RANGE(b, j);
            acc = acc + J[MAX_STATES * i + j] * bj;
        }
        out[i] = acc;
        s = s + acc;
    }
#ifdef ORC_PLANT_NORM_STRIDE
    for (unsigned int i = x; i < MAX_STATES; ++i)
        s = s + out[i];
#endif
    if (s <= (R)0.0)
        s = (R)1.0;
    for (unsigned int i = 0; i < x; ++i)
        out[i] = out[i] / s;
}

/* buf[k<n] *= msg[k] (NaN skipped); message entries k >= x read as 0 */
static inline void FN(combine)(R *buf, unsigned int n, const R *msg, unsigned int x)
{
    for (unsigned int k = 0; k < n; ++k) {
        R m = (k < x) ? msg[k] : (R)ORC_OUT_OF_RANGE(msg, k);
        if (m == m)
            buf[k] = buf[k] * m;
    }
}

/* graph.c:725-751 */
static void FN(init_previous_edge)(FN(view) *g)
{
    R *P = *g->prev;
    for (unsigned int v = 0; v < g->V; ++v) {
        for (unsigned int k = g->src_nodes[v]; k < FN(end)(g, g->src_nodes, v); ++k) {
            unsigned int e = g->src_edges[k];
            FN(send)(g, g->states + (size_t)MAX_STATES * v, g->num_vars[v], e, P + (size_t)MAX_STATES * e);
        }
    }
}

/* graph.c:1154-1175, one node: combine previous messages, send on out-edges */
static void FN(node_send)(FN(view) *g, unsigned int v, const R *P, R *C)
{
    R buf[MAX_STATES];
    unsigned int n = g->num_vars[v];
    for (unsigned int k = 0; k < n; ++k)
        buf[k] = g->states[(size_t)MAX_STATES * v + k];
    for (unsigned int k = g->dst_nodes[v]; k < FN(end)(g, g->dst_nodes, v); ++k) {
        unsigned int e = g->dst_edges[k];
        FN(combine)(buf, n, P + (size_t)MAX_STATES * e, g->xdim[e]);
    }
    for (unsigned int k = g->src_nodes[v]; k < FN(end)(g, g->src_nodes, v); ++k) {
        unsigned int e = g->src_edges[k];
        FN(send)(g, buf, n, e, C + (size_t)MAX_STATES * e);
    }
}

/* graph.c:985-1024, one node: multiply in the current messages, normalise */
static void FN(node_marginalise)(FN(view) *g, unsigned int v, const R *C)
{
    R t[MAX_STATES];
    unsigned int n = g->num_vars[v];
    R *s = g->states + (size_t)MAX_STATES * v;
    for (unsigned int k = 0; k < n; ++k)
        t[k] = (R)1.0;
    unsigned int b = g->dst_nodes[v], e_end = FN(end)(g, g->dst_nodes, v);
    for (unsigned int k = b; k < e_end; ++k) {
        unsigned int e = g->dst_edges[k];
        FN(combine)(t, n, C + (size_t)MAX_STATES * e, g->xdim[e]);
    }
    if (b < e_end)
        for (unsigned int k = 0; k < n; ++k)
            s[k] = s[k] * t[k];
    FN(normalise)(s, n);
}

/* graph.c:1121-1183; threads > 1 parallelises both node loops */
static void FN(loopy_iteration)(FN(view) *g, int threads)
{
    R *P = *g->prev, *C = *g->cur;
    long V = (long)g->V;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
#endif
    for (long v = 0; v < V; ++v)
        FN(node_send)(g, (unsigned int)v, P, C);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
#endif
    for (long v = 0; v < V; ++v)
        FN(node_marginalise)(g, (unsigned int)v, C);
    (void)threads;
    R **tmp = g->prev;
    g->prev = g->cur;
    g->cur = tmp;
}

/* graph.c:1300-1309 in R arithmetic, edge order */
static R FN(delta)(const FN(view) *g, const R *A, const R *B)
{
    R d = (R)0.0;
    for (unsigned int e = 0; e < g->E; ++e) {
        for (unsigned int k = 0; k < g->xdim[e]; ++k) {
            R q = A[(size_t)MAX_STATES * e + k] - B[(size_t)MAX_STATES * e + k];
            if (q != q)
                q = (R)0.0;
            d = d + (q < (R)0.0 ? -q : q);
        }
    }
    return d;
}

/* the same sum accumulated in double (the GPU's convergence delta) */
static double FN(delta64)(const FN(view) *g, const R *A, const R *B)
{
    double d = 0.0;
    for (unsigned int e = 0; e < g->E; ++e) {
        for (unsigned int k = 0; k < g->xdim[e]; ++k) {
            R q = A[(size_t)MAX_STATES * e + k] - B[(size_t)MAX_STATES * e + k];
            if (q != q)
                q = (R)0.0;
            d = d + (double)(q < (R)0.0 ? -q : q);
        }
    }
    return d;
}

/* graph.c:1278-1324 with the serial delta (one thread's sum, Q7 aside).
 * deltas (nullable) receives one entry per iteration run. */
static unsigned int FN(loopy_until)(FN(view) *g, R convergence, unsigned int max_iterations, int threads,
                                   R *deltas)
{
    R *A = *g->prev, *B = *g->cur; /* captured once, as graph.c:1285-1286 */
    R previous = (R)-1.0, d = (R)0.0;
    unsigned int i;
    for (i = 0; i < max_iterations; ++i) {
        FN(loopy_iteration)(g, threads);
        d = FN(delta)(g, A, B);
        if (deltas)
            deltas[i] = d;
        R change = d - previous;
        if (change < (R)0.0)
            change = -change;
        if (d < convergence || change < convergence)
            break;
        if (i < max_iterations - 1)
            previous = d;
    }
    return i;
}

/* ---- edge-centric schedule ------------------------------------------------
 * graph.c:1186-1276 with its defects fixed (it copies E*10 bytes instead of
 * floats, runs `parallel` without `for`, and normalises node 0 only):
 *   prev <- cur (all of it); cur[e] = send(states[src(e)], e) for every e;
 *   states[dst(e)][k] *= cur[e][k] for every e in edge order (no NaN skip,
 *   graph.c:1034-1043); every node divided by its sum when that is > 0
 *   (graph.c:1046-1061).  No pointer swap. */
static void FN(edge_iteration)(FN(view) *g)
{
    R *P = *g->prev, *Cm = *g->cur;
    memcpy(P, Cm, sizeof(R) * (size_t)g->E * MAX_STATES);
    for (unsigned int e = 0; e < g->E; ++e) {
        unsigned int u = g->src[e];
        FN(send)(g, g->states + (size_t)MAX_STATES * u, g->num_vars[u], e, Cm + (size_t)MAX_STATES * e);
    }
    for (unsigned int e = 0; e < g->E; ++e) {
        unsigned int v = g->dst[e], x = g->xdim[e], y = g->ydim[e];
        R *s = g->states + (size_t)MAX_STATES * v;
        for (unsigned int k = 0; k < y; ++k)
            s[k] = s[k] * ((k < x) ? Cm[(size_t)MAX_STATES * e + k] : (R)0.0);
    }
    for (unsigned int v = 0; v < g->V; ++v) {
        R *s = g->states + (size_t)MAX_STATES * v;
        R sum = (R)0.0;
        for (unsigned int k = 0; k < g->num_vars[v]; ++k)
            sum = sum + s[k];
        if (sum > (R)0.0)
            for (unsigned int k = 0; k < g->num_vars[v]; ++k)
                s[k] = s[k] / sum;
    }
}

/* ---- tree BP ---------------------------------------------------------- */

static unsigned int FN(level_end)(const FN(view) *g, unsigned int level)
{
    return (level + 1 == g->num_levels) ? 2 * g->V : g->levels[level + 1];
}

/* graph.c:376-424 */
static void FN(tree_start)(FN(view) *g)
{
    unsigned int b = g->levels[0], e = (g->num_levels == 1) ? 2 * g->V : g->levels[1];
    for (unsigned int k = b; k < e; ++k) {
        unsigned int v = g->levels[k];
        g->visited[v] = 1;
        for (unsigned int i = g->src_nodes[v]; i < FN(end)(g, g->src_nodes, v); ++i) {
            unsigned int edge = g->src_edges[i];
            FN(send)(g, g->states + (size_t)MAX_STATES * v, g->num_vars[v], edge, g->tree + (size_t)MAX_STATES * edge);
        }
    }
}

/* graph.c:461-538 */
static void FN(tree_level)(FN(view) *g, unsigned int level)
{
    R buf[MAX_STATES];
    for (unsigned int k = g->levels[level]; k < FN(level_end)(g, level); ++k) {
        unsigned int v = g->levels[k];
        unsigned int n = g->num_vars[v];
        g->visited[v] = 1;
        for (unsigned int i = 0; i < n; ++i)
            buf[i] = (R)1.0;
        for (unsigned int i = g->dst_nodes[v]; i < FN(end)(g, g->dst_nodes, v); ++i) {
            unsigned int edge = g->dst_edges[i];
            FN(combine)(buf, n, g->tree + (size_t)MAX_STATES * edge, g->xdim[edge]);
        }
        for (unsigned int i = g->src_nodes[v]; i < FN(end)(g, g->src_nodes, v); ++i) {
            unsigned int edge = g->src_edges[i];
            if (g->visited[g->dst[edge]] == 0)
                FN(send)(g, buf, n, edge, g->tree + (size_t)MAX_STATES * edge);
        }
    }
}

/* graph.c:540-596: the state is replaced by the product of the messages */
static void FN(tree_marginalise)(FN(view) *g)
{
    R t[MAX_STATES];
    for (unsigned int v = 0; v < g->V; ++v) {
        unsigned int n = g->num_vars[v];
        R *s = g->states + (size_t)MAX_STATES * v;
        for (unsigned int i = 0; i < n; ++i)
            t[i] = (R)1.0;
        unsigned int b = g->dst_nodes[v], e = FN(end)(g, g->dst_nodes, v);
        for (unsigned int i = b; i < e; ++i) {
            unsigned int edge = g->dst_edges[i];
            FN(combine)(t, n, g->tree + (size_t)MAX_STATES * edge, g->xdim[edge]);
        }
        if (b < e)
            for (unsigned int i = 0; i < n; ++i)
                s[i] = t[i];
        FN(normalise)(s, n);
    }
}

/* the driver sequence of src/c_benchmark/main.c:205-215 */
static void FN(tree_run)(FN(view) *g)
{
    memset(g->visited, 0, g->V);
    FN(tree_start)(g);
    for (unsigned int l = 1; l + 1 < g->num_levels; ++l)
        FN(tree_level)(g, l);
    memset(g->visited, 0, g->V);
    for (unsigned int l = g->num_levels - 1; l > 0; --l)
        FN(tree_level)(g, l);
    FN(tree_marginalise)(g);
}
