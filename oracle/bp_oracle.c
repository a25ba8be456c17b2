/*
 * bp_oracle.c -- TEST INFRASTRUCTURE: the CPU oracle for belief propagation.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * liboracle.so, and only as the checker (or the timed CPU baseline) -- the
 * product (libbp_amd.so) never links or calls it.
 *
 * It restates the reference's CPU algorithms (src/graph/graph.c) in plain C:
 *   orc_init_previous_edge       graph.c:725-751
 *   orc_loopy_iteration[_omp]    graph.c:1121-1183 (node loops :1153, :984)
 *   orc_loopy_until              graph.c:1278-1324 (serial delta, see below)
 *   orc_tree_start / _level      graph.c:376-538
 *   orc_tree_marginalize         graph.c:540-596
 *   orc_tree_run                 src/c_benchmark/main.c:205-215
 * in float on the Graph_t arrays themselves (bit-exact with the reference on
 * networks where neighbouring nodes share a state count; zero-extension
 * otherwise) and in double on a widened copy ("orc64_*", the FP64 oracle that
 * the GPU's FP32 beliefs are measured against, 1e-5 max-abs).
 *
 * Pinning (each checked in tests/test_oracle.py):
 *   - golden vectors written by the compiled reference (tests/golden/), per
 *     iteration for loopy BP and after the full sweep for tree BP, float and
 *     -Dfloat=double builds;
 *   - the reference's own known-answer test
 *     (src/example-belief-propagation/example-belief-propagation.c:72-107);
 *   - closed forms: one edge m = N(J.b); a node without in-edges keeps its
 *     normalised state; an all-zero product gives a zero message;
 *   - OpenMP vs serial bit-identity.
 * The delta differs from the reference's OpenMP build on purpose: that build
 * multiplies it by the thread count (graph.c:1300, a `parallel` without
 * `for`); orc_loopy_until sums once, like the serial build.
 */
#include <stdlib.h>
#include <string.h>

#include "../include/graph.h"

#pragma GCC diagnostic ignored "-Wunused-function"

#define R float
#define FN(x) orc_f32_##x
#include "bp_oracle_core.h"
#undef R
#undef FN
#define R double
#define FN(x) orc_f64_##x
#include "bp_oracle_core.h"
#undef R
#undef FN

static orc_f32_view f32_view(Graph_t g)
{
    orc_f32_view v;
    v.V = g->current_num_vertices;
    v.E = g->current_num_edges;
    v.num_vars = g->node_num_vars;
    v.src = g->edges_src_index;
    v.dst = g->edges_dest_index;
    v.xdim = g->edges_x_dim;
    v.ydim = g->edges_y_dim;
    v.src_nodes = g->src_nodes_to_edges_node_list;
    v.src_edges = g->src_nodes_to_edges_edge_list;
    v.dst_nodes = g->dest_nodes_to_edges_node_list;
    v.dst_edges = g->dest_nodes_to_edges_edge_list;
    v.levels = g->levels_to_nodes;
    v.num_levels = g->num_levels;
    v.visited = g->visited;
    v.states = g->node_states;
    v.J = g->edges_joint_probabilities;
    v.cur = g->current_edge_messages;
    v.prev = g->previous_edge_messages;
    v.tree = g->edges_messages;
    return v;
}

/* The pointer swap of an iteration is written back into the Graph_t. */
static void f32_commit(Graph_t g, const orc_f32_view *v)
{
    g->current_edge_messages = v->cur;
    g->previous_edge_messages = v->prev;
}

void orc_init_previous_edge(Graph_t g)
{
    orc_f32_view v = f32_view(g);
    orc_f32_init_previous_edge(&v);
}

void orc_loopy_iteration(Graph_t g)
{
    orc_f32_view v = f32_view(g);
    orc_f32_loopy_iteration(&v, 1);
    f32_commit(g, &v);
}

void orc_loopy_iteration_omp(Graph_t g, int threads)
{
    orc_f32_view v = f32_view(g);
    orc_f32_loopy_iteration(&v, threads);
    f32_commit(g, &v);
}

unsigned int orc_loopy_until(Graph_t g, float convergence, unsigned int max_iterations, int threads, float *deltas)
{
    orc_f32_view v = f32_view(g);
    unsigned int it = orc_f32_loopy_until(&v, convergence, max_iterations, threads, deltas);
    f32_commit(g, &v);
    return it;
}

void orc_edge_iteration(Graph_t g)
{
    orc_f32_view v = f32_view(g);
    orc_f32_edge_iteration(&v);
}

float orc_delta(Graph_t g)
{
    orc_f32_view v = f32_view(g);
    return orc_f32_delta(&v, *v.prev, *v.cur);
}

double orc_delta64(Graph_t g)
{
    orc_f32_view v = f32_view(g);
    return orc_f32_delta64(&v, *v.prev, *v.cur);
}

void orc_send_message(const float *b, unsigned int nb, Graph_t g, unsigned int edge, float *out)
{
    orc_f32_view v = f32_view(g);
    orc_f32_send(&v, b, nb, edge, out);
}

void orc_tree_start(Graph_t g)
{
    orc_f32_view v = f32_view(g);
    orc_f32_tree_start(&v);
}

void orc_tree_level(Graph_t g, unsigned int level)
{
    orc_f32_view v = f32_view(g);
    orc_f32_tree_level(&v, level);
}

void orc_tree_marginalize(Graph_t g)
{
    orc_f32_view v = f32_view(g);
    orc_f32_tree_marginalise(&v);
}

void orc_tree_run(Graph_t g)
{
    orc_f32_view v = f32_view(g);
    orc_f32_tree_run(&v);
}

/* ---- FP64 oracle on a widened copy ---------------------------------- */

typedef struct orc64 {
    orc_f64_view v;
    double *states, *J, *a, *b, *tree;
    double *cur_slot, *prev_slot; /* targets of v.cur / v.prev */
    char *visited;
} orc64;

void orc64_free(orc64 *o)
{
    if (!o)
        return;
    free(o->states);
    free(o->J);
    free(o->a);
    free(o->b);
    free(o->tree);
    free(o->visited);
    free(o);
}

/* Widens node states, tables and both message arrays; the structure arrays
 * are shared with g, which must outlive the copy. */
orc64 *orc64_create(Graph_t g)
{
    orc64 *o = (orc64 *)calloc(1, sizeof(orc64));
    if (!o)
        return NULL;
    size_t V = g->current_num_vertices, E = g->current_num_edges;
    o->states = (double *)calloc(V * MAX_STATES + 1, sizeof(double));
    o->J = (double *)calloc(E * MAX_STATES * MAX_STATES + 1, sizeof(double));
    o->a = (double *)calloc(E * MAX_STATES + 1, sizeof(double));
    o->b = (double *)calloc(E * MAX_STATES + 1, sizeof(double));
    o->tree = (double *)calloc(E * MAX_STATES + 1, sizeof(double));
    o->visited = (char *)calloc(V + 1, 1);
    if (!o->states || !o->J || !o->a || !o->b || !o->tree || !o->visited) {
        orc64_free(o);
        return NULL;
    }
    for (size_t i = 0; i < V * MAX_STATES; ++i)
        o->states[i] = g->node_states[i];
    for (size_t i = 0; i < E * MAX_STATES * MAX_STATES; ++i)
        o->J[i] = g->edges_joint_probabilities[i];
    for (size_t i = 0; i < E * MAX_STATES; ++i) {
        o->a[i] = (*g->current_edge_messages)[i];
        o->b[i] = (*g->previous_edge_messages)[i];
        o->tree[i] = g->edges_messages[i];
    }
    orc_f32_view f = f32_view(g);
    o->v.V = f.V;
    o->v.E = f.E;
    o->v.num_vars = f.num_vars;
    o->v.src = f.src;
    o->v.dst = f.dst;
    o->v.xdim = f.xdim;
    o->v.ydim = f.ydim;
    o->v.src_nodes = f.src_nodes;
    o->v.src_edges = f.src_edges;
    o->v.dst_nodes = f.dst_nodes;
    o->v.dst_edges = f.dst_edges;
    o->v.levels = f.levels;
    o->v.num_levels = f.num_levels;
    o->v.visited = o->visited;
    o->v.states = o->states;
    o->v.J = o->J;
    o->cur_slot = o->a;
    o->prev_slot = o->b;
    o->v.cur = &o->cur_slot;
    o->v.prev = &o->prev_slot;
    o->v.tree = o->tree;
    return o;
}

void orc64_init_previous_edge(orc64 *o) { orc_f64_init_previous_edge(&o->v); }
void orc64_loopy_iteration(orc64 *o, int threads) { orc_f64_loopy_iteration(&o->v, threads); }
void orc64_tree_run(orc64 *o) { orc_f64_tree_run(&o->v); }

/* copies out [V][MAX_STATES] states, or the current / previous messages */
void orc64_get_states(const orc64 *o, double *out)
{
    memcpy(out, o->states, sizeof(double) * o->v.V * MAX_STATES);
}

void orc64_get_messages(const orc64 *o, int previous, double *out)
{
    memcpy(out, previous ? *o->v.prev : *o->v.cur, sizeof(double) * o->v.E * MAX_STATES);
}

void orc64_get_tree_messages(const orc64 *o, double *out)
{
    memcpy(out, o->tree, sizeof(double) * o->v.E * MAX_STATES);
}
