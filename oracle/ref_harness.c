/*
 * ref_harness.c -- TEST INFRASTRUCTURE.  Driver linked against the
 * reference's own sources (compiled in place from /root/reference by
 * oracle/Makefile into oracle/_ref/; nothing of the reference is copied
 * into this repository).  It builds a graph through the reference's parsers
 * or graph API, runs the reference's propagation and writes every array the
 * tests compare against into a small tagged binary file, which
 * tests/golden/make_golden.py turns into .npz fixtures.
 *
 *   ref_harness <source> <path|-> <action> [args] <out.bin>
 *     source: xml | bif | example | replay
 *       example: the 4-node network of
 *                src/example-belief-propagation/example-belief-propagation.c:17-70
 *       replay:  a structure file written by tests/golden/make_golden.py
 *                (nodes, states, observed flags, edges with tables), rebuilt
 *                through graph_add_node / graph_set_node_state /
 *                graph_add_edge in the original order
 *     action:
 *       struct              structure + CSR (+ levels)
 *       loopy N             init_previous_edge, then N x
 *                           loopy_propagate_one_iteration; states, both
 *                           message arrays and the f32 delta per iteration
 *       until CONV MAX      init_previous_edge + loopy_propagate_until
 *       loopyq N            like loopy, but only the deltas and the final arrays
 *       tree [LEVELS]       levels + the driver's up/down sweeps + marginalize;
 *                           LEVELS (u32 num_levels, u32[2V]) replaces the
 *                           reference's O(V^2) level pass for large trees
 *       edge N              init_previous_edge + N x loopy_propagate_edge_one_iteration
 *       time N REPS         wall-clock of init + N iterations (best of REPS)
 *
 * Built with -Dfloat=double (HARNESS_FP64) the graph core runs in double; only
 * the example / replay sources exist there (the parsers read with "%f").
 */
#define _GNU_SOURCE
#include <assert.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "graph/graph.h"
#ifndef HARNESS_FP64
#include "bnf-parser/expression.h"
#include "bnf-parser/Parser.h"
#include "bnf-parser/Lexer.h"
#include "bnf-xml-parser/xml-expression.h"
int yyparse(struct expression **expr, yyscan_t scanner);
#endif

/* `float` is the graph core's real type (double under -Dfloat=double);
 * _Float32 is always the 4-byte type of the files. */
typedef _Float32 f32_t;

static FILE *out;

static void put(const char *name, char type, const void *data, uint64_t count, size_t elem)
{
    uint32_t n = (uint32_t)strlen(name);
    fwrite(&n, 4, 1, out);
    fwrite(name, 1, n, out);
    fwrite(&type, 1, 1, out);
    fwrite(&count, 8, 1, out);
    fwrite(data, elem, count, out);
}

static void put_u32(const char *name, const unsigned int *d, uint64_t n) { put(name, 'u', d, n, 4); }

/* real arrays are written as f32 (float build) or f64 (double build) */
static void put_real(const char *name, const float *d, uint64_t n)
{
    put(name, sizeof(float) == 8 ? 'd' : 'f', d, n, sizeof(float));
}

/* padded [count][MAX_STATES] array with entries beyond the valid prefix zeroed */
static void put_padded(const char *name, const float *d, uint64_t count, const unsigned int *valid)
{
    float *tmp = calloc(count * MAX_STATES + 1, sizeof(float));
    for (uint64_t i = 0; i < count; ++i)
        for (unsigned int k = 0; k < valid[i] && k < MAX_STATES; ++k)
            tmp[i * MAX_STATES + k] = d[i * MAX_STATES + k];
    put_real(name, tmp, count * MAX_STATES);
    free(tmp);
}

static void dump_structure(Graph_t g)
{
    unsigned int V = g->current_num_vertices, E = g->current_num_edges;
    unsigned int meta[6] = {V, E, g->total_num_vertices, g->total_num_edges, g->max_degree, g->num_levels};
    put_u32("meta", meta, 6);
    put_u32("node_num_vars", g->node_num_vars, V);
    put_padded("node_states", g->node_states, V, g->node_num_vars);
    put("observed", 'b', g->observed_nodes, V, 1);
    put("node_names", 'b', g->node_names, (uint64_t)V * CHAR_BUFFER_SIZE, 1);
    put_u32("edges_src", g->edges_src_index, E);
    put_u32("edges_dst", g->edges_dest_index, E);
    put_u32("edges_x", g->edges_x_dim, E);
    put_u32("edges_y", g->edges_y_dim, E);
    float *J = calloc((uint64_t)E * MAX_STATES * MAX_STATES + 1, sizeof(float));
    for (unsigned int e = 0; e < E; ++e)
        for (unsigned int i = 0; i < g->edges_x_dim[e]; ++i)
            for (unsigned int j = 0; j < g->edges_y_dim[e]; ++j)
                J[(uint64_t)e * 100 + i * 10 + j] = g->edges_joint_probabilities[(uint64_t)e * 100 + i * 10 + j];
    put_real("edges_J", J, (uint64_t)E * MAX_STATES * MAX_STATES);
    free(J);
    put_u32("src_node_list", g->src_nodes_to_edges_node_list, V);
    put_u32("src_edge_list", g->src_nodes_to_edges_edge_list, E);
    put_u32("dst_node_list", g->dest_nodes_to_edges_node_list, V);
    put_u32("dst_edge_list", g->dest_nodes_to_edges_edge_list, E);
}

static void dump_levels(Graph_t g)
{
    unsigned int V = g->current_num_vertices;
    put_u32("levels_to_nodes", g->levels_to_nodes, 2 * (uint64_t)V);
    put_u32("num_levels", &g->num_levels, 1);
}

static void dump_messages(Graph_t g, const char *cur_name, const char *prev_name)
{
    put_padded(cur_name, *g->current_edge_messages, g->current_num_edges, g->edges_x_dim);
    put_padded(prev_name, *g->previous_edge_messages, g->current_num_edges, g->edges_x_dim);
}

static float harness_delta(Graph_t g)
{
    /* graph.c:1300-1309, sequential */
    float d = 0.0;
    for (unsigned int e = 0; e < g->current_num_edges; ++e)
        for (unsigned int k = 0; k < g->edges_x_dim[e]; ++k) {
            float q = (*g->previous_edge_messages)[e * MAX_STATES + k] - (*g->current_edge_messages)[e * MAX_STATES + k];
            if (q != q)
                q = 0.0;
            d += fabs(q);
        }
    return d;
}

static Graph_t build_example(void)
{
    /* example-belief-propagation.c:17-70 */
    Graph_t g = create_graph(4, 5);
    float y2[2] = {1.0, 0.0};
    graph_add_node(g, 2, "x1");
    graph_add_node(g, 2, "x2");
    graph_add_node(g, 2, "x3");
    graph_add_and_set_node_state(g, 2, "y2", y2);
    float a[MAX_STATES * MAX_STATES], b[MAX_STATES * MAX_STATES], c[MAX_STATES * MAX_STATES];
    a[0] = 1.0; a[1] = 0.9; a[10] = 0.9; a[11] = 1.0;
    b[0] = 0.1; b[1] = 1.0; b[10] = 1.0; b[11] = 0.1;
    c[0] = 1.0; c[1] = 0.1; c[10] = 0.1; c[11] = 1.0;
    graph_add_edge(g, 0, 1, 2, 2, a);
    graph_add_edge(g, 1, 0, 2, 2, a);
    graph_add_edge(g, 1, 2, 2, 2, b);
    graph_add_edge(g, 2, 1, 2, 2, b);
    graph_add_edge(g, 3, 1, 2, 2, c);
    return g;
}

/* replay file: u32 V, u32 E, then per node u32 n, u8 observed, f32[n] states,
 * char[50] name; then per edge u32 src, dst, x, y, f32[x*y] table */
static Graph_t build_replay(const char *path)
{
    FILE *f = fopen(path, "rb");
    assert(f);
    unsigned int V, E;
    assert(fread(&V, 4, 1, f) == 1 && fread(&E, 4, 1, f) == 1);
    Graph_t g = create_graph(V, E);
    float st[MAX_STATES];
    for (unsigned int v = 0; v < V; ++v) {
        unsigned int n;
        unsigned char obs;
        f32_t raw[MAX_STATES];
        char name[CHAR_BUFFER_SIZE];
        assert(fread(&n, 4, 1, f) == 1 && fread(&obs, 1, 1, f) == 1);
        assert(n <= MAX_STATES && fread(raw, 4, n, f) == n);
        assert(fread(name, 1, CHAR_BUFFER_SIZE, f) == CHAR_BUFFER_SIZE);
        graph_add_node(g, n, name);
        if (obs) {
            for (unsigned int k = 0; k < n; ++k)
                st[k] = (float)raw[k];
            graph_set_node_state(g, v, n, st);
        }
    }
    float *J = malloc(sizeof(float) * MAX_STATES * MAX_STATES);
    for (unsigned int e = 0; e < E; ++e) {
        unsigned int hdr[4];
        f32_t raw[MAX_STATES * MAX_STATES];
        assert(fread(hdr, 4, 4, f) == 4);
        assert(fread(raw, 4, hdr[2] * hdr[3], f) == hdr[2] * hdr[3]);
        for (unsigned int i = 0; i < hdr[2]; ++i)
            for (unsigned int j = 0; j < hdr[3]; ++j)
                J[i * MAX_STATES + j] = (float)raw[i * hdr[3] + j];
        graph_add_edge(g, hdr[0], hdr[1], hdr[2], hdr[3], J);
    }
    free(J);
    fclose(f);
    return g;
}

#ifndef HARNESS_FP64
static Graph_t build_bif(const char *path)
{
    /* the parse sequence of src/bnf-parser/main.c:58-90 */
    struct expression *expression = NULL;
    yyscan_t scanner;
    FILE *in = fopen(path, "r");
    assert(in);
    assert(yylex_init(&scanner) == 0);
    yyset_in(in, scanner);
    assert(yyparse(&expression, scanner) == 0);
    yylex_destroy(scanner);
    fclose(in);
    assert(expression != NULL);
    Graph_t g = build_graph(expression);
    delete_expression(expression);
    return g;
}
#endif

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s <xml|bif|example|replay> <path|-> <action> [args] <out>\n", argv[0]);
        return 2;
    }
    const char *source = argv[1], *path = argv[2], *action = argv[3];
    out = fopen(argv[argc - 1], "wb");
    assert(out);
    Graph_t g = NULL;
    if (strcmp(source, "example") == 0)
        g = build_example();
    else if (strcmp(source, "replay") == 0)
        g = build_replay(path);
#ifndef HARNESS_FP64
    else if (strcmp(source, "xml") == 0)
        g = parse_xml_file(path);
    else if (strcmp(source, "bif") == 0)
        g = build_bif(path);
#endif
    assert(g);
    set_up_src_nodes_to_edges(g);
    set_up_dest_nodes_to_edges(g);

    if (strcmp(action, "struct") == 0) {
        dump_structure(g);
        if (argc > 5 && strcmp(argv[4], "levels") == 0) {
            init_levels_to_nodes(g);
            dump_levels(g);
        }
    } else if (strcmp(action, "loopy") == 0) {
        unsigned int n = (unsigned int)atoi(argv[4]);
        dump_structure(g);
        init_previous_edge(g);
        put_padded("init_prev", *g->previous_edge_messages, g->current_num_edges, g->edges_x_dim);
        float *deltas = calloc(n + 1, sizeof(float));
        char name[64];
        for (unsigned int it = 0; it < n; ++it) {
            loopy_propagate_one_iteration(g);
            deltas[it] = harness_delta(g);
            snprintf(name, sizeof(name), "states_%u", it + 1);
            put_padded(name, g->node_states, g->current_num_vertices, g->node_num_vars);
            char cur[64], prev[64];
            snprintf(cur, sizeof(cur), "cur_%u", it + 1);
            snprintf(prev, sizeof(prev), "prev_%u", it + 1);
            dump_messages(g, cur, prev);
        }
        put_real("deltas", deltas, n);
        free(deltas);
    } else if (strcmp(action, "loopyq") == 0) {
        unsigned int n = (unsigned int)atoi(argv[4]);
        init_previous_edge(g);
        float *deltas = calloc(n + 1, sizeof(float));
        for (unsigned int it = 0; it < n; ++it) {
            loopy_propagate_one_iteration(g);
            deltas[it] = harness_delta(g);
        }
        put_real("deltas", deltas, n);
        put_padded("states", g->node_states, g->current_num_vertices, g->node_num_vars);
        dump_messages(g, "cur", "prev");
        free(deltas);
    } else if (strcmp(action, "edge") == 0) {
        /* the reference's edge-centric iteration, defects included */
        unsigned int n = (unsigned int)atoi(argv[4]);
        init_previous_edge(g);
        for (unsigned int it = 0; it < n; ++it)
            loopy_propagate_edge_one_iteration(g);
        put_padded("states", g->node_states, g->current_num_vertices, g->node_num_vars);
        put_padded("cur", *g->current_edge_messages, g->current_num_edges, g->edges_x_dim);
    } else if (strcmp(action, "until") == 0) {
        float conv = (float)atof(argv[4]);
        unsigned int max = (unsigned int)atoi(argv[5]);
        dump_structure(g);
        init_previous_edge(g);
        unsigned int it = loopy_propagate_until(g, conv, max);
        put_u32("iterations", &it, 1);
        put_padded("states", g->node_states, g->current_num_vertices, g->node_num_vars);
        dump_messages(g, "cur", "prev");
    } else if (strcmp(action, "tree") == 0) {
        if (argc > 5 && strcmp(argv[4], "reflevels") == 0) {
            /* the reference's own level scan (graph.c:806-827), with no
             * structure dump: used on inputs too large to dump whole */
            init_levels_to_nodes(g);
        } else if (argc > 5) {
            FILE *lf = fopen(argv[4], "rb");
            assert(lf);
            assert(fread(&g->num_levels, 4, 1, lf) == 1);
            assert(fread(g->levels_to_nodes, 4, 2 * (size_t)g->current_num_vertices, lf) ==
                   2 * (size_t)g->current_num_vertices);
            fclose(lf);
        } else {
            dump_structure(g);
            init_levels_to_nodes(g);
        }
        dump_levels(g);
        /* src/c_benchmark/main.c:205-215 */
        propagate_using_levels_start(g);
        for (unsigned int i = 1; i < g->num_levels - 1; ++i)
            propagate_using_levels(g, i);
        reset_visited(g);
        for (unsigned int i = g->num_levels - 1; i > 0; --i)
            propagate_using_levels(g, i);
        put_padded("tree_messages", g->edges_messages, g->current_num_edges, g->edges_x_dim);
        marginalize(g);
        put_padded("states", g->node_states, g->current_num_vertices, g->node_num_vars);
    } else if (strcmp(action, "time") == 0) {
        unsigned int n = (unsigned int)atoi(argv[4]), reps = (unsigned int)atoi(argv[5]);
        double best = 1e30, best_iter = 1e30;
        float *s0 = malloc(sizeof(float) * (uint64_t)g->current_num_vertices * MAX_STATES);
        memcpy(s0, g->node_states, sizeof(float) * (uint64_t)g->current_num_vertices * MAX_STATES);
        for (unsigned int r = 0; r < reps; ++r) {
            memcpy(g->node_states, s0, sizeof(float) * (uint64_t)g->current_num_vertices * MAX_STATES);
            double t0 = now();
            init_previous_edge(g);
            double t1 = now();
            for (unsigned int it = 0; it < n; ++it)
                loopy_propagate_one_iteration(g);
            double t2 = now();
            if (t2 - t0 < best)
                best = t2 - t0;
            if (t2 - t1 < best_iter)
                best_iter = t2 - t1;
        }
        double times[2] = {best, best_iter};
        put("seconds", 'd', times, 2, 8);
        put_padded("states", g->node_states, g->current_num_vertices, g->node_num_vars);
        free(s0);
    } else {
        fprintf(stderr, "unknown action %s\n", action);
        return 2;
    }
    fclose(out);
    graph_destroy(g);
    return 0;
}
