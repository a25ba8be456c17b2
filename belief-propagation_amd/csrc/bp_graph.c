/*
 * bp_graph.c -- host-side graph construction, CSR and level assignment.
 *
 * Semantics follow src/graph/graph.c of the reference and are bit-exact with
 * it (field by field); the data structures behind them are this library's
 * own:
 *   - per-node edge lists are not kept while edges are added; the CSR views
 *     are built by a stable counting sort over the edge arrays, which yields
 *     the reference's per-node insertion order (graph.c:162-208,249-313);
 *   - node names are looked up in an open-addressing table where the first
 *     node inserted under a name wins, like hsearch_r ENTER (graph.c:213-246);
 *   - levels are assigned in O(V+E) with an in-list flag instead of the
 *     reference's linear membership scan (graph.c:774-827), visiting nodes in
 *     the same order and producing the same levels_to_nodes array.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bp_internal.h"

/* ------------------------------------------------------------------------ */
/* allocation                                                                */

static void *bp_zalloc(size_t count, size_t size)
{
    if (count == 0)
        count = 1; /* keep every array dereferenceable */
    return calloc(count, size);
}

Graph_t create_graph(unsigned int num_vertices, unsigned int num_edges)
{
    struct bp_graph_impl *impl = (struct bp_graph_impl *)calloc(1, sizeof(*impl));
    if (!impl) {
        bp_fail("create_graph: out of host memory");
        return NULL;
    }
    Graph_t g = &impl->pub;
    size_t V = num_vertices, E = num_edges;

    g->edges_src_index = (unsigned int *)bp_zalloc(E, sizeof(unsigned int));
    g->edges_dest_index = (unsigned int *)bp_zalloc(E, sizeof(unsigned int));
    g->edges_x_dim = (unsigned int *)bp_zalloc(E, sizeof(unsigned int));
    g->edges_y_dim = (unsigned int *)bp_zalloc(E, sizeof(unsigned int));
    g->edges_joint_probabilities = (float *)bp_zalloc(E * MAX_STATES * MAX_STATES, sizeof(float));
    g->edges_messages = (float *)bp_zalloc(E * MAX_STATES, sizeof(float));
    g->last_edges_messages = (float *)bp_zalloc(E * MAX_STATES, sizeof(float));
    g->node_states = (float *)bp_zalloc(V * MAX_STATES, sizeof(float));
    g->node_num_vars = (unsigned int *)bp_zalloc(V, sizeof(unsigned int));
    g->src_nodes_to_edges_node_list = (unsigned int *)bp_zalloc(V, sizeof(unsigned int));
    g->src_nodes_to_edges_edge_list = (unsigned int *)bp_zalloc(E, sizeof(unsigned int));
    g->dest_nodes_to_edges_node_list = (unsigned int *)bp_zalloc(V, sizeof(unsigned int));
    g->dest_nodes_to_edges_edge_list = (unsigned int *)bp_zalloc(E, sizeof(unsigned int));
    g->node_names = (char *)bp_zalloc(V * CHAR_BUFFER_SIZE, 1);
    g->visited = (char *)bp_zalloc(V, 1);
    g->observed_nodes = (char *)bp_zalloc(V, 1);
    g->variable_names = (char *)bp_zalloc(V * CHAR_BUFFER_SIZE * MAX_STATES, 1);
    g->levels_to_nodes = (unsigned int *)bp_zalloc(2 * V, sizeof(unsigned int));

    if (!g->edges_src_index || !g->edges_dest_index || !g->edges_x_dim || !g->edges_y_dim ||
        !g->edges_joint_probabilities || !g->edges_messages || !g->last_edges_messages ||
        !g->node_states || !g->node_num_vars || !g->src_nodes_to_edges_node_list ||
        !g->src_nodes_to_edges_edge_list || !g->dest_nodes_to_edges_node_list ||
        !g->dest_nodes_to_edges_edge_list || !g->node_names || !g->visited ||
        !g->observed_nodes || !g->variable_names || !g->levels_to_nodes) {
        graph_destroy(g);
        bp_fail("create_graph: out of host memory for %u nodes / %u edges", num_vertices, num_edges);
        return NULL;
    }

    g->current_edge_messages = &g->edges_messages;
    g->previous_edge_messages = &g->last_edges_messages;
    g->node_hash_table_created = 0;
    g->edge_tables_created = 0;
    g->num_levels = 0;
    g->total_num_vertices = num_vertices;
    g->total_num_edges = num_edges;
    g->current_num_vertices = 0;
    g->current_num_edges = 0;
    g->diameter = -1;
    g->max_degree = 0;
    return g;
}

void graph_destroy(Graph_t g)
{
    if (!g)
        return;
    struct bp_graph_impl *impl = bp_impl(g);
    bp_dev_cache_release(impl);
    free(impl->name_slots);
    free(g->edges_src_index);
    free(g->edges_dest_index);
    free(g->edges_x_dim);
    free(g->edges_y_dim);
    free(g->edges_joint_probabilities);
    free(g->edges_messages);
    free(g->last_edges_messages);
    free(g->src_nodes_to_edges_node_list);
    free(g->src_nodes_to_edges_edge_list);
    free(g->dest_nodes_to_edges_node_list);
    free(g->dest_nodes_to_edges_edge_list);
    free(g->node_names);
    free(g->visited);
    free(g->observed_nodes);
    free(g->variable_names);
    free(g->levels_to_nodes);
    free(g->node_num_vars);
    free(g->node_states);
    free(impl);
}

/* ------------------------------------------------------------------------ */
/* nodes                                                                     */

void initialize_node(Graph_t g, unsigned int node_index, unsigned int num_variables)
{
    if (num_variables > MAX_STATES) {
        bp_fail("initialize_node: %u states exceed MAX_STATES=%d", num_variables, MAX_STATES);
        return;
    }
    for (unsigned int i = 0; i < num_variables; ++i)
        g->node_states[(size_t)node_index * MAX_STATES + i] = (float)DEFAULT_STATE;
    g->node_num_vars[node_index] = num_variables;
}

void node_set_state(Graph_t g, unsigned int node_index, unsigned int num_variables, float *state)
{
    if (num_variables > MAX_STATES) {
        bp_fail("node_set_state: %u states exceed MAX_STATES=%d", num_variables, MAX_STATES);
        return;
    }
    for (unsigned int i = 0; i < num_variables; ++i)
        g->node_states[(size_t)node_index * MAX_STATES + i] = state[i];
}

void graph_add_node(Graph_t g, unsigned int num_variables, const char *name)
{
    unsigned int v = g->current_num_vertices;
    if (v >= g->total_num_vertices) {
        bp_fail("graph_add_node: graph was created for %u nodes", g->total_num_vertices);
        return;
    }
    if (num_variables > MAX_STATES) {
        bp_fail("graph_add_node: %u states exceed MAX_STATES=%d", num_variables, MAX_STATES);
        return;
    }
    initialize_node(g, v, num_variables);
    /* strncpy semantics: zero padded, not terminated at 50 characters */
    strncpy(&g->node_names[(size_t)v * CHAR_BUFFER_SIZE], name, CHAR_BUFFER_SIZE);
    g->current_num_vertices = v + 1;
    bp_impl(g)->generation++;
}

void graph_add_and_set_node_state(Graph_t g, unsigned int num_variables, const char *name, float *state)
{
    unsigned int v = g->current_num_vertices;
    if (v >= g->total_num_vertices) {
        bp_fail("graph_add_and_set_node_state: graph was created for %u nodes", g->total_num_vertices);
        return;
    }
    g->observed_nodes[v] = 1;
    graph_add_node(g, num_variables, name);
    node_set_state(g, v, num_variables, state);
}

void graph_set_node_state(Graph_t g, unsigned int node_index, unsigned int num_states, float *state)
{
    if (node_index >= g->current_num_vertices) {
        bp_fail("graph_set_node_state: node %u does not exist (%u nodes)", node_index, g->current_num_vertices);
        return;
    }
    if (num_states > g->node_num_vars[node_index]) {
        bp_fail("graph_set_node_state: %u states given for node %u with %u states", num_states, node_index,
                g->node_num_vars[node_index]);
        return;
    }
    g->observed_nodes[node_index] = 1;
    node_set_state(g, node_index, num_states, state);
}

/* ------------------------------------------------------------------------ */
/* edges                                                                     */

void init_edge(Graph_t g, unsigned int edge_index, unsigned int src_index, unsigned int dest_index,
               unsigned int dim_x, unsigned int dim_y, float *joint_probabilities)
{
    if (dim_x > MAX_STATES || dim_y > MAX_STATES) {
        bp_fail("init_edge: table %ux%u exceeds MAX_STATES=%d", dim_x, dim_y, MAX_STATES);
        return;
    }
    g->edges_src_index[edge_index] = src_index;
    g->edges_dest_index[edge_index] = dest_index;
    g->edges_x_dim[edge_index] = dim_x;
    g->edges_y_dim[edge_index] = dim_y;
    float *J = &g->edges_joint_probabilities[(size_t)edge_index * MAX_STATES * MAX_STATES];
    float *m = &g->edges_messages[(size_t)edge_index * MAX_STATES];
    float *p = &g->last_edges_messages[(size_t)edge_index * MAX_STATES];
    for (unsigned int i = 0; i < dim_x; ++i) {
        /* the caller's table is MAX_STATES-strided (graph.c:99-102) */
        for (unsigned int j = 0; j < dim_y; ++j)
            J[MAX_STATES * i + j] = joint_probabilities[MAX_STATES * i + j];
        m[i] = 0.0f;
        p[i] = 0.0f;
    }
}

void graph_add_edge(Graph_t g, unsigned int src_index, unsigned int dest_index, unsigned int dim_x,
                    unsigned int dim_y, float *joint_probabilities)
{
    unsigned int e = g->current_num_edges;
    if (e >= g->total_num_edges) {
        bp_fail("graph_add_edge: graph was created for %u edges", g->total_num_edges);
        return;
    }
    if (src_index >= g->current_num_vertices || dest_index >= g->current_num_vertices) {
        bp_fail("graph_add_edge: edge %u->%u names a node that does not exist (%u nodes)", src_index,
                dest_index, g->current_num_vertices);
        return;
    }
    if (g->node_num_vars[src_index] != dim_x || g->node_num_vars[dest_index] != dim_y) {
        bp_fail("graph_add_edge: table %ux%u does not match state counts %u (node %u) and %u (node %u)",
                dim_x, dim_y, g->node_num_vars[src_index], src_index, g->node_num_vars[dest_index], dest_index);
        return;
    }
    init_edge(g, e, src_index, dest_index, dim_x, dim_y, joint_probabilities);
    g->edge_tables_created = 1;
    g->current_num_edges = e + 1;
    bp_impl(g)->generation++;
}

int graph_vertex_count(Graph_t g) { return (int)g->current_num_vertices; }
int graph_edge_count(Graph_t g) { return (int)g->current_num_edges; }

/* ------------------------------------------------------------------------ */
/* CSR views: stable counting sort by source / destination                  */

static void build_csr(Graph_t g, const unsigned int *key, unsigned int *node_list, unsigned int *edge_list,
                      unsigned int *max_count)
{
    unsigned int V = g->total_num_vertices, E = g->current_num_edges;
    for (unsigned int v = 0; v < V; ++v)
        node_list[v] = 0;
    for (unsigned int e = 0; e < E; ++e)
        node_list[key[e]]++;
    unsigned int running = 0;
    for (unsigned int v = 0; v < V; ++v) {
        unsigned int c = node_list[v];
        if (max_count && c > *max_count)
            *max_count = c;
        node_list[v] = running;
        running += c;
    }
    /* second pass scatters in edge-index (= insertion) order; node_list is
     * used as a cursor and restored afterwards */
    for (unsigned int e = 0; e < E; ++e)
        edge_list[node_list[key[e]]++] = e;
    for (unsigned int v = V; v-- > 0;)
        node_list[v] = (v == 0) ? 0 : node_list[v - 1];
}

void set_up_src_nodes_to_edges(Graph_t g)
{
    if (g->current_num_vertices != g->total_num_vertices) {
        bp_fail("set_up_src_nodes_to_edges: %u of %u nodes added", g->current_num_vertices,
                g->total_num_vertices);
        return;
    }
    /* max_degree only grows, as in graph.c:279-281 */
    build_csr(g, g->edges_src_index, g->src_nodes_to_edges_node_list, g->src_nodes_to_edges_edge_list,
              &g->max_degree);
    bp_impl(g)->generation++;
}

void set_up_dest_nodes_to_edges(Graph_t g)
{
    if (g->current_num_vertices != g->total_num_vertices) {
        bp_fail("set_up_dest_nodes_to_edges: %u of %u nodes added", g->current_num_vertices,
                g->total_num_vertices);
        return;
    }
    build_csr(g, g->edges_dest_index, g->dest_nodes_to_edges_node_list, g->dest_nodes_to_edges_edge_list, NULL);
    bp_impl(g)->generation++;
}

/* ------------------------------------------------------------------------ */
/* names                                                                     */

static uint64_t name_hash(const char *s)
{
    uint64_t h = 1469598103934665603ull; /* FNV-1a */
    size_t n = strnlen(s, CHAR_BUFFER_SIZE);
    for (size_t i = 0; i < n; ++i) {
        h ^= (unsigned char)s[i];
        h *= 1099511628211ull;
    }
    return h;
}

static int name_equal(const char *a, const char *b)
{
    return strncmp(a, b, CHAR_BUFFER_SIZE) == 0;
}

void fill_in_node_hash_table(Graph_t g)
{
    struct bp_graph_impl *impl = bp_impl(g);
    unsigned int V = g->current_num_vertices;
    if (impl->name_slots && impl->name_count == V)
        return;
    unsigned int cap = 16;
    while (cap < 2u * V + 2u)
        cap <<= 1;
    unsigned int *slots = (unsigned int *)malloc(sizeof(unsigned int) * cap);
    if (!slots) {
        bp_fail("fill_in_node_hash_table: out of host memory");
        return;
    }
    memset(slots, 0xff, sizeof(unsigned int) * cap);
    for (unsigned int v = 0; v < V; ++v) {
        const char *name = &g->node_names[(size_t)v * CHAR_BUFFER_SIZE];
        unsigned int pos = (unsigned int)name_hash(name) & (cap - 1);
        for (;;) {
            unsigned int cur = slots[pos];
            if (cur == 0xffffffffu) {
                slots[pos] = v;
                break;
            }
            if (name_equal(&g->node_names[(size_t)cur * CHAR_BUFFER_SIZE], name))
                break; /* first node with this name keeps it */
            pos = (pos + 1) & (cap - 1);
        }
    }
    free(impl->name_slots);
    impl->name_slots = slots;
    impl->name_cap = cap;
    impl->name_count = V;
    g->node_hash_table_created = 1;
}

void fill_in_node_table(Graph_t g) { fill_in_node_hash_table(g); }

unsigned int find_node_by_name(char *name, Graph_t g)
{
    struct bp_graph_impl *impl = bp_impl(g);
    fill_in_node_hash_table(g);
    if (!impl->name_slots) {
        return 0;
    }
    char key[CHAR_BUFFER_SIZE + 1];
    strncpy(key, name, CHAR_BUFFER_SIZE);
    key[CHAR_BUFFER_SIZE] = '\0';
    unsigned int pos = (unsigned int)name_hash(key) & (impl->name_cap - 1);
    for (;;) {
        unsigned int cur = impl->name_slots[pos];
        if (cur == 0xffffffffu)
            break;
        if (name_equal(&g->node_names[(size_t)cur * CHAR_BUFFER_SIZE], key))
            return cur;
        pos = (pos + 1) & (impl->name_cap - 1);
    }
    bp_fail("find_node_by_name: no node named '%s'", key);
    return 0xffffffffu;
}

/* ------------------------------------------------------------------------ */
/* levels (graph.c:753-827)                                                  */

void reset_visited(Graph_t g)
{
    memset(g->visited, 0, g->current_num_vertices);
}

void fill_in_leaf_nodes_in_index(Graph_t g, unsigned int *start_index, unsigned int *end_index,
                                 unsigned int max_count)
{
    g->levels_to_nodes[0] = *start_index;
    for (unsigned int v = 0; v < g->current_num_vertices; ++v) {
        unsigned int begin = g->dest_nodes_to_edges_node_list[v];
        unsigned int end = bp_csr_end(g, g->dest_nodes_to_edges_node_list, v);
        if (end - begin <= max_count)
            g->levels_to_nodes[(*end_index)++] = v;
    }
}

/* Standalone form with the reference's membership scan over
 * levels_to_nodes[V, *end_index); init_levels_to_nodes() uses a flag array. */
void visit_node(Graph_t g, unsigned int buffer_index, unsigned int *end_index)
{
    unsigned int v = g->levels_to_nodes[buffer_index];
    if (g->visited[v])
        return;
    g->visited[v] = 1;
    unsigned int begin = g->src_nodes_to_edges_node_list[v];
    unsigned int end = bp_csr_end(g, g->src_nodes_to_edges_node_list, v);
    for (unsigned int i = begin; i < end; ++i) {
        unsigned int w = g->edges_dest_index[g->src_nodes_to_edges_edge_list[i]];
        int listed = 0;
        for (unsigned int j = g->current_num_vertices; j < *end_index; ++j) {
            if (g->levels_to_nodes[j] == w) {
                listed = 1;
                break;
            }
        }
        if (!listed)
            g->levels_to_nodes[(*end_index)++] = w;
    }
}

void init_levels_to_nodes(Graph_t g)
{
    unsigned int V = g->current_num_vertices;
    unsigned int *L = g->levels_to_nodes;
    unsigned char *listed = (unsigned char *)calloc(V ? V : 1, 1);
    if (!listed) {
        bp_fail("init_levels_to_nodes: out of host memory");
        return;
    }
    reset_visited(g);
    g->num_levels = 0;

    unsigned int start = V, end = V;
    fill_in_leaf_nodes_in_index(g, &start, &end, 2);
    for (unsigned int i = V; i < end; ++i)
        listed[L[i]] = 1;

    /* Each round of the reference walks [V, end) and expands the nodes not
     * yet visited; those are exactly the ones appended by the previous
     * round, so this walks [frontier, copy_end). */
    unsigned int frontier = V;
    while (end < 2 * V) {
        unsigned int copy_end = end;
        for (unsigned int i = frontier; i < copy_end; ++i) {
            unsigned int v = L[i];
            if (g->visited[v])
                continue;
            g->visited[v] = 1;
            unsigned int begin = g->src_nodes_to_edges_node_list[v];
            unsigned int stop = bp_csr_end(g, g->src_nodes_to_edges_node_list, v);
            for (unsigned int k = begin; k < stop; ++k) {
                unsigned int w = g->edges_dest_index[g->src_nodes_to_edges_edge_list[k]];
                if (!listed[w]) {
                    listed[w] = 1;
                    L[end++] = w;
                }
            }
        }
        if (end == copy_end) {
            /* the reference loops forever here (graph.c:815-822) */
            free(listed);
            reset_visited(g);
            bp_fail("init_levels_to_nodes: %u of %u nodes are unreachable from the level-0 nodes", 2 * V - end, V);
            return;
        }
        frontier = copy_end;
        g->num_levels += 1;
        L[g->num_levels] = copy_end;
    }
    g->num_levels += 1;
    free(listed);
    reset_visited(g);
    bp_impl(g)->generation++;
}

/* ------------------------------------------------------------------------ */
/* diameter: largest finite shortest-path length over ordered pairs,         */
/* including i == j through a cycle, as the Floyd-Warshall of graph.c:1549   */
/* computes; here one BFS per source.                                        */

void calculate_diameter(Graph_t g)
{
    unsigned int V = g->current_num_vertices;
    int *dist = (int *)malloc(sizeof(int) * (V ? V : 1));
    unsigned int *queue = (unsigned int *)malloc(sizeof(unsigned int) * (V ? V : 1));
    if (!dist || !queue) {
        free(dist);
        free(queue);
        bp_fail("calculate_diameter: out of host memory");
        return;
    }
    int best = -1;
    for (unsigned int s = 0; s < V; ++s) {
        for (unsigned int v = 0; v < V; ++v)
            dist[v] = -1;
        unsigned int head = 0, tail = 0;
        int self = -1;
        /* seed with the direct successors of s (paths of at least one edge) */
        unsigned int b = g->src_nodes_to_edges_node_list[s], e = bp_csr_end(g, g->src_nodes_to_edges_node_list, s);
        for (unsigned int k = b; k < e; ++k) {
            unsigned int w = g->edges_dest_index[g->src_nodes_to_edges_edge_list[k]];
            if (dist[w] < 0) {
                dist[w] = 1;
                queue[tail++] = w;
            }
        }
        while (head < tail) {
            unsigned int u = queue[head++];
            if (u == s && self < 0)
                self = dist[u];
            unsigned int ub = g->src_nodes_to_edges_node_list[u], ue = bp_csr_end(g, g->src_nodes_to_edges_node_list, u);
            for (unsigned int k = ub; k < ue; ++k) {
                unsigned int w = g->edges_dest_index[g->src_nodes_to_edges_edge_list[k]];
                if (dist[w] < 0) {
                    dist[w] = dist[u] + 1;
                    queue[tail++] = w;
                }
            }
        }
        for (unsigned int v = 0; v < V; ++v)
            if (dist[v] > best && dist[v] < WEIGHT_INFINITY)
                best = dist[v];
    }
    g->diameter = best;
    free(dist);
    free(queue);
}

/* ------------------------------------------------------------------------ */
/* debug output                                                              */

void print_node(Graph_t g, unsigned int v)
{
    unsigned int n = g->node_num_vars[v];
    printf("Node %.*s [\n", CHAR_BUFFER_SIZE, &g->node_names[(size_t)v * CHAR_BUFFER_SIZE]);
    for (unsigned int i = 0; i < n; ++i)
        printf("%.*s:\t%.6lf\n", CHAR_BUFFER_SIZE,
               &g->variable_names[(size_t)v * CHAR_BUFFER_SIZE * MAX_STATES + (size_t)i * CHAR_BUFFER_SIZE],
               (double)g->node_states[(size_t)MAX_STATES * v + i]);
    printf("]\n");
}

void print_edge(Graph_t g, unsigned int e)
{
    unsigned int x = g->edges_x_dim[e], y = g->edges_y_dim[e];
    unsigned int s = g->edges_src_index[e], d = g->edges_dest_index[e];
    printf("Edge  %.*s -> %.*s [\n", CHAR_BUFFER_SIZE, &g->node_names[(size_t)s * CHAR_BUFFER_SIZE],
           CHAR_BUFFER_SIZE, &g->node_names[(size_t)d * CHAR_BUFFER_SIZE]);
    printf("Joint probability matrix: [\n");
    for (unsigned int i = 0; i < x; ++i) {
        printf("[");
        for (unsigned int j = 0; j < y; ++j)
            printf("\t%.6lf", (double)g->edges_joint_probabilities[(size_t)e * MAX_STATES * MAX_STATES + MAX_STATES * i + j]);
        printf("\t]\n");
    }
    printf("]\nMessage:\n[");
    for (unsigned int i = 0; i < x; ++i)
        printf("\t%.6lf", (double)g->edges_messages[(size_t)e * MAX_STATES + i]);
    printf("\t]\n]\n");
}

void print_nodes(Graph_t g)
{
    for (unsigned int v = 0; v < g->current_num_vertices; ++v)
        print_node(g, v);
}

void print_edges(Graph_t g)
{
    for (unsigned int e = 0; e < g->current_num_edges; ++e)
        print_edge(g, e);
}

static void print_csr(Graph_t g, const char *title, const unsigned int *node_list, const unsigned int *edge_list)
{
    printf("%s\n", title);
    for (unsigned int v = 0; v < g->total_num_vertices; ++v) {
        printf("Node -----\n");
        print_node(g, v);
        printf("Edges-------\n");
        for (unsigned int k = node_list[v]; k < bp_csr_end(g, node_list, v); ++k)
            print_edge(g, edge_list[k]);
        printf("---------\n");
    }
}

void print_src_nodes_to_edges(Graph_t g)
{
    print_csr(g, "src index -> edge index", g->src_nodes_to_edges_node_list, g->src_nodes_to_edges_edge_list);
}

void print_dest_nodes_to_edges(Graph_t g)
{
    print_csr(g, "dest index -> edge index", g->dest_nodes_to_edges_node_list, g->dest_nodes_to_edges_edge_list);
}

void print_levels_to_nodes(Graph_t g)
{
    for (unsigned int l = 0; l < g->num_levels; ++l) {
        printf("Level: %u\n", l);
        printf("---------------\n");
        unsigned int b = g->levels_to_nodes[l];
        unsigned int e = (l + 1 == g->num_levels) ? 2 * g->current_num_vertices : g->levels_to_nodes[l + 1];
        printf("Nodes:-----------\n");
        for (unsigned int k = b; k < e; ++k)
            print_node(g, g->levels_to_nodes[k]);
        printf("-------------------\n");
    }
}
