/*
 * graph.h -- drop-in replacement for the reference graph API
 * (src/graph/graph.h:22-133).
 *
 * The struct layout below is binary interface: reference callers read and
 * write its fields directly (e.g. src/c_benchmark/main.c:206,314 reads
 * num_levels / current_num_edges, src/example-belief-propagation/
 * example-belief-propagation.c:30-40 reads node_states and node_num_vars),
 * so every field keeps the reference's name, type and order.
 *
 * Construction, CSR, level assignment and the parsers run on the host and are
 * bit-exact with the reference.  Every propagation entry point
 * (init_previous_edge, loopy_propagate_*, propagate_using_levels*,
 * marginalize, send_message) executes on the GPU through the HIP kernels of
 * libbp_amd.so and writes its results back into these fields; there is no
 * host fallback -- without a usable GPU they report an error.
 *
 * Differences from the reference, all in cases where the reference aborts,
 * hangs or reads uninitialised memory:
 *   - per-node edge lists grow dynamically (the reference asserts fewer than
 *     MAX_DEGREE entries, src/graph/graph.c:184,201);
 *   - errors go through bp_last_error() and, unless bp_set_abort_on_error(0)
 *     was called, print the message and abort() like the reference's asserts;
 *   - init_levels_to_nodes() reports unreachable nodes instead of looping
 *     forever (src/graph/graph.c:815-822);
 *   - the API is reentrant (the reference keeps static key buffers,
 *     src/graph/graph.c:9);
 *   - all arrays are zero-initialised, which defines the reference's
 *     uninitialised reads for networks whose neighbouring nodes have
 *     different state counts (zero-extension).
 */
#ifndef BP_GRAPH_H_
#define BP_GRAPH_H_

#include "constants.h"

#ifdef __cplusplus
extern "C" {
#endif

struct hsearch_data; /* only pointers to it appear in the layout */

/* src/graph/graph.h:22-74 */
struct graph {
    unsigned int total_num_vertices;   /* capacity given to create_graph */
    unsigned int total_num_edges;      /* capacity given to create_graph */

    unsigned int current_num_vertices; /* nodes added so far; node index = insertion order */
    unsigned int current_num_edges;    /* directed message edges added so far */
    unsigned int max_degree;           /* max out-degree, set by set_up_src_nodes_to_edges */

    unsigned int *edges_src_index;     /* [E] */
    unsigned int *edges_dest_index;    /* [E] */
    unsigned int *edges_x_dim;         /* [E] state count of the source */
    unsigned int *edges_y_dim;         /* [E] state count of the destination */
    float *edges_joint_probabilities;  /* [E][MAX_STATES][MAX_STATES], J[src_state][dst_state] */

    float *edges_messages;             /* [E][MAX_STATES] */
    float *last_edges_messages;        /* [E][MAX_STATES] */

    float **current_edge_messages;     /* points at one of the two arrays above */
    float **previous_edge_messages;    /* points at the other one */

    float *node_states;                /* [V][MAX_STATES] beliefs */
    unsigned int *node_num_vars;       /* [V] state counts */

    unsigned int *src_nodes_to_edges_node_list;  /* [V] start of each node's out-edges */
    unsigned int *src_nodes_to_edges_edge_list;  /* [E] edge indices grouped by source */

    unsigned int *dest_nodes_to_edges_node_list; /* [V] start of each node's in-edges */
    unsigned int *dest_nodes_to_edges_edge_list; /* [E] edge indices grouped by destination */

    unsigned int *levels_to_nodes;     /* [2V]: [0,L) level starts, [V,2V) nodes in level order */
    unsigned int num_levels;

    int diameter;

    char *visited;                     /* [V] */
    char *node_names;                  /* [V][CHAR_BUFFER_SIZE] */

    char *variable_names;              /* [V][MAX_STATES][CHAR_BUFFER_SIZE] */

    char *observed_nodes;              /* [V] 1 when a state was set through the API */

    char graph_name[CHAR_BUFFER_SIZE];

    char node_hash_table_created;
    struct hsearch_data *node_hash_table;          /* unused by this library (always NULL) */

    struct hsearch_data *src_node_to_edge_table;   /* unused by this library (always NULL) */

    struct hsearch_data *dest_node_to_edge_table;  /* unused by this library (always NULL) */
    char edge_tables_created;
};
typedef struct graph *Graph_t;

/* Public in the reference header (src/graph/graph.h:76-79); kept for source compatibility. */
struct htable_entry {
    unsigned int indices[MAX_DEGREE];
    unsigned int count;
};

/* ---- construction (host, bit-exact) ------------------------------------ */
Graph_t create_graph(unsigned int num_vertices, unsigned int num_edges);          /* graph.c:12-70 */

void graph_add_node(Graph_t, unsigned int num_variables, const char *name);        /* graph.c:116-125 */
void graph_add_and_set_node_state(Graph_t, unsigned int, const char *, float *);    /* graph.c:128-136 */
void graph_set_node_state(Graph_t, unsigned int node, unsigned int num_states, float *state); /* graph.c:138-147 */

void graph_add_edge(Graph_t, unsigned int src, unsigned int dest, unsigned int dim_x,
                    unsigned int dim_y, float *joint_probabilities);                /* graph.c:149-211 */

void set_up_src_nodes_to_edges(Graph_t);                                            /* graph.c:249-283 */
void set_up_dest_nodes_to_edges(Graph_t);                                           /* graph.c:285-313 */
void init_levels_to_nodes(Graph_t);                                                 /* graph.c:806-827 */
void calculate_diameter(Graph_t);                                                   /* graph.c:608-723 */

void initialize_node(Graph_t, unsigned int node, unsigned int num_variables);        /* graph.c:72-79 */
void node_set_state(Graph_t, unsigned int node, unsigned int num_variables, float *state); /* graph.c:107-114 */

void init_edge(Graph_t, unsigned int edge, unsigned int src, unsigned int dest,
               unsigned int dim_x, unsigned int dim_y, float *joint_probabilities);  /* graph.c:81-105 */

/* One message on one edge, computed on the GPU (graph.c:426-448). */
void send_message(float *states, unsigned int offset, unsigned int edge_index,
                  float *edge_joint_probabilities, float *edge_messages,
                  unsigned int *edge_num_src, unsigned int *edge_num_dest);

void fill_in_node_table(Graph_t);            /* declared but never defined by the reference (graph.h:101) */
void fill_in_node_hash_table(Graph_t);       /* graph.c:213-229 */
unsigned int find_node_by_name(char *name, Graph_t);                                /* graph.c:231-246 */

int graph_vertex_count(Graph_t);
int graph_edge_count(Graph_t);

void graph_destroy(Graph_t);                                                        /* graph.c:323-374 */

/* Level helpers the reference exports without a prototype (graph.c:753,774). */
void fill_in_leaf_nodes_in_index(Graph_t, unsigned int *start_index, unsigned int *end_index,
                                 unsigned int max_count);
void visit_node(Graph_t, unsigned int buffer_index, unsigned int *end_index);

/* ---- tree BP (GPU) ------------------------------------------------------ */
void propagate_using_levels_start(Graph_t);                  /* graph.c:376-424 */
void propagate_using_levels(Graph_t, unsigned int level);    /* graph.c:524-538 */
void reset_visited(Graph_t);                                 /* graph.c:598-605 (host flags) */
void marginalize(Graph_t);                                   /* graph.c:588-596 */

/* ---- loopy BP (GPU) ----------------------------------------------------- */
void init_previous_edge(Graph_t);                            /* graph.c:725-751 */
void loopy_propagate_one_iteration(Graph_t);                 /* graph.c:1121-1183 */
unsigned int loopy_propagate_until(Graph_t, float convergence, unsigned int max_iterations); /* graph.c:1278-1324 */

/* Edge-centric variants (graph.c:1186-1276,1422,1527).  The reference's
 * versions are defective (they copy E*10 bytes instead of floats, combine
 * with races and normalise node 0 only); these run the edge-centric schedule
 * with defined semantics on the GPU. */
void loopy_propagate_edge_one_iteration(Graph_t);
unsigned int loopy_propagate_until_edge(Graph_t, float, unsigned int);
unsigned int loopy_progagate_until_acc(Graph_t, float convergence, unsigned int max_iterations);
unsigned int loopy_progagate_until_edge_acc(Graph_t, float, unsigned int);

/* ---- debug output (host) ------------------------------------------------- */
void print_node(Graph_t, unsigned int);
void print_edge(Graph_t, unsigned int);
void print_nodes(Graph_t);
void print_edges(Graph_t);
void print_src_nodes_to_edges(Graph_t);
void print_dest_nodes_to_edges(Graph_t);
void print_levels_to_nodes(Graph_t);

#ifdef __cplusplus
}
#endif

#endif /* BP_GRAPH_H_ */
