/*
 * bp_model.h -- Bayes-net model, seeded workload generators and XMLBIF
 * writer of libbp_amd.so.
 *
 * A bp_bn holds what an XMLBIF file holds: variables (name, state count,
 * outcome names) and definitions (FOR, GIVEN list, TABLE).  bp_bn_to_graph()
 * turns it into a Graph_t with exactly the steps parse_xml_file() takes
 * (src/bnf-xml-parser/xml-expression.c:421-466), so a generated network
 * built in memory and the same network written by bp_bn_write_xmlbif() and
 * read back by parse_xml_file() give bit-identical graphs.  Tables are read
 * child-state-major, TABLE[k*slice + parent_configuration] with the last
 * GIVEN varying fastest (xml-expression.c:393).
 *
 * This interface has no counterpart in the reference: the reference ships a
 * Python 2 generator (generate_belief_network.py) whose distribution is not
 * the one these workloads need.
 */
#ifndef BP_MODEL_H_
#define BP_MODEL_H_

#include <stdint.h>

#include "graph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bp_bn {
    unsigned int num_vars;
    char *names;            /* [num_vars][CHAR_BUFFER_SIZE] */
    unsigned int *num_states; /* [num_vars] */

    unsigned int num_defs;
    unsigned int *def_for;          /* [num_defs] variable index */
    uint64_t *def_given_begin;      /* [num_defs+1] offsets into def_given */
    unsigned int *def_given;        /* parents in GIVEN order */
    uint64_t *def_table_begin;      /* [num_defs+1] offsets into def_table */
    float *def_table;               /* child-state-major tables */

    uint64_t given_cap, table_cap, def_cap; /* growth bookkeeping */
} bp_bn;

/* Workload families. */
enum bp_workload {
    BP_WL_DAG = 0,      /* random DAG, parents drawn from lower-indexed nodes */
    BP_WL_TREE = 1,     /* random recursive tree */
    BP_WL_LATTICE = 2,  /* R x C grid, parents (r,c-1) and (r-1,c) */
};

/* Parameters of a seeded network.  See bp_generate.c for the exact draws. */
typedef struct bp_gen_params {
    int kind;                    /* enum bp_workload */
    unsigned int num_nodes;      /* DAG / tree node count */
    unsigned int rows, cols;     /* lattice */
    int indegree_dist;           /* 0: uniform {indeg_lo..indeg_hi}; 1: min(Poisson(indeg_mean), indeg_hi) */
    unsigned int indeg_lo, indeg_hi;
    double indeg_mean;
    unsigned int states_lo, states_hi; /* node state counts uniform in {lo..hi} */
    int cpt_dist;                /* 0: U(0,1) normalised; 1: Dirichlet(alpha); 2: softmax of N(0,1) */
    double alpha;
    double evidence_fraction;    /* floor(f*V) one-hot observed nodes */
    unsigned int max_children;   /* cap per node so edge lists stay below MAX_DEGREE */
    uint64_t seed;
} bp_gen_params;

/* Fills *p with the defaults of a named workload: "c1", "c2", "c2b", "tree",
 * "c3", "c5".  Returns 0, or -1 for an unknown name. */
int bp_gen_preset(const char *name, bp_gen_params *p);

/* Generates a network; NULL on failure (bp_last_error()). */
bp_bn *bp_bn_generate(const bp_gen_params *p);

/* Evidence set: a second seeded one-hot evidence draw appended as extra
 * GIVEN-less definitions (used for batches of evidence sets). */
int bp_bn_add_evidence(bp_bn *bn, double fraction, uint64_t seed);

/* Removes all definitions appended after the first `keep` ones. */
void bp_bn_truncate_defs(bp_bn *bn, unsigned int keep);

Graph_t bp_bn_to_graph(const bp_bn *bn);
int bp_bn_write_xmlbif(const bp_bn *bn, const char *path);
void bp_bn_free(bp_bn *bn);

/* Builders used by bp_bn_generate and tests. */
bp_bn *bp_bn_create(unsigned int num_vars);
int bp_bn_set_var(bp_bn *bn, unsigned int v, const char *name, unsigned int num_states);
int bp_bn_add_def(bp_bn *bn, unsigned int child, const unsigned int *parents, unsigned int num_parents,
                  const float *table, uint64_t table_len);

/* The two passes of the XMLBIF reader, on an existing graph; shared by
 * parse_xml_file() and bp_bn_to_graph().  Return 0 or -1. */
int bp_graph_add_observation(Graph_t g, unsigned int node, const float *table, unsigned int table_len);
int bp_graph_add_cpt_edges(Graph_t g, unsigned int child, const unsigned int *parents, unsigned int num_parents,
                           const float *table, uint64_t table_len);

#ifdef __cplusplus
}
#endif

#endif /* BP_MODEL_H_ */
