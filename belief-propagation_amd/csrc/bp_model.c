/*
 * bp_model.c -- Bayes-net model -> Graph_t, and the XMLBIF writer.
 *
 * bp_graph_add_observation / bp_graph_add_cpt_edges restate the two passes
 * of the reference XMLBIF reader:
 *   pass 1, a DEFINITION without GIVEN: its TABLE (fewer than MAX_STATES
 *     entries) becomes the node's state through graph_set_node_state, which
 *     marks it observed (src/bnf-xml-parser/xml-expression.c:290-317);
 *   pass 2, a DEFINITION with GIVENs: for each parent, last GIVEN first, the
 *     pairwise table J[parent_state][child_state] is the sequential f32 sum
 *     of the TABLE entries over all configurations of the other parents,
 *     reading TABLE child-state-major (xml-expression.c:330-419).  The
 *     forward edge parent->child carries J; the reverse edge child->parent
 *     carries its transpose and is only added when the parent is not
 *     observed (xml-expression.c:402-411).
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bp_internal.h"
#include "../../include/bp_model.h"

int bp_graph_add_observation(Graph_t g, unsigned int node, const float *table, unsigned int table_len)
{
    float probabilities[MAX_STATES];
    if (table_len >= MAX_STATES) {
        /* xml-expression.c:306 asserts num_probabilities < MAX_STATES */
        bp_fail("observation for node %u has %u entries (limit %d)", node, table_len, MAX_STATES - 1);
        return -1;
    }
    for (unsigned int i = 0; i < MAX_STATES; ++i)
        probabilities[i] = 0.0f;
    for (unsigned int i = 0; i < table_len; ++i)
        probabilities[i] = table[i];
    graph_set_node_state(g, node, table_len, probabilities);
    return bp_has_error() ? -1 : 0;
}

int bp_graph_add_cpt_edges(Graph_t g, unsigned int child, const unsigned int *parents, unsigned int num_parents,
                           const float *table, uint64_t table_len)
{
    float forward[MAX_STATES * MAX_STATES];
    float reverse[MAX_STATES * MAX_STATES];
    unsigned int n_child = g->node_num_vars[child];
    if (n_child == 0) {
        bp_fail("definition for node %u: node has no states", child);
        return -1;
    }
    uint64_t slice = table_len / n_child;
    uint64_t offset = 1; /* product of the state counts of the parents handled so far */

    for (unsigned int pi = num_parents; pi-- > 0;) {
        unsigned int parent = parents[pi];
        unsigned int n_par = g->node_num_vars[parent];
        uint64_t stride = (uint64_t)n_par * offset;

        memset(forward, 0, sizeof(forward));
        memset(reverse, 0, sizeof(reverse));
        for (unsigned int k = 0; k < n_child; ++k) {
            for (unsigned int j = 0; j < n_par; ++j) {
                /* Blocks of `offset` consecutive entries where this parent is
                 * in state j; block 0 is taken when it starts inside the
                 * slice, later blocks when they end inside it. */
                float acc = forward[MAX_STATES * j + k];
                for (uint64_t b = 0;; ++b) {
                    uint64_t start = (uint64_t)j * offset + b * stride;
                    if (b == 0 ? start > slice : start + offset > slice)
                        break;
                    for (uint64_t idx = start; idx < start + offset; ++idx) {
                        uint64_t at = idx + (uint64_t)k * slice;
                        if (at >= table_len) {
                            bp_fail("definition for node %u: TABLE has %llu entries, too few for its parents",
                                    child, (unsigned long long)table_len);
                            return -1;
                        }
                        acc += table[at];
                    }
                }
                forward[MAX_STATES * j + k] = acc;
            }
        }
        for (unsigned int j = 0; j < n_par; ++j)
            for (unsigned int k = 0; k < n_child; ++k)
                reverse[MAX_STATES * k + j] = forward[MAX_STATES * j + k];

        graph_add_edge(g, parent, child, n_par, n_child, forward);
        if (bp_has_error())
            return -1;
        if (g->observed_nodes[parent] != 1) {
            graph_add_edge(g, child, parent, n_child, n_par, reverse);
            if (bp_has_error())
                return -1;
        }
        offset *= n_par;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* model container                                                           */

bp_bn *bp_bn_create(unsigned int num_vars)
{
    bp_bn *bn = (bp_bn *)calloc(1, sizeof(bp_bn));
    if (!bn)
        return NULL;
    bn->num_vars = num_vars;
    bn->names = (char *)calloc((size_t)(num_vars ? num_vars : 1), CHAR_BUFFER_SIZE);
    bn->num_states = (unsigned int *)calloc(num_vars ? num_vars : 1, sizeof(unsigned int));
    bn->def_cap = num_vars + 16;
    bn->def_for = (unsigned int *)malloc(sizeof(unsigned int) * bn->def_cap);
    bn->def_given_begin = (uint64_t *)malloc(sizeof(uint64_t) * (bn->def_cap + 1));
    bn->def_table_begin = (uint64_t *)malloc(sizeof(uint64_t) * (bn->def_cap + 1));
    bn->given_cap = 2ull * num_vars + 16;
    bn->def_given = (unsigned int *)malloc(sizeof(unsigned int) * bn->given_cap);
    bn->table_cap = 8ull * num_vars + 64;
    bn->def_table = (float *)malloc(sizeof(float) * bn->table_cap);
    if (!bn->names || !bn->num_states || !bn->def_for || !bn->def_given_begin || !bn->def_table_begin ||
        !bn->def_given || !bn->def_table) {
        bp_bn_free(bn);
        return NULL;
    }
    bn->def_given_begin[0] = 0;
    bn->def_table_begin[0] = 0;
    return bn;
}

void bp_bn_free(bp_bn *bn)
{
    if (!bn)
        return;
    free(bn->names);
    free(bn->num_states);
    free(bn->def_for);
    free(bn->def_given_begin);
    free(bn->def_given);
    free(bn->def_table_begin);
    free(bn->def_table);
    free(bn);
}

int bp_bn_set_var(bp_bn *bn, unsigned int v, const char *name, unsigned int num_states)
{
    if (v >= bn->num_vars || num_states == 0 || num_states > MAX_STATES) {
        bp_fail("bp_bn_set_var: bad variable %u / %u states", v, num_states);
        return -1;
    }
    strncpy(&bn->names[(size_t)v * CHAR_BUFFER_SIZE], name, CHAR_BUFFER_SIZE);
    bn->num_states[v] = num_states;
    return 0;
}

static int grow(void **ptr, uint64_t *cap, uint64_t need, size_t elem)
{
    if (need <= *cap)
        return 0;
    uint64_t c = *cap ? *cap : 16;
    while (c < need)
        c *= 2;
    void *p = realloc(*ptr, (size_t)(c * elem));
    if (!p)
        return -1;
    *ptr = p;
    *cap = c;
    return 0;
}

int bp_bn_add_def(bp_bn *bn, unsigned int child, const unsigned int *parents, unsigned int num_parents,
                  const float *table, uint64_t table_len)
{
    uint64_t d = bn->num_defs;
    uint64_t g0 = bn->def_given_begin[d], t0 = bn->def_table_begin[d];
    if (d + 2 > bn->def_cap) {
        uint64_t cap = bn->def_cap, cap2 = bn->def_cap, cap3 = bn->def_cap;
        if (grow((void **)&bn->def_for, &cap, d + 2, sizeof(unsigned int)) ||
            grow((void **)&bn->def_given_begin, &cap2, d + 3, sizeof(uint64_t)) ||
            grow((void **)&bn->def_table_begin, &cap3, d + 3, sizeof(uint64_t))) {
            bp_fail("bp_bn_add_def: out of host memory");
            return -1;
        }
        bn->def_cap = cap;
    }
    if (grow((void **)&bn->def_given, &bn->given_cap, g0 + num_parents, sizeof(unsigned int)) ||
        grow((void **)&bn->def_table, &bn->table_cap, t0 + table_len, sizeof(float))) {
        bp_fail("bp_bn_add_def: out of host memory");
        return -1;
    }
    bn->def_for[d] = child;
    memcpy(bn->def_given + g0, parents, sizeof(unsigned int) * num_parents);
    memcpy(bn->def_table + t0, table, sizeof(float) * table_len);
    bn->def_given_begin[d + 1] = g0 + num_parents;
    bn->def_table_begin[d + 1] = t0 + table_len;
    bn->num_defs = (unsigned int)(d + 1);
    return 0;
}

void bp_bn_truncate_defs(bp_bn *bn, unsigned int keep)
{
    if (keep < bn->num_defs)
        bn->num_defs = keep;
}

Graph_t bp_bn_to_graph(const bp_bn *bn)
{
    bp_clear_error();
    uint64_t givens = bn->def_given_begin[bn->num_defs];
    if (2 * givens > 0xffffffffull) {
        bp_fail("bp_bn_to_graph: %llu GIVEN entries exceed the 32-bit edge index", (unsigned long long)givens);
        return NULL;
    }
    Graph_t g = create_graph(bn->num_vars, (unsigned int)(2 * givens));
    if (!g)
        return NULL;
    for (unsigned int v = 0; v < bn->num_vars; ++v) {
        graph_add_node(g, bn->num_states[v], &bn->names[(size_t)v * CHAR_BUFFER_SIZE]);
        if (bp_has_error())
            goto fail;
    }
    for (unsigned int d = 0; d < bn->num_defs; ++d) {
        if (bn->def_given_begin[d + 1] != bn->def_given_begin[d])
            continue;
        uint64_t t0 = bn->def_table_begin[d], t1 = bn->def_table_begin[d + 1];
        if (bp_graph_add_observation(g, bn->def_for[d], bn->def_table + t0, (unsigned int)(t1 - t0)))
            goto fail;
    }
    for (unsigned int d = 0; d < bn->num_defs; ++d) {
        uint64_t g0 = bn->def_given_begin[d], g1 = bn->def_given_begin[d + 1];
        if (g0 == g1)
            continue;
        uint64_t t0 = bn->def_table_begin[d], t1 = bn->def_table_begin[d + 1];
        if (bp_graph_add_cpt_edges(g, bn->def_for[d], bn->def_given + g0, (unsigned int)(g1 - g0),
                                   bn->def_table + t0, t1 - t0))
            goto fail;
    }
    return g;
fail:
    graph_destroy(g);
    return NULL;
}

/* ------------------------------------------------------------------------ */
/* XMLBIF 0.3 writer (the layout of src/benchmark_files/dog.xml)             */

static void write_escaped(FILE *f, const char *s, size_t max)
{
    for (size_t i = 0; i < max && s[i]; ++i) {
        switch (s[i]) {
        case '<': fputs("&lt;", f); break;
        case '>': fputs("&gt;", f); break;
        case '&': fputs("&amp;", f); break;
        default: fputc(s[i], f);
        }
    }
}

int bp_bn_write_xmlbif(const bp_bn *bn, const char *path)
{
    FILE *f = fopen(path, "w");
    if (!f) {
        bp_fail("bp_bn_write_xmlbif: cannot open %s: %s", path, strerror(errno));
        return -1;
    }
    setvbuf(f, NULL, _IOFBF, 1 << 20);
    fputs("<?xml version=\"1.0\"?>\n<BIF VERSION=\"0.3\">\n<NETWORK>\n<NAME>generated</NAME>\n", f);
    for (unsigned int v = 0; v < bn->num_vars; ++v) {
        fputs("<VARIABLE TYPE=\"nature\">\n<NAME>", f);
        write_escaped(f, &bn->names[(size_t)v * CHAR_BUFFER_SIZE], CHAR_BUFFER_SIZE);
        fputs("</NAME>\n", f);
        for (unsigned int s = 0; s < bn->num_states[v]; ++s)
            fprintf(f, "<OUTCOME>s%u</OUTCOME>\n", s);
        fputs("</VARIABLE>\n", f);
    }
    for (unsigned int d = 0; d < bn->num_defs; ++d) {
        fputs("<DEFINITION>\n<FOR>", f);
        write_escaped(f, &bn->names[(size_t)bn->def_for[d] * CHAR_BUFFER_SIZE], CHAR_BUFFER_SIZE);
        fputs("</FOR>\n", f);
        for (uint64_t i = bn->def_given_begin[d]; i < bn->def_given_begin[d + 1]; ++i) {
            fputs("<GIVEN>", f);
            write_escaped(f, &bn->names[(size_t)bn->def_given[i] * CHAR_BUFFER_SIZE], CHAR_BUFFER_SIZE);
            fputs("</GIVEN>\n", f);
        }
        fputs("<TABLE>", f);
        for (uint64_t i = bn->def_table_begin[d]; i < bn->def_table_begin[d + 1]; ++i)
            /* 9 significant digits round-trip every float through strtof */
            fprintf(f, i + 1 < bn->def_table_begin[d + 1] ? "%.9g " : "%.9g", (double)bn->def_table[i]);
        fputs("</TABLE>\n</DEFINITION>\n", f);
    }
    fputs("</NETWORK>\n</BIF>\n", f);
    if (fclose(f) != 0) {
        bp_fail("bp_bn_write_xmlbif: write to %s failed", path);
        return -1;
    }
    return 0;
}
