/*
 * bp_generate.c -- seeded synthetic Bayes nets for the loopy / tree BP
 * workloads.
 *
 * Random numbers: splitmix64 expands the seed into the state of
 * xoshiro256** (Blackman & Vigna); doubles take the top 53 bits.
 *
 * Structure:
 *   DAG      node i draws its in-degree (uniform {lo..hi} or
 *            min(Poisson(mean), hi)) and that many distinct parents uniformly
 *            from the lower-indexed nodes that still have fewer than
 *            max_children children; index order is a topological order.
 *   TREE     node i >= 1 draws one parent uniformly from the lower-indexed
 *            nodes with fewer than max_children children.
 *   LATTICE  node (r,c) = r*cols+c has parents (r,c-1) and (r-1,c), in that
 *            GIVEN order.
 * Parents are listed in ascending index order.  max_children keeps every
 * per-node edge list below MAX_DEGREE entries so the reference itself can
 * run the same network (src/graph/graph.c:184,201).
 *
 * Tables: one child distribution per parent configuration, stored
 * child-state-major with the last parent fastest (the layout the XMLBIF
 * reader expects, src/bnf-xml-parser/xml-expression.c:393).  Rows are
 * U(0,1) draws, Dirichlet(alpha) draws (Marsaglia-Tsang gamma sampler with
 * the U^(1/alpha) boost for alpha < 1) or softmax of N(0,1) draws
 * (Box-Muller), normalised in f32 by a sequential sum.  Roots get a prior
 * (a GIVEN-less definition, which marks them observed); then
 * floor(fraction*V) distinct nodes get one-hot evidence with a uniform
 * state, appended after all priors.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bp_internal.h"
#include "../../include/bp_model.h"

typedef struct {
    uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i)
        r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t rng_next(rng_t *r)
{
    uint64_t *s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

static double rng_unif(rng_t *r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }

/* uniform integer in [0, n) without modulo bias */
static uint64_t rng_below(rng_t *r, uint64_t n)
{
    if (n <= 1)
        return 0;
    uint64_t limit = UINT64_MAX - UINT64_MAX % n;
    for (;;) {
        uint64_t x = rng_next(r);
        if (x < limit)
            return x % n;
    }
}

static unsigned int rng_poisson(rng_t *r, double mean)
{
    double L = exp(-mean), p = 1.0;
    unsigned int k = 0;
    do {
        k++;
        p *= rng_unif(r);
    } while (p > L);
    return k - 1;
}

static double rng_normal(rng_t *r)
{
    double u1 = rng_unif(r), u2 = rng_unif(r);
    if (u1 < 1e-300)
        u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static double rng_gamma(rng_t *r, double alpha)
{
    if (alpha < 1.0) {
        double u = rng_unif(r);
        return rng_gamma(r, alpha + 1.0) * pow(u > 0 ? u : 1e-300, 1.0 / alpha);
    }
    double d = alpha - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    for (;;) {
        double x, v;
        do {
            x = rng_normal(r);
            v = 1.0 + c * x;
        } while (v <= 0.0);
        v = v * v * v;
        double u = rng_unif(r);
        if (u < 1.0 - 0.0331 * x * x * x * x)
            return d * v;
        if (log(u) < 0.5 * x * x + d * (1.0 - v + log(v)))
            return d * v;
    }
}

/* one child distribution of n entries, f32-normalised */
static void draw_row(rng_t *r, const bp_gen_params *p, unsigned int n, float *row)
{
    float sum = 0.0f;
    if (p->cpt_dist == 2) {
        double th[MAX_STATES], mx = -1e300;
        for (unsigned int k = 0; k < n; ++k) {
            th[k] = rng_normal(r);
            if (th[k] > mx)
                mx = th[k];
        }
        for (unsigned int k = 0; k < n; ++k)
            row[k] = (float)exp(th[k] - mx);
    } else {
        for (unsigned int k = 0; k < n; ++k)
            row[k] = (float)(p->cpt_dist == 1 ? rng_gamma(r, p->alpha) : rng_unif(r));
    }
    for (unsigned int k = 0; k < n; ++k)
        sum += row[k];
    if (!(sum > 0.0f)) {
        for (unsigned int k = 0; k < n; ++k)
            row[k] = 1.0f;
        sum = (float)n;
    }
    for (unsigned int k = 0; k < n; ++k)
        row[k] = row[k] / sum;
}

int bp_gen_preset(const char *name, bp_gen_params *p)
{
    memset(p, 0, sizeof(*p));
    p->seed = 20170205ull;
    p->alpha = 1.0;
    p->indeg_mean = 2.0;
    if (strcmp(name, "c1") == 0) {
        p->kind = BP_WL_DAG;
        p->num_nodes = 1000;
        p->indegree_dist = 0;
        p->indeg_lo = 0;
        p->indeg_hi = 3;
        p->states_lo = p->states_hi = 2;
        p->cpt_dist = 0;
        p->evidence_fraction = 0.05;
        p->max_children = 16;
        p->seed = 1001;
    } else if (strcmp(name, "c2") == 0 || strcmp(name, "c2b") == 0) {
        p->kind = BP_WL_DAG;
        p->num_nodes = 500000;
        p->indegree_dist = 1;
        p->indeg_mean = 2.0;
        p->indeg_hi = 6;
        p->states_lo = 2;
        p->states_hi = (strcmp(name, "c2") == 0) ? 4 : 2;
        p->cpt_dist = 0;
        p->evidence_fraction = 0.02;
        p->max_children = 13;
        p->seed = (strcmp(name, "c2") == 0) ? 2002 : 2003;
    } else if (strcmp(name, "tree") == 0) {
        p->kind = BP_WL_TREE;
        p->num_nodes = 1000000;
        p->states_lo = p->states_hi = 2;
        p->cpt_dist = 0;
        p->evidence_fraction = 0.0;
        p->max_children = 18;
        p->seed = 3003;
    } else if (strcmp(name, "c3") == 0) {
        p->kind = BP_WL_LATTICE;
        p->rows = p->cols = 1024;
        p->states_lo = p->states_hi = 4;
        p->cpt_dist = 1;
        p->alpha = 0.5;
        p->evidence_fraction = 0.01;
        p->max_children = 2;
        p->seed = 4004;
    } else if (strcmp(name, "c4net") == 0) {
        p->kind = BP_WL_DAG;
        p->num_nodes = 1000;
        p->indegree_dist = 0;
        p->indeg_lo = 0;
        p->indeg_hi = 4;
        p->states_lo = 2;
        p->states_hi = 8;
        p->cpt_dist = 0;
        p->evidence_fraction = 0.03;
        p->max_children = 15;
        p->seed = 5005;
    } else if (strcmp(name, "c5") == 0) {
        p->kind = BP_WL_DAG;
        p->num_nodes = 200000;
        p->indegree_dist = 1;
        p->indeg_mean = 2.0;
        p->indeg_hi = 6;
        p->states_lo = p->states_hi = 3;
        p->cpt_dist = 2;
        p->evidence_fraction = 0.05;
        p->max_children = 13;
        p->seed = 6006;
    } else {
        return -1;
    }
    return 0;
}

static int add_table_def(bp_bn *bn, rng_t *r, const bp_gen_params *p, unsigned int child,
                         const unsigned int *parents, unsigned int np, float *scratch, uint64_t scratch_len)
{
    unsigned int n = bn->num_states[child];
    uint64_t slice = 1;
    for (unsigned int i = 0; i < np; ++i)
        slice *= bn->num_states[parents[i]];
    uint64_t len = slice * n;
    if (len > scratch_len) {
        bp_fail("generator: table of %llu entries too large", (unsigned long long)len);
        return -1;
    }
    float row[MAX_STATES];
    for (uint64_t c = 0; c < slice; ++c) {
        draw_row(r, p, n, row);
        for (unsigned int k = 0; k < n; ++k)
            scratch[(uint64_t)k * slice + c] = row[k];
    }
    return bp_bn_add_def(bn, child, parents, np, scratch, len);
}

static int cmp_uint(const void *a, const void *b)
{
    unsigned int x = *(const unsigned int *)a, y = *(const unsigned int *)b;
    return (x > y) - (x < y);
}

int bp_bn_add_evidence(bp_bn *bn, double fraction, uint64_t seed)
{
    rng_t r;
    rng_seed(&r, seed ^ 0x5eed0e71de5cull);
    unsigned int V = bn->num_vars;
    unsigned int count = (unsigned int)floor(fraction * (double)V);
    if (count == 0)
        return 0;
    unsigned int *perm = (unsigned int *)malloc(sizeof(unsigned int) * V);
    if (!perm) {
        bp_fail("generator: out of host memory");
        return -1;
    }
    for (unsigned int v = 0; v < V; ++v)
        perm[v] = v;
    float onehot[MAX_STATES];
    int rc = 0;
    for (unsigned int i = 0; i < count; ++i) {
        unsigned int j = i + (unsigned int)rng_below(&r, V - i);
        unsigned int t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
        unsigned int v = perm[i];
        unsigned int n = bn->num_states[v];
        unsigned int s = (unsigned int)rng_below(&r, n);
        for (unsigned int k = 0; k < n; ++k)
            onehot[k] = (k == s) ? 1.0f : 0.0f;
        if (bp_bn_add_def(bn, v, NULL, 0, onehot, n)) {
            rc = -1;
            break;
        }
    }
    free(perm);
    return rc;
}

bp_bn *bp_bn_generate(const bp_gen_params *p)
{
    bp_clear_error();
    unsigned int V = (p->kind == BP_WL_LATTICE) ? p->rows * p->cols : p->num_nodes;
    if (V == 0 || p->states_lo == 0 || p->states_hi < p->states_lo || p->states_hi > MAX_STATES) {
        bp_fail("generator: bad parameters");
        return NULL;
    }
    bp_bn *bn = bp_bn_create(V);
    if (!bn) {
        bp_fail("generator: out of host memory");
        return NULL;
    }
    rng_t r;
    rng_seed(&r, p->seed);
    char name[CHAR_BUFFER_SIZE];
    for (unsigned int v = 0; v < V; ++v) {
        unsigned int n = p->states_lo + (unsigned int)rng_below(&r, p->states_hi - p->states_lo + 1);
        snprintf(name, sizeof(name), "n%u", v);
        bp_bn_set_var(bn, v, name, n);
    }

    unsigned int *children = (unsigned int *)calloc(V, sizeof(unsigned int));
    uint64_t scratch_len = 1ull << 20;
    float *scratch = (float *)malloc(sizeof(float) * scratch_len);
    unsigned int parents[64];
    if (!children || !scratch) {
        free(children);
        free(scratch);
        bp_bn_free(bn);
        bp_fail("generator: out of host memory");
        return NULL;
    }
    int rc = 0;
    for (unsigned int v = 0; v < V && rc == 0; ++v) {
        unsigned int np = 0;
        if (p->kind == BP_WL_LATTICE) {
            unsigned int row = v / p->cols, col = v % p->cols;
            if (col > 0)
                parents[np++] = v - 1;
            if (row > 0)
                parents[np++] = v - p->cols;
        } else if (v > 0) {
            unsigned int want;
            if (p->kind == BP_WL_TREE)
                want = 1;
            else if (p->indegree_dist == 1) {
                want = rng_poisson(&r, p->indeg_mean);
                if (want > p->indeg_hi)
                    want = p->indeg_hi;
            } else
                want = p->indeg_lo + (unsigned int)rng_below(&r, p->indeg_hi - p->indeg_lo + 1);
            if (want > v)
                want = v;
            if (want > 63)
                want = 63;
            /* rejection draws; a node whose candidates are exhausted gets fewer parents */
            for (unsigned int tries = 0; np < want && tries < 64u * want + 64u; ++tries) {
                unsigned int u = (unsigned int)rng_below(&r, v);
                if (children[u] >= p->max_children)
                    continue;
                int dup = 0;
                for (unsigned int i = 0; i < np; ++i)
                    dup |= (parents[i] == u);
                if (!dup)
                    parents[np++] = u;
            }
            qsort(parents, np, sizeof(unsigned int), cmp_uint);
        }
        for (unsigned int i = 0; i < np; ++i)
            children[parents[i]]++;
        rc = add_table_def(bn, &r, p, v, parents, np, scratch, scratch_len);
    }
    free(children);
    free(scratch);
    if (rc == 0 && p->evidence_fraction > 0.0)
        rc = bp_bn_add_evidence(bn, p->evidence_fraction, p->seed);
    if (rc != 0) {
        bp_bn_free(bn);
        return NULL;
    }
    return bn;
}
