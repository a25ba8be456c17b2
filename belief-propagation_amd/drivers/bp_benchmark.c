/*
 * bp_benchmark -- the reference's benchmark driver (src/c_benchmark/main.c:
 * 184-347, 366-380, 483-509) over libbp_amd.so: for each network file, tree
 * BP ("regular"), loopy BP until convergence ("loopy") and the edge-centric
 * loopy schedule ("loopy-edge", XMLBIF files only, as in the reference),
 * one CSV row each in the reference's schema:
 *   File Name,Propagation Type,Number of Nodes,Number of Edges,Diameter,
 *   Number of Iterations,BP Run Time(s)
 * The propagation runs on the GPU through the drop-in graph API.  Times
 * are wall-clock seconds (the reference uses clock(), which sums CPU time
 * over threads); they cover levels + sweeps + marginalisation for tree BP
 * and init_previous_edge + loopy_propagate_until(_edge) for loopy BP, the
 * spans the reference times.
 *
 * usage: bp_benchmark [-o out.csv] [-r repeats] file.xml|file.bif ...
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/constants.h"
#include "../../include/expression.h"
#include "../../include/graph.h"
#include "../../include/xml-expression.h"

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int is_xml(const char *path)
{
    size_t n = strlen(path);
    return n >= 4 && strcmp(path + n - 4, ".xml") == 0;
}

static Graph_t load(const char *path)
{
    Graph_t g = is_xml(path) ? parse_xml_file(path) : parse_bif_file(path);
    if (g) {
        set_up_src_nodes_to_edges(g);
        set_up_dest_nodes_to_edges(g);
    }
    return g;
}

static void run_tree(const char *path, FILE *out)
{
    Graph_t g = load(path);
    if (!g) {
        fprintf(stderr, "bp_benchmark: cannot read %s\n", path);
        return;
    }
    double t0 = now();
    init_levels_to_nodes(g);
    propagate_using_levels_start(g);
    for (unsigned int i = 1; i + 1 < g->num_levels; ++i)
        propagate_using_levels(g, i);
    reset_visited(g);
    for (unsigned int i = g->num_levels - 1; i > 0; --i)
        propagate_using_levels(g, i);
    marginalize(g);
    double t = now() - t0;
    fprintf(out, "%s,regular,%d,%d,%d,2,%lf\n", path, g->current_num_vertices, g->current_num_edges, g->diameter, t);
    fflush(out);
    graph_destroy(g);
}

static void run_loopy(const char *path, int edge, FILE *out)
{
    Graph_t g = load(path);
    if (!g) {
        fprintf(stderr, "bp_benchmark: cannot read %s\n", path);
        return;
    }
    double t0 = now();
    init_previous_edge(g);
    unsigned int iterations = edge ? loopy_propagate_until_edge(g, PRECISION, NUM_ITERATIONS)
                                   : loopy_propagate_until(g, PRECISION, NUM_ITERATIONS);
    double t = now() - t0;
    fprintf(out, "%s,%s,%d,%d,%d,%d,%lf\n", path, edge ? "loopy-edge" : "loopy", g->current_num_vertices,
            g->current_num_edges, g->diameter, iterations, t);
    fflush(out);
    graph_destroy(g);
}

int main(int argc, char **argv)
{
    const char *out_path = NULL;
    unsigned int repeats = 1;
    int first = 1;
    while (first < argc && argv[first][0] == '-') {
        if (strcmp(argv[first], "-o") == 0 && first + 1 < argc) {
            out_path = argv[first + 1];
            first += 2;
        } else if (strcmp(argv[first], "-r") == 0 && first + 1 < argc) {
            repeats = (unsigned int)strtoul(argv[first + 1], NULL, 10);
            first += 2;
        } else {
            fprintf(stderr, "usage: %s [-o out.csv] [-r repeats] file.xml|file.bif ...\n", argv[0]);
            return strcmp(argv[first], "-h") == 0 || strcmp(argv[first], "--help") == 0 ? 0 : 2;
        }
    }
    FILE *out = out_path ? fopen(out_path, "w") : stdout;
    if (!out) {
        fprintf(stderr, "bp_benchmark: cannot open %s\n", out_path);
        return 1;
    }
    fprintf(out, "File Name,Propagation Type,Number of Nodes,Number of Edges,Diameter,Number of Iterations,BP Run "
                 "Time(s)\n");
    fflush(out);
    for (int f = first; f < argc; ++f) {
        for (unsigned int r = 0; r < repeats; ++r)
            run_tree(argv[f], out);
        for (unsigned int r = 0; r < repeats; ++r)
            run_loopy(argv[f], 0, out);
        if (is_xml(argv[f]))
            for (unsigned int r = 0; r < repeats; ++r)
                run_loopy(argv[f], 1, out);
    }
    if (out != stdout)
        fclose(out);
    return 0;
}
