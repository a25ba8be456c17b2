/*
 * expression.h -- drop-in BIF reader (src/bnf-parser/expression.h:17-58).
 *
 * struct expression and eType keep the reference's layout and values: the
 * reference's BIF parser (flex/bison, src/bnf-parser/Lexer.l, Parser.y)
 * produces a tree of these nodes and build_graph() turns it into a Graph_t.
 * This library's reader builds the same tree shapes (left-recursive lists
 * whose head holds the last element) without flex or bison, and its
 * build_graph() reproduces the reference's graph bit for bit: declarations
 * walked in reverse file order, outcome names reversed and then partially
 * swapped, table / default / entry rows, pairwise tables for every parent.
 */
#ifndef BP_EXPRESSION_H_
#define BP_EXPRESSION_H_

#include "constants.h"
#include "graph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum expressionType {
    COMPILATION_UNIT,
    NETWORK_DECLARATION,
    NETWORK_CONTENT,
    PROPERTY_LIST,
    PROPERTY,
    BLANK,
    VARIABLE_OR_PROBABILITY_DECLARATION,
    VARIABLE_OR_PROBABILITY,
    VARIABLE_DECLARATION,
    VARIABLE_CONTENT,
    VARIABLE_DISCRETE,
    VARIABLE_VALUES_LIST,
    PROBABILITY_DECLARATION,
    PROBABILITY_VARIABLES_LIST,
    PROBABILITY_VARIABLE_NAMES,
    PROBABILITY_CONTENT,
    PROBABILITY_CONTENT_LIST,
    PROBABILITY_DEFAULT_ENTRY,
    PROBABILITY_ENTRY,
    PROBABILITY_VALUES_LIST,
    PROBABILITY_VALUES,
    PROBABILITY_TABLE,
    FLOATING_POINT_LIST
} eType;

struct expression {
    eType type;
    char value[CHAR_BUFFER_SIZE];
    float float_value;
    int int_value;
    struct expression *left;
    struct expression *right;
};

struct expression *create_expression(eType, struct expression *left, struct expression *right); /* expression.c:23-32 */
void delete_expression(struct expression *);                                                     /* expression.c:51-67 */
void print_expression(struct expression *);                                                      /* expression.c:69-121 */
Graph_t build_graph(struct expression *);                                                        /* expression.c:919-937 */

/* Reads a BIF file (or a NUL-terminated BIF text) into an expression tree;
 * NULL with bp_last_error() on a syntax error.  These replace the
 * yylex_init / yyset_in / yyparse sequence of the generated parser
 * (src/bnf-parser/main.c:58-77). */
struct expression *parse_bif_expression(const char *file_name);
struct expression *parse_bif_string(const char *text);

/* parse_bif_expression + build_graph + delete_expression. */
Graph_t parse_bif_file(const char *file_name);

#ifdef __cplusplus
}
#endif

#endif /* BP_EXPRESSION_H_ */
