/*
 * bp_bif.c -- self-contained BIF reader: the drop-in for the reference's
 * flex/bison parser (src/bnf-parser/Lexer.l, Parser.y) and build_graph
 * (src/bnf-parser/expression.c:919-937).
 *
 * Lexer: the token rules of Lexer.l:20-76 with flex's matching discipline --
 * at each position the longest match wins, ties go to the earlier rule;
 * whitespace, ',', '|', comments and any other single character (quotes
 * included) are skipped.  Numbers convert like sscanf("%f") / sscanf("%d").
 *
 * Parser: recursive descent over the grammar of Parser.y:86-256, building
 * the tree shapes bison builds: every list is left-recursive, so a list
 * node's `left` is the list without its last element and the head holds the
 * last element.
 *
 * build_graph: the reference walks declaration chains from their head, so
 * declarations are visited in reverse file order (nodes, then priors, then
 * edges); outcome names are stored last-first and then swapped pairwise
 * j <-> n/2 - j (expression.c:902-917); a conditional's rows come from
 * "table" lists, "default" lists (filling only unset entries) and
 * "(states) values;" entries placed child-state-major at the offset of the
 * named parent states (expression.c:398-622); then every parent gets a
 * pairwise table exactly as in the XMLBIF reader.
 */
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "bp_internal.h"
#include "../../include/bp_model.h"
#include "../../include/expression.h"

/* ------------------------------------------------------------------------ */
/* tree nodes                                                                */

struct expression *create_expression(eType type, struct expression *left, struct expression *right)
{
    struct expression *e = (struct expression *)calloc(1, sizeof(*e));
    if (!e) {
        bp_fail("create_expression: out of host memory");
        return NULL;
    }
    e->type = type;
    e->left = left;
    e->right = right;
    return e;
}

void delete_expression(struct expression *root)
{
    /* explicit stack: declaration lists can be very long */
    size_t cap = 64, n = 0;
    struct expression **stack = (struct expression **)malloc(sizeof(*stack) * cap);
    if (!stack)
        return;
    if (root)
        stack[n++] = root;
    while (n) {
        struct expression *e = stack[--n];
        if (n + 2 > cap) {
            cap *= 2;
            struct expression **s = (struct expression **)realloc(stack, sizeof(*stack) * cap);
            if (!s)
                break;
            stack = s;
        }
        if (e->left)
            stack[n++] = e->left;
        if (e->right)
            stack[n++] = e->right;
        free(e);
    }
    free(stack);
}

static const char *type_name(eType t)
{
    static const char *names[] = {"Compilation Unit", "Network Declaration", "Network Content", "Property List",
                                  "Property", "Blank", "Variable or Probability Declaration",
                                  "Variable or Probability", "Variable Declaration", "Variable Content",
                                  "Variable Discrete", "Variable Values List", "Probability Declaration",
                                  "Probability Variables List", "Probability Variable Names",
                                  "Probability Content", "Probability Content List", "Probability Default Entry",
                                  "Probability Entry", "Probability Values List", "Probability Values",
                                  "Probability Table", "Floating Point List"};
    return (t >= COMPILATION_UNIT && t <= FLOATING_POINT_LIST) ? names[t] : "?";
}

void print_expression(struct expression *e)
{
    printf("Expression: {\nType: %s\n", type_name(e->type));
    switch (e->type) {
    case NETWORK_DECLARATION:
    case PROPERTY:
    case VARIABLE_DECLARATION:
    case VARIABLE_VALUES_LIST:
    case PROBABILITY_VARIABLE_NAMES:
    case PROBABILITY_VALUES:
        printf("Content: %.*s\n", CHAR_BUFFER_SIZE, e->value);
        break;
    case FLOATING_POINT_LIST:
        printf("Float value: %f\n", (double)e->float_value);
        break;
    case VARIABLE_DISCRETE:
        printf("Int value: %d\n", e->int_value);
        break;
    default:
        break;
    }
    printf("}\n");
}

/* ------------------------------------------------------------------------ */
/* lexer                                                                     */

enum tok {
    T_EOF = 0, T_NETWORK, T_VARIABLE, T_PROBABILITY, T_PROPERTY, T_TYPE, T_DISCRETE, T_DEFAULT, T_TABLE,
    T_LBRACE, T_RBRACE, T_LBRACKET, T_RBRACKET, T_LPAREN, T_RPAREN, T_SEMI, T_FLOAT, T_DECIMAL, T_WORD,
    T_SKIP
};

typedef struct {
    const char *p, *end;
    int tok;
    char word[CHAR_BUFFER_SIZE];
    float f;
    int i;
    unsigned line;
    int failed;
} lex_t;

static int is_digit(char c) { return c >= '0' && c <= '9'; }
static int is_alpha_(char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_' || c == '-'; }

static size_t m_exponent(const char *p, const char *end)
{
    if (p >= end || (*p != 'e' && *p != 'E'))
        return 0;
    const char *q = p + 1;
    if (q < end && (*q == '+' || *q == '-'))
        ++q;
    const char *d = q;
    while (q < end && is_digit(*q))
        ++q;
    return q > d ? (size_t)(q - p) : 0;
}

static size_t m_float(const char *p, const char *end)
{
    size_t best = 0;
    const char *q = p;
    while (q < end && is_digit(*q))
        ++q;
    size_t nd = (size_t)(q - p);
    if (q < end && *q == '.') {
        const char *r = q + 1;
        while (r < end && is_digit(*r))
            ++r;
        if (r > q + 1) { /* [0-9]+.[0-9]+ or .[0-9]+ (exponent optional) */
            size_t len = (size_t)(r - p) + m_exponent(r, end);
            if (nd > 0 || q == p)
                best = len;
        }
    }
    if (nd > 0) { /* [0-9]+ exponent */
        size_t e = m_exponent(q, end);
        if (e && nd + e > best)
            best = nd + e;
    }
    return best;
}

static size_t m_word(const char *p, const char *end)
{
    const char *q = p;
    while (q < end && is_digit(*q))
        ++q;
    if (q >= end || !is_alpha_(*q))
        return 0;
    while (q < end && (is_alpha_(*q) || is_digit(*q)))
        ++q;
    return (size_t)(q - p);
}

static size_t m_lit(const char *p, const char *end, const char *lit)
{
    size_t n = strlen(lit);
    return ((size_t)(end - p) >= n && memcmp(p, lit, n) == 0) ? n : 0;
}

static void next_token(lex_t *L)
{
    for (;;) {
        const char *p = L->p, *end = L->end;
        if (p >= end) {
            L->tok = T_EOF;
            return;
        }
        size_t best = 0;
        int tok = T_SKIP;
#define CAND(len_expr, t)                  \
    do {                                   \
        size_t len_ = (len_expr);          \
        if (len_ > best) {                 \
            best = len_;                   \
            tok = (t);                     \
        }                                  \
    } while (0)
        /* rules in the order of Lexer.l:56-76 */
        {
            const char *q = p;
            while (q < end && (*q == ' ' || *q == '\n' || *q == '\t' || *q == '\r'))
                ++q;
            CAND((size_t)(q - p), T_SKIP);
        }
        CAND(*p == ',' ? 1u : 0u, T_SKIP);
        CAND(*p == '|' ? 1u : 0u, T_SKIP);
        if (m_lit(p, end, "//")) { /* ends with \n or \r\n */
            const char *q = p + 2;
            while (q < end && *q != '\n' && *q != '\r')
                ++q;
            size_t len = 0;
            if (q < end && *q == '\n')
                len = (size_t)(q + 1 - p);
            else if (q + 1 < end && q[0] == '\r' && q[1] == '\n')
                len = (size_t)(q + 2 - p);
            CAND(len, T_SKIP);
        }
        if (m_lit(p, end, "/*")) {
            const char *q = p + 2;
            size_t len = 0;
            while (q + 1 < end) {
                if (q[0] == '*' && q[1] == '/') {
                    len = (size_t)(q + 2 - p);
                    break;
                }
                ++q;
            }
            CAND(len, T_SKIP);
        }
        CAND(m_lit(p, end, "network"), T_NETWORK);
        CAND(m_lit(p, end, "variable"), T_VARIABLE);
        CAND(m_lit(p, end, "probability"), T_PROBABILITY);
        if (m_lit(p, end, "property")) {
            const char *q = memchr(p + 8, ';', (size_t)(end - p - 8));
            if (q)
                CAND((size_t)(q + 1 - p), T_PROPERTY);
        }
        CAND(m_lit(p, end, "type"), T_TYPE);
        CAND(m_lit(p, end, "discrete"), T_DISCRETE);
        CAND(m_lit(p, end, "default"), T_DEFAULT);
        CAND(m_lit(p, end, "table"), T_TABLE);
        CAND(*p == '{' ? 1u : 0u, T_LBRACE);
        CAND(*p == '}' ? 1u : 0u, T_RBRACE);
        CAND(*p == '[' ? 1u : 0u, T_LBRACKET);
        CAND(*p == ']' ? 1u : 0u, T_RBRACKET);
        CAND(*p == '(' ? 1u : 0u, T_LPAREN);
        CAND(*p == ')' ? 1u : 0u, T_RPAREN);
        CAND(*p == ';' ? 1u : 0u, T_SEMI);
        CAND(m_float(p, end), T_FLOAT);
        {
            const char *q = p;
            while (q < end && is_digit(*q))
                ++q;
            CAND((size_t)(q - p), T_DECIMAL);
        }
        CAND(m_word(p, end), T_WORD);
        CAND(1u, T_SKIP); /* '.' : any other character */
#undef CAND
        for (const char *q = p; q < p + best; ++q)
            if (*q == '\n')
                L->line++;
        L->p = p + best;
        if (tok == T_SKIP)
            continue;
        char text[256];
        size_t n = best < sizeof(text) - 1 ? best : sizeof(text) - 1;
        memcpy(text, p, n);
        text[n] = '\0';
        if (tok == T_WORD || tok == T_PROPERTY) {
            memset(L->word, 0, sizeof(L->word));
            memcpy(L->word, p, best < CHAR_BUFFER_SIZE ? best : CHAR_BUFFER_SIZE);
        } else if (tok == T_FLOAT) {
            L->f = strtof(text, NULL);
        } else if (tok == T_DECIMAL) {
            long v = strtol(text, NULL, 10);
            L->i = (int)v;
        }
        L->tok = tok;
        return;
    }
}

/* ------------------------------------------------------------------------ */
/* parser (Parser.y:86-256)                                                  */

typedef struct {
    lex_t L;
    int failed;
} parser_t;

static struct expression *syntax(parser_t *P, const char *what)
{
    if (!P->failed)
        bp_fail("BIF syntax error at line %u: %s", P->L.line + 1, what);
    P->failed = 1;
    return NULL;
}

static int accept(parser_t *P, int tok)
{
    if (P->L.tok != tok)
        return 0;
    next_token(&P->L);
    return 1;
}

static struct expression *node(parser_t *P, eType t, struct expression *l, struct expression *r)
{
    struct expression *e = create_expression(t, l, r);
    if (!e)
        P->failed = 1;
    return e;
}

static void set_value_from_token(struct expression *e, const lex_t *L)
{
    if (L->tok == T_DECIMAL)
        snprintf(e->value, CHAR_BUFFER_SIZE, "%d", L->i);
    else
        memcpy(e->value, L->word, CHAR_BUFFER_SIZE);
}

/* WORD | DECIMAL list as a left-recursive chain of `t` nodes */
static struct expression *name_list(parser_t *P, eType t)
{
    struct expression *head = NULL;
    while (P->L.tok == T_WORD || P->L.tok == T_DECIMAL) {
        struct expression *e = node(P, t, head, NULL);
        if (!e)
            return head;
        set_value_from_token(e, &P->L);
        head = e;
        next_token(&P->L);
    }
    if (!head)
        return syntax(P, "expected a name");
    return head;
}

static struct expression *float_list(parser_t *P)
{
    struct expression *head = NULL;
    while (P->L.tok == T_FLOAT || P->L.tok == T_DECIMAL) {
        struct expression *e = node(P, FLOATING_POINT_LIST, head, NULL);
        if (!e)
            return head;
        e->float_value = (P->L.tok == T_FLOAT) ? P->L.f : (float)P->L.i;
        head = e;
        next_token(&P->L);
    }
    if (!head)
        return syntax(P, "expected a number");
    return head;
}

static struct expression *property(parser_t *P)
{
    struct expression *e = node(P, PROPERTY, NULL, NULL);
    if (e)
        memcpy(e->value, P->L.word, CHAR_BUFFER_SIZE);
    next_token(&P->L);
    return e;
}

/* consecutive properties form one list (bison shifts rather than reduces) */
static struct expression *property_list(parser_t *P)
{
    struct expression *head = node(P, PROPERTY_LIST, property(P), NULL);
    while (head && P->L.tok == T_PROPERTY)
        head = node(P, PROPERTY_LIST, head, property(P));
    return head;
}

static struct expression *variable_discrete(parser_t *P)
{
    next_token(&P->L); /* type */
    if (!accept(P, T_DISCRETE) || !accept(P, T_LBRACKET) || P->L.tok != T_DECIMAL)
        return syntax(P, "expected 'discrete [ n ]'");
    int n = P->L.i;
    next_token(&P->L);
    if (!accept(P, T_RBRACKET) || !accept(P, T_LBRACE))
        return syntax(P, "expected '] {'");
    struct expression *values = name_list(P, VARIABLE_VALUES_LIST);
    if (P->failed)
        return values;
    if (!accept(P, T_RBRACE) || !accept(P, T_SEMI)) {
        delete_expression(values);
        return syntax(P, "expected '};' after the values");
    }
    struct expression *e = node(P, VARIABLE_DISCRETE, values, NULL);
    if (e)
        e->int_value = n;
    return e;
}

static struct expression *variable_declaration(parser_t *P)
{
    next_token(&P->L); /* variable */
    if (P->L.tok != T_WORD)
        return syntax(P, "expected a variable name");
    struct expression *decl = node(P, VARIABLE_DECLARATION, NULL, NULL);
    if (!decl)
        return NULL;
    memcpy(decl->value, P->L.word, CHAR_BUFFER_SIZE);
    next_token(&P->L);
    if (!accept(P, T_LBRACE)) {
        delete_expression(decl);
        return syntax(P, "expected '{'");
    }
    struct expression *items = NULL;
    while (!P->failed && (P->L.tok == T_PROPERTY || P->L.tok == T_TYPE)) {
        struct expression *it = (P->L.tok == T_PROPERTY) ? property(P) : variable_discrete(P);
        items = items ? node(P, VARIABLE_OR_PROBABILITY, items, it) : node(P, VARIABLE_OR_PROBABILITY, it, NULL);
    }
    if (P->failed || !accept(P, T_RBRACE)) {
        delete_expression(items);
        delete_expression(decl);
        return syntax(P, "expected '}' after a variable");
    }
    decl->left = items ? node(P, VARIABLE_CONTENT, items, NULL) : NULL;
    return decl;
}

static struct expression *probability_declaration(parser_t *P)
{
    next_token(&P->L); /* probability */
    if (!accept(P, T_LPAREN))
        return syntax(P, "expected '('");
    struct expression *names = name_list(P, PROBABILITY_VARIABLE_NAMES);
    if (P->failed || !accept(P, T_RPAREN) || !accept(P, T_LBRACE)) {
        delete_expression(names);
        return syntax(P, "expected ') {'");
    }
    struct expression *vars = node(P, PROBABILITY_VARIABLES_LIST, names, NULL);
    struct expression *items = NULL;
    while (!P->failed && P->L.tok != T_RBRACE) {
        struct expression *it = NULL;
        if (P->L.tok == T_PROPERTY) {
            it = property_list(P);
        } else if (P->L.tok == T_DEFAULT) {
            next_token(&P->L);
            struct expression *f = float_list(P);
            if (!P->failed && !accept(P, T_SEMI))
                syntax(P, "expected ';' after default");
            it = node(P, PROBABILITY_DEFAULT_ENTRY, f, NULL);
        } else if (P->L.tok == T_TABLE) {
            next_token(&P->L);
            struct expression *f = float_list(P);
            if (!P->failed && !accept(P, T_SEMI))
                syntax(P, "expected ';' after table");
            it = node(P, PROBABILITY_TABLE, f, NULL);
        } else if (P->L.tok == T_LPAREN) {
            next_token(&P->L);
            struct expression *v = name_list(P, PROBABILITY_VALUES);
            if (!P->failed && !accept(P, T_RPAREN))
                syntax(P, "expected ')' after entry states");
            struct expression *vl = node(P, PROBABILITY_VALUES_LIST, v, NULL);
            struct expression *f = P->failed ? NULL : float_list(P);
            if (!P->failed && !accept(P, T_SEMI))
                syntax(P, "expected ';' after an entry");
            it = node(P, PROBABILITY_ENTRY, vl, f);
        } else {
            syntax(P, "unexpected token in a probability block");
            break;
        }
        items = items ? node(P, PROBABILITY_CONTENT_LIST, items, it) : node(P, PROBABILITY_CONTENT_LIST, it, NULL);
    }
    struct expression *content = NULL;
    if (!P->failed && accept(P, T_RBRACE))
        content = items ? node(P, PROBABILITY_CONTENT, items, NULL) : NULL;
    else {
        delete_expression(items);
        syntax(P, "expected '}' after a probability block");
    }
    return node(P, PROBABILITY_DECLARATION, vars, content);
}

static struct expression *compilation_unit(parser_t *P)
{
    if (!accept(P, T_NETWORK) || P->L.tok != T_WORD)
        return syntax(P, "expected 'network <name>'");
    struct expression *net = node(P, NETWORK_DECLARATION, NULL, NULL);
    if (!net)
        return NULL;
    memcpy(net->value, P->L.word, CHAR_BUFFER_SIZE);
    next_token(&P->L);
    if (!accept(P, T_LBRACE)) {
        delete_expression(net);
        return syntax(P, "expected '{' after the network name");
    }
    struct expression *props = (P->L.tok == T_PROPERTY) ? property_list(P) : NULL;
    if (!accept(P, T_RBRACE)) {
        delete_expression(props);
        delete_expression(net);
        return syntax(P, "expected '}' after the network block");
    }
    net->left = node(P, NETWORK_CONTENT, props, NULL);
    struct expression *decls = NULL;
    while (!P->failed && (P->L.tok == T_VARIABLE || P->L.tok == T_PROBABILITY)) {
        struct expression *d = (P->L.tok == T_VARIABLE) ? variable_declaration(P) : probability_declaration(P);
        decls = decls ? node(P, VARIABLE_OR_PROBABILITY_DECLARATION, decls, d)
                      : node(P, VARIABLE_OR_PROBABILITY_DECLARATION, d, NULL);
    }
    struct expression *unit = node(P, COMPILATION_UNIT, net, decls);
    if (!P->failed && P->L.tok != T_EOF)
        syntax(P, "unexpected token after the declarations");
    return unit;
}

struct expression *parse_bif_string(const char *text)
{
    bp_clear_error();
    parser_t P;
    memset(&P, 0, sizeof(P));
    P.L.p = text;
    P.L.end = text + strlen(text);
    next_token(&P.L);
    struct expression *root = compilation_unit(&P);
    if (P.failed) {
        delete_expression(root);
        return NULL;
    }
    return root;
}

struct expression *parse_bif_expression(const char *file_name)
{
    FILE *f = fopen(file_name, "rb");
    if (!f) {
        bp_fail("parse_bif_expression: cannot open %s: %s", file_name, strerror(errno));
        return NULL;
    }
    struct stat st;
    if (fstat(fileno(f), &st) != 0) {
        fclose(f);
        bp_fail("parse_bif_expression: cannot stat %s", file_name);
        return NULL;
    }
    char *buf = (char *)malloc((size_t)st.st_size + 1);
    if (!buf || fread(buf, 1, (size_t)st.st_size, f) != (size_t)st.st_size) {
        free(buf);
        fclose(f);
        bp_fail("parse_bif_expression: cannot read %s", file_name);
        return NULL;
    }
    fclose(f);
    buf[st.st_size] = '\0';
    struct expression *root = parse_bif_string(buf);
    free(buf);
    return root;
}

/* ------------------------------------------------------------------------ */
/* build_graph                                                               */

typedef struct {
    struct expression **items;
    size_t n, cap;
    struct expression *network;
} decls_t;

static int push_decl(decls_t *D, struct expression *e)
{
    if (D->n == D->cap) {
        size_t cap = D->cap ? D->cap * 2 : 64;
        struct expression **p = (struct expression **)realloc(D->items, cap * sizeof(*p));
        if (!p)
            return -1;
        D->items = p;
        D->cap = cap;
    }
    D->items[D->n++] = e;
    return 0;
}

/* Declarations in the order the reference visits them (expression.c:
 * 297-326,858-884): along a declaration chain the element held by each
 * link first, then the chain's tail -- reverse file order. */
static int collect(struct expression *e, decls_t *D)
{
    if (!e)
        return 0;
    if (e->type == NETWORK_DECLARATION) {
        D->network = e;
        return 0;
    }
    if (e->type == VARIABLE_DECLARATION || e->type == PROBABILITY_DECLARATION)
        return push_decl(D, e);
    if (e->type == VARIABLE_OR_PROBABILITY_DECLARATION) {
        struct expression *next = e;
        while (next && next->type == VARIABLE_OR_PROBABILITY_DECLARATION) {
            if (collect(next->right, D))
                return -1;
            next = next->left;
        }
        return next ? collect(next, D) : 0;
    }
    if (collect(e->left, D))
        return -1;
    return collect(e->right, D);
}

/* the outcome names of a variable, in the order the reference stores them
 * (expression.c:225-296): items of the content list last-first, values of
 * a discrete list last-first */
static int variable_states(struct expression *e, Graph_t g, unsigned int v, unsigned int *count)
{
    if (!e)
        return 0;
    if (e->type == VARIABLE_DISCRETE) {
        if (e->int_value < 0 || e->int_value > MAX_STATES) {
            bp_fail("build_graph: variable %u declares %d states (limit %d)", v, e->int_value, MAX_STATES);
            return -1;
        }
        for (struct expression *x = e->left; x; x = x->left) {
            if (*count >= MAX_STATES) {
                bp_fail("build_graph: variable %u lists more than %d values", v, MAX_STATES);
                return -1;
            }
            strncpy(&g->variable_names[(size_t)v * MAX_STATES * CHAR_BUFFER_SIZE + (size_t)*count * CHAR_BUFFER_SIZE],
                    x->value, CHAR_BUFFER_SIZE);
            *count += 1;
        }
        if ((int)*count != e->int_value) {
            bp_fail("build_graph: variable %u declares %d states but lists %u values", v, e->int_value, *count);
            return -1;
        }
        return 0;
    }
    if (e->type == VARIABLE_OR_PROBABILITY) {
        struct expression *next = e;
        while (next && next->type == VARIABLE_OR_PROBABILITY) {
            if (variable_states(next->right, g, v, count))
                return -1;
            next = next->left;
        }
        return next ? variable_states(next, g, v, count) : 0;
    }
    if (variable_states(e->left, g, v, count))
        return -1;
    return variable_states(e->right, g, v, count);
}

/* names of a probability declaration, head first (expression.c:354-389) */
static unsigned int count_names(struct expression *e)
{
    if (!e || e->type == FLOATING_POINT_LIST)
        return 0;
    if (e->type == PROBABILITY_VARIABLE_NAMES) {
        unsigned int c = 0;
        for (struct expression *x = e; x; x = x->left)
            c++;
        return c;
    }
    return count_names(e->left) + count_names(e->right);
}

static void gather_names(struct expression *e, char *buf, unsigned int *idx, unsigned int cap)
{
    if (!e || e->type == PROBABILITY_CONTENT)
        return;
    if (e->type == PROBABILITY_VARIABLE_NAMES && *idx < cap) {
        strncpy(&buf[(size_t)*idx * CHAR_BUFFER_SIZE], e->value, CHAR_BUFFER_SIZE);
        *idx += 1;
    }
    gather_names(e->left, buf, idx, cap);
    gather_names(e->right, buf, idx, cap);
}

static void reverse_names(char *buf, unsigned int n)
{
    char tmp[CHAR_BUFFER_SIZE];
    for (unsigned int i = 0; i < n / 2; ++i) {
        char *a = &buf[(size_t)i * CHAR_BUFFER_SIZE], *b = &buf[(size_t)(n - 1 - i) * CHAR_BUFFER_SIZE];
        strncpy(tmp, a, CHAR_BUFFER_SIZE);
        strncpy(a, b, CHAR_BUFFER_SIZE);
        strncpy(b, tmp, CHAR_BUFFER_SIZE);
    }
}

static void reverse_floats(float *v, uint64_t n)
{
    for (uint64_t i = 0; i < n / 2; ++i) {
        float t = v[i];
        v[i] = v[n - 1 - i];
        v[n - 1 - i] = t;
    }
}

typedef struct {
    float *buf;
    uint64_t num;
    int failed;
} fill_t;

/* "table" lists: values head (last) first from index 0 (expression.c:415-433,477-495) */
static void fill_tables(struct expression *e, fill_t *F)
{
    if (!e || e->type == FLOATING_POINT_LIST || F->failed)
        return;
    if (e->type == PROBABILITY_TABLE) {
        uint64_t idx = 0;
        for (struct expression *x = e->left; x && x->type == FLOATING_POINT_LIST; x = x->left) {
            if (idx >= F->num) {
                bp_fail("build_graph: a table lists more than %llu values", (unsigned long long)F->num);
                F->failed = 1;
                return;
            }
            F->buf[idx++] = x->float_value;
        }
    }
    fill_tables(e->left, F);
    fill_tables(e->right, F);
}

/* "default" lists: the list (head first) repeated over the unset entries
 * (expression.c:398-413,457-475) */
static void fill_defaults(struct expression *e, fill_t *F)
{
    if (!e || e->type == FLOATING_POINT_LIST || F->failed)
        return;
    if (e->type == PROBABILITY_DEFAULT_ENTRY && e->left) {
        uint64_t idx = 0;
        while (idx < F->num) {
            for (struct expression *x = e->left; x && x->type == FLOATING_POINT_LIST; x = x->left) {
                if (idx >= F->num) {
                    bp_fail("build_graph: a default row does not divide the %llu table entries",
                            (unsigned long long)F->num);
                    F->failed = 1;
                    return;
                }
                if (F->buf[idx] < 0)
                    F->buf[idx] = x->float_value;
                idx++;
            }
        }
    }
    fill_defaults(e->left, F);
    fill_defaults(e->right, F);
}

/* "(states) values;" entries (expression.c:435-455,523-622) */
static void fill_entries(struct expression *e, fill_t *F, Graph_t g, const char *names, unsigned int num_names,
                         unsigned int child_states)
{
    if (!e || e->type == FLOATING_POINT_LIST || F->failed)
        return;
    if (e->type == PROBABILITY_ENTRY) {
        uint64_t jump = F->num / child_states;
        unsigned int count = 0;
        for (struct expression *x = e->left ? e->left->left : NULL; x; x = x->left)
            count++;
        if (count >= num_names) {
            bp_fail("build_graph: an entry names %u states for %u parents", count, num_names - 1);
            F->failed = 1;
            return;
        }
        /* parent states in file order */
        unsigned int idx_state[MAX_STATES + 64];
        unsigned int k = count;
        uint64_t pos = 0, step = 1;
        for (struct expression *x = e->left ? e->left->left : NULL; x; x = x->left) {
            k--; /* the chain head is the last state named */
            unsigned int node = find_node_by_name((char *)&names[(size_t)(k + 1) * CHAR_BUFFER_SIZE], g);
            if (bp_has_error()) {
                F->failed = 1;
                return;
            }
            unsigned int j, n = g->node_num_vars[node];
            for (j = 0; j < n; ++j)
                if (strncmp(x->value, &g->variable_names[(size_t)node * MAX_STATES * CHAR_BUFFER_SIZE +
                                                         (size_t)j * CHAR_BUFFER_SIZE],
                            CHAR_BUFFER_SIZE) == 0)
                    break;
            if (j == n) {
                bp_fail("build_graph: '%.*s' is not a state of '%.*s'", CHAR_BUFFER_SIZE, x->value,
                        CHAR_BUFFER_SIZE, &names[(size_t)(k + 1) * CHAR_BUFFER_SIZE]);
                F->failed = 1;
                return;
            }
            if (k < sizeof(idx_state) / sizeof(idx_state[0]))
                idx_state[k] = j;
        }
        /* mixed radix over the named parents, the last one fastest */
        for (unsigned int q = count; q > 0; --q) {
            pos += (uint64_t)idx_state[q - 1] * step;
            unsigned int node = find_node_by_name((char *)&names[(size_t)q * CHAR_BUFFER_SIZE], g);
            step *= g->node_num_vars[node];
        }
        /* values head (last) first into child states n-1, n-2, ... */
        unsigned int cur = 0;
        for (struct expression *x = e->right; x && x->type == FLOATING_POINT_LIST; x = x->left) {
            if (cur >= child_states) {
                bp_fail("build_graph: an entry lists more values than the child has states");
                F->failed = 1;
                return;
            }
            uint64_t at = ((uint64_t)(child_states - cur - 1) * jump + pos) % F->num;
            F->buf[at] = x->float_value;
            cur++;
        }
    }
    fill_entries(e->left, F, g, names, num_names, child_states);
    fill_entries(e->right, F, g, names, num_names, child_states);
}

static uint64_t table_size(Graph_t g, const char *names, unsigned int n, int *failed)
{
    uint64_t num = 1;
    for (unsigned int i = 0; i < n; ++i) {
        unsigned int node = find_node_by_name((char *)&names[(size_t)i * CHAR_BUFFER_SIZE], g);
        if (bp_has_error()) {
            *failed = 1;
            return 0;
        }
        num *= g->node_num_vars[node];
    }
    return num;
}

static int add_probability(Graph_t g, struct expression *decl, int edges)
{
    unsigned int n = count_names(decl);
    if (edges ? n <= 1 : n != 1)
        return 0;
    char *names = (char *)calloc(n, CHAR_BUFFER_SIZE);
    if (!names) {
        bp_fail("build_graph: out of host memory");
        return -1;
    }
    unsigned int idx = 0;
    gather_names(decl, names, &idx, n);
    if (edges)
        reverse_names(names, n); /* [child, parent 1, ..., parent m] */
    int failed = 0;
    uint64_t num = table_size(g, names, n, &failed);
    fill_t F = {NULL, num, 0};
    if (!failed)
        F.buf = (float *)malloc(sizeof(float) * (num ? num : 1));
    if (failed || !F.buf) {
        if (!failed)
            bp_fail("build_graph: out of host memory");
        free(names);
        return -1;
    }
    for (uint64_t i = 0; i < num; ++i)
        F.buf[i] = -1.0f;
    fill_tables(decl, &F);
    fill_defaults(decl, &F);
    int rc = 0;
    if (!edges) {
        /* priors (expression.c:624-669): unset -> 0, then file order */
        for (uint64_t i = 0; i < num; ++i)
            if (F.buf[i] < 0.0f)
                F.buf[i] = 0.0f;
        reverse_floats(F.buf, num);
        unsigned int node = find_node_by_name(names, g);
        if (F.failed || bp_has_error())
            rc = -1;
        else {
            graph_set_node_state(g, node, (unsigned int)num, F.buf);
            rc = bp_has_error() ? -1 : 0;
        }
    } else {
        /* conditionals (expression.c:766-816) */
        reverse_floats(F.buf, num);
        unsigned int child = find_node_by_name(names, g);
        unsigned int child_states = g->node_num_vars[child];
        if (!F.failed && child_states)
            fill_entries(decl, &F, g, names, n, child_states);
        for (uint64_t i = 0; i < num; ++i)
            if (F.buf[i] < 0.0f)
                F.buf[i] = 0.0f;
        unsigned int *parents = (unsigned int *)malloc(sizeof(unsigned int) * n);
        if (F.failed || !parents || bp_has_error()) {
            rc = -1;
        } else {
            for (unsigned int i = 1; i < n; ++i)
                parents[i - 1] = find_node_by_name(&names[(size_t)i * CHAR_BUFFER_SIZE], g);
            rc = bp_has_error() ? -1 : bp_graph_add_cpt_edges(g, child, parents, n - 1, F.buf, num);
        }
        free(parents);
    }
    free(F.buf);
    free(names);
    return rc;
}

Graph_t build_graph(struct expression *root)
{
    bp_clear_error();
    decls_t D;
    memset(&D, 0, sizeof(D));
    if (collect(root, &D)) {
        free(D.items);
        bp_fail("build_graph: out of host memory");
        return NULL;
    }
    unsigned int V = 0;
    uint64_t E = 0;
    for (size_t i = 0; i < D.n; ++i) {
        if (D.items[i]->type == VARIABLE_DECLARATION)
            V++;
        else {
            unsigned int c = count_names(D.items[i]);
            E += c ? c - 1 : 0;
        }
    }
    Graph_t g = create_graph(V, (unsigned int)(2 * E));
    if (!g) {
        free(D.items);
        return NULL;
    }
    if (D.network)
        strncpy(g->graph_name, D.network->value, CHAR_BUFFER_SIZE);
    /* nodes */
    for (size_t i = 0; i < D.n; ++i) {
        struct expression *d = D.items[i];
        if (d->type != VARIABLE_DECLARATION)
            continue;
        unsigned int count = 0, v = g->current_num_vertices;
        if (variable_states(d->left, g, v, &count))
            goto fail;
        graph_add_node(g, count, d->value);
        if (bp_has_error())
            goto fail;
    }
    /* outcome names: swap j <-> n/2 - j (expression.c:902-917) */
    for (unsigned int v = 0; v < g->current_num_vertices; ++v) {
        unsigned int n = g->node_num_vars[v];
        char tmp[CHAR_BUFFER_SIZE];
        for (unsigned int j = 0; j < n / 2; ++j) {
            unsigned int k = n / 2 - j;
            if (j == k)
                continue;
            char *a = &g->variable_names[(size_t)v * MAX_STATES * CHAR_BUFFER_SIZE + (size_t)j * CHAR_BUFFER_SIZE];
            char *b = &g->variable_names[(size_t)v * MAX_STATES * CHAR_BUFFER_SIZE + (size_t)k * CHAR_BUFFER_SIZE];
            strncpy(tmp, a, CHAR_BUFFER_SIZE);
            strncpy(a, b, CHAR_BUFFER_SIZE);
            strncpy(b, tmp, CHAR_BUFFER_SIZE);
        }
    }
    for (size_t i = 0; i < D.n; ++i)
        if (D.items[i]->type == PROBABILITY_DECLARATION && add_probability(g, D.items[i], 0))
            goto fail;
    for (size_t i = 0; i < D.n; ++i)
        if (D.items[i]->type == PROBABILITY_DECLARATION && add_probability(g, D.items[i], 1))
            goto fail;
    free(D.items);
    return g;
fail:
    free(D.items);
    graph_destroy(g);
    return NULL;
}

Graph_t parse_bif_file(const char *file_name)
{
    struct expression *root = parse_bif_expression(file_name);
    if (!root)
        return NULL;
    Graph_t g = build_graph(root);
    delete_expression(root);
    return g;
}
