/*
 * bp_xml.c -- self-contained XMLBIF reader (drop-in for
 * src/bnf-xml-parser/xml-expression.c:441-466, without libxml2).
 *
 * The reference evaluates these XPath queries with libxml2:
 *   //NETWORK/VARIABLE                 nodes, document order (:198-213)
 *   .//NAME/text(), .//OUTCOME/text()  per variable (:156-196)
 *   //NETWORK/DEFINITION/GIVEN         edge capacity = 2 x count (:137-150,459)
 *   //NETWORK/DEFINITION               definitions, two passes (:421-439)
 *   .//FOR/text(), .//GIVEN/text(), .//TABLE/text() per definition; TABLE
 *     uses its last text node (:216-288)
 * This reader scans the document once, keeping for every VARIABLE and
 * DEFINITION that is a child of a NETWORK element the text nodes found below
 * NAME / OUTCOME / FOR / GIVEN / TABLE elements.  A text node is a run of
 * character data (entity and character references expanded) ended by any
 * markup; a CDATA section is a text node of its own -- libxml2's model.
 * TABLE tokens are split on " \t\n\r" and converted like sscanf("%f")
 * (strtof); names are truncated to CHAR_BUFFER_SIZE like strncpy.
 */
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "bp_internal.h"
#include "../../include/bp_model.h"
#include "../../include/xml-expression.h"

/* a text node: [begin, end) in the decoded text arena */
typedef struct {
    uint64_t begin, end;
} span_t;

enum { F_NAME = 1, F_OUTCOME, F_FOR, F_GIVEN, F_TABLE };

typedef struct {
    int is_def;     /* 0: VARIABLE, 1: DEFINITION */
    uint64_t first; /* index of its first text record */
    uint64_t count;
} rec_t;

typedef struct {
    int field;   /* F_* */
    span_t text;
} trec_t;

typedef struct {
    const char *p, *end;
    char *arena;
    uint64_t arena_len, arena_cap;
    trec_t *texts;
    uint64_t ntexts, texts_cap;
    rec_t *recs;
    uint64_t nrecs, recs_cap;
    uint64_t given_direct; /* GIVEN elements that are children of a NETWORK-child DEFINITION */
    int failed;
} xml_t;

static int xml_error(xml_t *x, const char *what)
{
    if (!x->failed)
        bp_fail("parse_xml_file: %s", what);
    x->failed = 1;
    return -1;
}

static int arena_put(xml_t *x, const char *s, size_t n)
{
    if (x->arena_len + n + 1 > x->arena_cap) {
        uint64_t cap = x->arena_cap ? x->arena_cap : 1 << 16;
        while (cap < x->arena_len + n + 1)
            cap *= 2;
        char *a = (char *)realloc(x->arena, cap);
        if (!a)
            return xml_error(x, "out of host memory");
        x->arena = a;
        x->arena_cap = cap;
    }
    memcpy(x->arena + x->arena_len, s, n);
    x->arena_len += n;
    return 0;
}

static int push_text(xml_t *x, int field, span_t s)
{
    if (x->ntexts == x->texts_cap) {
        uint64_t cap = x->texts_cap ? x->texts_cap * 2 : 1024;
        trec_t *t = (trec_t *)realloc(x->texts, cap * sizeof(trec_t));
        if (!t)
            return xml_error(x, "out of host memory");
        x->texts = t;
        x->texts_cap = cap;
    }
    x->texts[x->ntexts].field = field;
    x->texts[x->ntexts].text = s;
    x->ntexts++;
    x->recs[x->nrecs - 1].count++;
    return 0;
}

static int push_rec(xml_t *x, int is_def)
{
    if (x->nrecs == x->recs_cap) {
        uint64_t cap = x->recs_cap ? x->recs_cap * 2 : 1024;
        rec_t *r = (rec_t *)realloc(x->recs, cap * sizeof(rec_t));
        if (!r)
            return xml_error(x, "out of host memory");
        x->recs = r;
        x->recs_cap = cap;
    }
    x->recs[x->nrecs].is_def = is_def;
    x->recs[x->nrecs].first = x->ntexts;
    x->recs[x->nrecs].count = 0;
    x->nrecs++;
    return 0;
}

static void utf8_put(xml_t *x, unsigned long cp)
{
    char buf[4];
    size_t n;
    if (cp < 0x80) {
        buf[0] = (char)cp;
        n = 1;
    } else if (cp < 0x800) {
        buf[0] = (char)(0xc0 | (cp >> 6));
        buf[1] = (char)(0x80 | (cp & 0x3f));
        n = 2;
    } else if (cp < 0x10000) {
        buf[0] = (char)(0xe0 | (cp >> 12));
        buf[1] = (char)(0x80 | ((cp >> 6) & 0x3f));
        buf[2] = (char)(0x80 | (cp & 0x3f));
        n = 3;
    } else {
        buf[0] = (char)(0xf0 | (cp >> 18));
        buf[1] = (char)(0x80 | ((cp >> 12) & 0x3f));
        buf[2] = (char)(0x80 | ((cp >> 6) & 0x3f));
        buf[3] = (char)(0x80 | (cp & 0x3f));
        n = 4;
    }
    arena_put(x, buf, n);
}

/* Element stack entry: which field a text node below it belongs to, and
 * whether it opened a record. */
typedef struct {
    char name[16];
    int field;
    int opens_rec;
    int is_network;
} elem_t;

static int name_is(const char *s, size_t n, const char *lit)
{
    return strlen(lit) == n && memcmp(s, lit, n) == 0;
}

static const char *skip_until(const char *p, const char *end, const char *lit)
{
    size_t n = strlen(lit);
    while (p + n <= end) {
        if (memcmp(p, lit, n) == 0)
            return p + n;
        ++p;
    }
    return NULL;
}

/* <!DOCTYPE ... [ internal subset ] > with quoted strings and comments */
static const char *skip_doctype(const char *p, const char *end)
{
    int depth = 0;
    while (p < end) {
        char c = *p;
        if (c == '"' || c == '\'') {
            const char *q = memchr(p + 1, c, (size_t)(end - p - 1));
            if (!q)
                return NULL;
            p = q + 1;
            continue;
        }
        if (c == '<' && p + 4 <= end && memcmp(p, "<!--", 4) == 0) {
            p = skip_until(p + 4, end, "-->");
            if (!p)
                return NULL;
            continue;
        }
        if (c == '[')
            depth++;
        else if (c == ']')
            depth--;
        else if (c == '>' && depth <= 0)
            return p + 1;
        ++p;
    }
    return NULL;
}

static int scan(xml_t *x)
{
    elem_t stack[256];
    int depth = 0;
    int cur_field = 0;   /* field of the element whose direct text is being read */
    int in_rec = 0;      /* inside a NETWORK-child VARIABLE/DEFINITION */
    const char *p = x->p, *end = x->end;
    int text_open = 0;
    span_t text = {0, 0};

#define CLOSE_TEXT()                                                        \
    do {                                                                    \
        if (text_open) {                                                    \
            text.end = x->arena_len;                                        \
            if (in_rec && cur_field && push_text(x, cur_field, text))       \
                return -1;                                                  \
            text_open = 0;                                                  \
        }                                                                   \
    } while (0)
#define OPEN_TEXT()                                                         \
    do {                                                                    \
        if (!text_open) {                                                   \
            text_open = 1;                                                  \
            text.begin = x->arena_len;                                      \
        }                                                                   \
    } while (0)

    while (p < end) {
        if (*p != '<') {
            if (depth == 0) { /* prolog / epilog whitespace */
                ++p;
                continue;
            }
            const char *q = p;
            while (q < end && *q != '<' && *q != '&')
                ++q;
            if (in_rec && cur_field) {
                OPEN_TEXT();
                if (arena_put(x, p, (size_t)(q - p)))
                    return -1;
            } else if (q > p) {
                OPEN_TEXT(); /* text outside any field still splits nothing we keep */
            }
            p = q;
            if (p < end && *p == '&') {
                const char *semi = memchr(p, ';', (size_t)(end - p) < 16 ? (size_t)(end - p) : 16);
                if (!semi)
                    return xml_error(x, "malformed entity reference");
                size_t n = (size_t)(semi - p - 1);
                const char *ent = p + 1;
                OPEN_TEXT();
                if (in_rec && cur_field) {
                    if (name_is(ent, n, "lt")) arena_put(x, "<", 1);
                    else if (name_is(ent, n, "gt")) arena_put(x, ">", 1);
                    else if (name_is(ent, n, "amp")) arena_put(x, "&", 1);
                    else if (name_is(ent, n, "quot")) arena_put(x, "\"", 1);
                    else if (name_is(ent, n, "apos")) arena_put(x, "'", 1);
                    else if (n > 1 && ent[0] == '#') {
                        unsigned long cp = (ent[1] == 'x') ? strtoul(ent + 2, NULL, 16) : strtoul(ent + 1, NULL, 10);
                        utf8_put(x, cp);
                    } else
                        return xml_error(x, "unknown entity reference");
                }
                p = semi + 1;
            }
            continue;
        }
        /* markup */
        if (p + 4 <= end && memcmp(p, "<!--", 4) == 0) {
            CLOSE_TEXT();
            p = skip_until(p + 4, end, "-->");
            if (!p)
                return xml_error(x, "unterminated comment");
            continue;
        }
        if (p + 9 <= end && memcmp(p, "<![CDATA[", 9) == 0) {
            CLOSE_TEXT();
            const char *q = skip_until(p + 9, end, "]]>");
            if (!q)
                return xml_error(x, "unterminated CDATA section");
            if (in_rec && cur_field) {
                span_t s;
                s.begin = x->arena_len;
                if (arena_put(x, p + 9, (size_t)(q - 3 - (p + 9))))
                    return -1;
                s.end = x->arena_len;
                if (push_text(x, cur_field, s))
                    return -1;
            }
            p = q;
            continue;
        }
        if (p + 2 <= end && p[1] == '?') {
            CLOSE_TEXT();
            p = skip_until(p + 2, end, "?>");
            if (!p)
                return xml_error(x, "unterminated processing instruction");
            continue;
        }
        if (p + 2 <= end && p[1] == '!') {
            CLOSE_TEXT();
            p = skip_doctype(p + 2, end);
            if (!p)
                return xml_error(x, "unterminated declaration");
            continue;
        }
        CLOSE_TEXT();
        int closing = (p + 1 < end && p[1] == '/');
        const char *nm = p + (closing ? 2 : 1);
        const char *q = nm;
        while (q < end && *q != '>' && *q != '/' && *q != ' ' && *q != '\t' && *q != '\n' && *q != '\r')
            ++q;
        size_t nlen = (size_t)(q - nm);
        /* find the end of the tag, honouring quoted attribute values */
        const char *t = q;
        int self_closing = 0;
        while (t < end && *t != '>') {
            if (*t == '"' || *t == '\'') {
                const char *r = memchr(t + 1, *t, (size_t)(end - t - 1));
                if (!r)
                    return xml_error(x, "unterminated attribute value");
                t = r + 1;
                continue;
            }
            ++t;
        }
        if (t >= end)
            return xml_error(x, "unterminated tag");
        if (!closing && t > p && t[-1] == '/')
            self_closing = 1;
        p = t + 1;

        if (closing) {
            if (depth == 0)
                return xml_error(x, "unbalanced closing tag");
            depth--;
            if (stack[depth].opens_rec)
                in_rec = 0;
            /* text() selects direct text children only */
            cur_field = depth > 0 ? stack[depth - 1].field : 0;
            continue;
        }
        if (depth >= 255)
            return xml_error(x, "elements nested too deeply");
        elem_t *el = &stack[depth];
        memset(el, 0, sizeof(*el));
        if (nlen < sizeof(el->name)) {
            memcpy(el->name, nm, nlen);
            el->name[nlen] = '\0';
        }
        el->is_network = name_is(nm, nlen, "NETWORK");
        int parent_is_network = depth > 0 && stack[depth - 1].is_network;
        if (parent_is_network && !in_rec && (name_is(nm, nlen, "VARIABLE") || name_is(nm, nlen, "DEFINITION"))) {
            if (push_rec(x, name_is(nm, nlen, "DEFINITION")))
                return -1;
            el->opens_rec = 1;
            in_rec = 1;
        }
        if (name_is(nm, nlen, "GIVEN") && depth >= 2 && stack[depth - 1].opens_rec &&
            x->recs[x->nrecs - 1].is_def && stack[depth - 2].is_network)
            x->given_direct++;
        if (in_rec) {
            if (name_is(nm, nlen, "NAME")) el->field = F_NAME;
            else if (name_is(nm, nlen, "OUTCOME")) el->field = F_OUTCOME;
            else if (name_is(nm, nlen, "FOR")) el->field = F_FOR;
            else if (name_is(nm, nlen, "GIVEN")) el->field = F_GIVEN;
            else if (name_is(nm, nlen, "TABLE")) el->field = F_TABLE;
        }
        if (self_closing) {
            if (el->opens_rec)
                in_rec = 0;
            continue;
        }
        depth++;
        cur_field = el->field;
    }
    CLOSE_TEXT();
    if (depth != 0)
        return xml_error(x, "document ends inside an element");
    return 0;
#undef CLOSE_TEXT
#undef OPEN_TEXT
}

/* copy a text node into a CHAR_BUFFER_SIZE buffer with strncpy semantics */
static void copy_name(const xml_t *x, span_t s, char *out)
{
    size_t n = (size_t)(s.end - s.begin);
    memset(out, 0, CHAR_BUFFER_SIZE);
    /* the text stops at an embedded NUL like a C string would */
    const char *src = x->arena + s.begin;
    size_t len = strnlen(src, n);
    memcpy(out, src, len < CHAR_BUFFER_SIZE ? len : CHAR_BUFFER_SIZE);
}

static int is_sep(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }

/* count and convert the tokens of a TABLE text node */
static uint64_t table_tokens(const xml_t *x, span_t s, float *out)
{
    uint64_t count = 0;
    const char *p = x->arena + s.begin, *end = x->arena + s.end;
    char tok[128];
    while (p < end) {
        while (p < end && is_sep(*p))
            ++p;
        if (p >= end)
            break;
        const char *q = p;
        while (q < end && !is_sep(*q))
            ++q;
        if (out) {
            size_t n = (size_t)(q - p);
            if (n >= sizeof(tok))
                n = sizeof(tok) - 1;
            memcpy(tok, p, n);
            tok[n] = '\0';
            char *stop;
            float v = strtof(tok, &stop);
            if (stop != tok)
                out[count] = v; /* a token sscanf cannot read leaves the entry as it was (0) */
        }
        count++;
        p = q;
    }
    return count;
}

Graph_t parse_xml_file(const char *file_name)
{
    bp_clear_error();
    FILE *f = fopen(file_name, "rb");
    if (!f) {
        bp_fail("parse_xml_file: cannot open %s: %s", file_name, strerror(errno));
        return NULL;
    }
    struct stat st;
    if (fstat(fileno(f), &st) != 0) {
        fclose(f);
        bp_fail("parse_xml_file: cannot stat %s", file_name);
        return NULL;
    }
    size_t size = (size_t)st.st_size;
    char *buf = (char *)malloc(size + 1);
    if (!buf || fread(buf, 1, size, f) != size) {
        free(buf);
        fclose(f);
        bp_fail("parse_xml_file: cannot read %s", file_name);
        return NULL;
    }
    fclose(f);
    buf[size] = '\0';

    xml_t x;
    memset(&x, 0, sizeof(x));
    x.p = buf;
    x.end = buf + size;
    Graph_t g = NULL;
    float *table = NULL;
    unsigned int *givens = NULL;
    uint64_t table_cap = 0, givens_cap = 0;
    if (scan(&x) != 0)
        goto done;

    unsigned int V = 0;
    for (uint64_t r = 0; r < x.nrecs; ++r)
        if (!x.recs[r].is_def)
            V++;
    if (x.given_direct * 2 > 0xffffffffull) {
        bp_fail("parse_xml_file: too many GIVEN elements");
        goto done;
    }
    g = create_graph(V, (unsigned int)(x.given_direct * 2));
    if (!g)
        goto done;

    /* nodes, document order */
    for (uint64_t r = 0; r < x.nrecs; ++r) {
        const rec_t *rc = &x.recs[r];
        if (rc->is_def)
            continue;
        unsigned int v = g->current_num_vertices;
        unsigned int outcomes = 0, names = 0;
        char name[CHAR_BUFFER_SIZE];
        for (uint64_t i = rc->first; i < rc->first + rc->count; ++i) {
            const trec_t *t = &x.texts[i];
            if (t->field == F_OUTCOME) {
                if (outcomes < MAX_STATES)
                    copy_name(&x, t->text,
                              &g->variable_names[(size_t)v * MAX_STATES * CHAR_BUFFER_SIZE +
                                                 (size_t)outcomes * CHAR_BUFFER_SIZE]);
                outcomes++;
            } else if (t->field == F_NAME) {
                if (names == 0)
                    copy_name(&x, t->text, name);
                names++;
            }
        }
        if (names != 1) {
            bp_fail("parse_xml_file: VARIABLE %u has %u NAME text nodes (expected 1)", v, names);
            goto fail;
        }
        if (outcomes == 0 || outcomes > MAX_STATES) {
            bp_fail("parse_xml_file: VARIABLE '%.*s' has %u outcomes (1..%d supported)", CHAR_BUFFER_SIZE, name,
                    outcomes, MAX_STATES);
            goto fail;
        }
        graph_add_node(g, outcomes, name);
        if (bp_has_error())
            goto fail;
    }

    /* two passes over the definitions */
    for (int pass = 1; pass <= 2; ++pass) {
        for (uint64_t r = 0; r < x.nrecs; ++r) {
            const rec_t *rc = &x.recs[r];
            if (!rc->is_def)
                continue;
            uint64_t ngiven = 0, nfor = 0, ntable = 0;
            span_t for_text = {0, 0}, table_text = {0, 0};
            for (uint64_t i = rc->first; i < rc->first + rc->count; ++i) {
                const trec_t *t = &x.texts[i];
                if (t->field == F_GIVEN)
                    ngiven++;
                else if (t->field == F_FOR) {
                    for_text = t->text;
                    nfor++;
                } else if (t->field == F_TABLE) {
                    table_text = t->text; /* the last one is used */
                    ntable++;
                }
            }
            if ((pass == 1) != (ngiven == 0))
                continue;
            if (nfor != 1) {
                bp_fail("parse_xml_file: DEFINITION has %llu FOR text nodes (expected 1)", (unsigned long long)nfor);
                goto fail;
            }
            if (ntable == 0) {
                bp_fail("parse_xml_file: DEFINITION without TABLE text");
                goto fail;
            }
            char name[CHAR_BUFFER_SIZE + 1];
            copy_name(&x, for_text, name);
            name[CHAR_BUFFER_SIZE] = '\0';
            unsigned int child = find_node_by_name(name, g);
            if (bp_has_error())
                goto fail;
            uint64_t n = table_tokens(&x, table_text, NULL);
            if (n + 1 > table_cap) {
                table_cap = n + 1;
                free(table);
                table = (float *)malloc(sizeof(float) * table_cap);
                if (!table) {
                    bp_fail("parse_xml_file: out of host memory");
                    goto fail;
                }
            }
            memset(table, 0, sizeof(float) * (n + 1));
            table_tokens(&x, table_text, table);
            if (pass == 1) {
                if (bp_graph_add_observation(g, child, table, (unsigned int)(n < 0xffffffffu ? n : 0xffffffffu)))
                    goto fail;
                continue;
            }
            if (ngiven > givens_cap) {
                givens_cap = ngiven;
                free(givens);
                givens = (unsigned int *)malloc(sizeof(unsigned int) * givens_cap);
                if (!givens) {
                    bp_fail("parse_xml_file: out of host memory");
                    goto fail;
                }
            }
            uint64_t k = 0;
            for (uint64_t i = rc->first; i < rc->first + rc->count; ++i) {
                const trec_t *t = &x.texts[i];
                if (t->field != F_GIVEN)
                    continue;
                copy_name(&x, t->text, name);
                givens[k++] = find_node_by_name(name, g);
                if (bp_has_error())
                    goto fail;
            }
            if (bp_graph_add_cpt_edges(g, child, givens, (unsigned int)ngiven, table, n))
                goto fail;
        }
    }
    goto done;
fail:
    graph_destroy(g);
    g = NULL;
done:
    free(table);
    free(givens);
    free(x.arena);
    free(x.texts);
    free(x.recs);
    free(buf);
    return g;
}
