/*
 * xml-expression.h -- drop-in XMLBIF reader (src/bnf-xml-parser/
 * xml-expression.h:9).  Self-contained: no libxml2 dependency.
 *
 * parse_xml_file() builds the same Graph_t as the reference
 * (src/bnf-xml-parser/xml-expression.c:441-466): nodes in document order,
 * edge capacity 2 x #GIVEN, pass 1 sets priors / evidence (marking the node
 * observed), pass 2 emits pairwise tables for every GIVEN, last GIVEN first,
 * with a transposed reverse edge unless the parent is observed.
 */
#ifndef BP_XML_EXPRESSION_H_
#define BP_XML_EXPRESSION_H_

#include "graph.h"

#ifdef __cplusplus
extern "C" {
#endif

Graph_t parse_xml_file(const char *file_name);

#ifdef __cplusplus
}
#endif

#endif /* BP_XML_EXPRESSION_H_ */
