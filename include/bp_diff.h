/*
 * bp_diff.h -- differentiable loopy belief propagation on the GPU (C-ABI of
 * libbp_amd.so).  No counterpart in the reference: it differentiates the
 * reference's unrolled loopy recursion (src/graph/graph.c:725-751 and
 * 1121-1183) with respect to the conditional probability tables that the
 * XMLBIF reader turns into pairwise edge tables (src/bnf-xml-parser/
 * xml-expression.c:290-419).
 *
 * Model.  A bp_bn whose first num_param_defs definitions are the
 * parameters (one CPT per variable, root priors included) and whose later
 * definitions are GIVEN-less observations (evidence, constants).  The
 * parameter vector is the concatenation of those definitions' TABLEs in
 * bp_bn order, child-state-major as in the XMLBIF file.  Setting it
 * rebuilds, on the device, exactly what the reader would build from those
 * numbers: each pairwise table is the reader's sequential f32 sum, and a
 * root prior becomes the root's initial belief unless a later observation
 * overrides it.
 *
 * Forward.  init_previous_edge + K iterations of the node-centric schedule
 * (bit-identical beliefs to bp_loopy_iterate on the same tables), keeping
 * the beliefs and combine buffers of every c-th iteration.
 *
 * Backward.  Reverse-mode adjoint of the same recursion, in exact
 * arithmetic terms:
 *   m_t[s] = N(J_s . buf_{t-1}[src s]),  T_t[v] = prod_{s in in(v)} m_t[s],
 *   st_t[v] = N(st_{t-1}[v] (.) T_t[v]),  buf_t[v] = st_t[v] (.) T_t[v],
 * with m_0 and buf_0 from st_0 (init).  Iterates between checkpoints are
 * recomputed by the forward kernel, segment by segment.  Returns dL/dTABLE
 * for every parameter entry given dL/dbeliefs after iteration K.
 *
 * Limits: every node has the same state count, and no node has more than
 * 128 in-edges (64 for more than 4 states); bp_diff_create reports others.
 * NaN factors, which the forward skips, are treated as absent.
 *
 * Device pointers are plain device addresses; stream is a hipStream_t
 * (NULL: the null stream).  Every function returns 0, or -1 with the
 * message in bp_last_error().
 */
#ifndef BP_DIFF_H_
#define BP_DIFF_H_

#include <stdint.h>

#include "bp_model.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bp_diff bp_diff;

int bp_diff_create(const bp_bn *bn, unsigned int num_param_defs, int device, bp_diff **out);
void bp_diff_free(bp_diff *d);

/* info[0]=V, info[1]=E (message edges), info[2]=S (states per node),
 * info[3]=belief row stride, info[4]=parameter entries,
 * info[5]=parameter rows (definition x parent configuration) */
int bp_diff_info(const bp_diff *d, uint64_t *info);

/* The softmax rows of the parameter vector: entry k < row_len[r] of row r
 * is parameter row_first[r] + k * row_stride[r] (host arrays of info[5]). */
int bp_diff_rows(const bp_diff *d, uint64_t *row_first, uint32_t *row_stride, uint32_t *row_len);

/* Pairwise tables and initial beliefs from the parameter vector (device,
 * info[4] floats). */
int bp_diff_set_params(bp_diff *d, const float *params, void *stream);

/* The same from log-parameters: each row of bp_diff_rows gets the softmax
 * of its entries of theta (device, info[4] floats), and those are the
 * parameters. */
int bp_diff_set_log_params(bp_diff *d, const float *theta, void *stream);

/* K iterations from the initial beliefs, checkpoint interval c >= 1;
 * beliefs (device, [V][stride]) receives the beliefs after iteration K. */
int bp_diff_forward(bp_diff *d, unsigned int K, unsigned int c, float *beliefs, void *stream);

/* dL/dparams (device, info[4] floats, overwritten) from dL/dbeliefs
 * (device, [V][stride]) of the last bp_diff_forward. */
int bp_diff_backward(bp_diff *d, const float *grad_beliefs, float *grad_params, void *stream);

/* dL/dtheta (device, info[4] floats) after bp_diff_set_log_params: the
 * adjoint, then the softmax's transpose per row. */
int bp_diff_backward_log(bp_diff *d, const float *grad_beliefs, float *grad_theta, void *stream);

/* Same as bp_diff_backward, also returning dL/dJ per message edge (device,
 * [E][S*S] in the reference's edge order, entry [i][j] with i the source
 * state) and dL/d(initial beliefs) (device, [V][stride]); either may be
 * NULL. */
int bp_diff_backward_ex(bp_diff *d, const float *grad_beliefs, float *grad_params, float *grad_tables,
                        float *grad_initial, void *stream);

/* The pairwise tables of the current parameters in the reference's edge
 * order, [E][S*S] host floats (for checks against the reader). */
int bp_diff_tables(bp_diff *d, float *out);

#ifdef __cplusplus
}
#endif

#endif /* BP_DIFF_H_ */
