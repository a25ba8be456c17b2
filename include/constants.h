/*
 * constants.h -- compile-time limits of the belief-propagation drop-in.
 *
 * Values are part of the binary interface: every per-node vector in
 * `struct graph` is padded to MAX_STATES floats and every per-edge table to
 * MAX_STATES x MAX_STATES floats, exactly as in the reference
 * (src/constants.h:9-41).  Callers that index `node_states[MAX_STATES*v+k]`
 * or pass MAX_STATES-strided tables to graph_add_edge() rely on them.
 */
#ifndef BP_CONSTANTS_H_
#define BP_CONSTANTS_H_

#define CHAR_BUFFER_SIZE 50   /* bytes per node / state name (src/constants.h:9) */
#define MAX_STATES 10         /* padding of every state vector (src/constants.h:11) */
#define DEFAULT_STATE 1.0     /* initial state entry of an unobserved node (:13) */
#define WEIGHT_INFINITY 99999 /* diameter sentinel (:15) */
#define BATCH_SIZE 10         /* iterations between convergence checks in the GPU driver (:17) */
#define NUM_ITERATIONS 200    /* default max loopy iterations used by the drivers (:19) */
#define PRECISION 1E-3f       /* default convergence threshold used by the drivers (:21) */
#define MAX_DEGREE 20         /* per-node edge-list limit of the reference; see graph.h (:39) */
#define CHARS_IN_KEY 20       /* (:41) */

#endif /* BP_CONSTANTS_H_ */
