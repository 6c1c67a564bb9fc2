/*
 * Expression-program format of the gearhip C ABI (plain C, no HIP types).
 *
 * A program is a list of scalar nodes in topological order over n inputs:
 * nodes 0..n-1 are the inputs x_i (op GH_OP_INPUT, a = i), every later node
 * reads only earlier nodes.  Row r of the specification is node out[r].
 * The node vocabulary is the subset of the reference's operator set that
 * nonlinear programs lower to (auto_LiRPA/operators: add_sub.py, bivariate.py,
 * activations.py BoundSqr, s_shaped.py BoundPow, convex_concave.py BoundExp /
 * BoundLog / BoundReciprocal, trigonometric.py BoundSin / BoundCos); division is
 * MUL(a, RECIP(b)) as optimize_graph.py:103-122 lowers it.
 */
#ifndef GEARHIP_PROGRAM_H
#define GEARHIP_PROGRAM_H

#include <stdint.h>

enum {
    GH_OP_INPUT = 0, /* x_a */
    GH_OP_CONST = 1, /* c */
    GH_OP_ADD = 2,   /* a + b */
    GH_OP_SUB = 3,   /* a - b */
    GH_OP_NEG = 4,   /* -a */
    GH_OP_MULC = 5,  /* a * c */
    GH_OP_ADDC = 6,  /* a + c */
    GH_OP_MUL = 7,   /* a * b, both perturbed */
    GH_OP_SQR = 8,   /* a * a */
    GH_OP_POW = 9,   /* a ** c: c an integer >= 3, or a non-integer c > 0 (on a >= 0) */
    GH_OP_EXP = 10,
    GH_OP_LOG = 11,
    GH_OP_RECIP = 12, /* 1 / a */
    GH_OP_SIN = 13,
    GH_OP_COS = 14
};

typedef struct {
    int32_t op, a, b, pad;
    double c;
} gh_node;

typedef struct {
    int32_t n;               /* inputs */
    int32_t V;               /* nodes (including the n inputs) */
    int32_t R;               /* rows */
    int32_t pad;
    const gh_node *nodes;    /* V */
    const int32_t *out;      /* R */
} gh_prog;

#endif /* GEARHIP_PROGRAM_H */
