"""Expression programs: nonlinear programs whose rows are scalar expression
DAGs over x in a box, in the node format of include/gearhip_program.h.

Row 0 is the objective (minimised), rows 1.. are constraints f_r(x) <= rhs_r.
`Program` builds the node list; `nlp_n8_m3_problem` and `uneven_nlp_instances`
are the seeded synthetic generators of the small NLP and the uneven batch.
"""
from dataclasses import dataclass, field

import numpy as np

INPUT, CONST, ADD, SUB, NEG, MULC, ADDC, MUL, SQR, POW, EXP, LOG, RECIP, SIN, COS = range(15)
OP_NAMES = ["input", "const", "add", "sub", "neg", "mulc", "addc", "mul", "sqr", "pow", "exp",
            "log", "recip", "sin", "cos"]
RELAXED = {MUL, SQR, POW, EXP, LOG, RECIP, SIN, COS}
BINARY = {ADD, SUB, MUL}

NODE_DTYPE = np.dtype([("op", "<i4"), ("a", "<i4"), ("b", "<i4"), ("pad", "<i4"), ("c", "<f8")])


class Program:
    def __init__(self, n, name="program"):
        self.n = n
        self.name = name
        self.nodes = [(INPUT, i, 0, 0.0) for i in range(n)]
        self.out = []
        self.rhs = []
        self.xl = None
        self.xu = None

    # -- builders
    def _node(self, op, a=0, b=0, c=0.0):
        self.nodes.append((op, int(a), int(b), float(c)))
        return len(self.nodes) - 1

    def x(self, i):
        return int(i)

    def const(self, c):
        return self._node(CONST, c=c)

    def add(self, a, b):
        return self._node(ADD, a, b)

    def sub(self, a, b):
        return self._node(SUB, a, b)

    def neg(self, a):
        return self._node(NEG, a)

    def mulc(self, a, c):
        return self._node(MULC, a, c=c)

    def addc(self, a, c):
        return self._node(ADDC, a, c=c)

    def mul(self, a, b):
        return self._node(MUL, a, b)

    def sqr(self, a):
        return self._node(SQR, a)

    def pow(self, a, p):
        """a ** p for a constant p: integers >= 1 (x ** 2 is SQR, as
        convert_sqr lowers it) and non-integers p > 0 (on a >= 0)."""
        p = float(p)
        if p == 2:
            return self.sqr(a)
        if p == 1:
            return a
        if not (p > 0) or p > 31 or (p == int(p) and p < 3):
            raise ValueError("exponents: integers 1 .. 31 or non-integers p > 0")
        return self._node(POW, a, c=p)

    def exp(self, a):
        return self._node(EXP, a)

    def log(self, a):
        return self._node(LOG, a)

    def recip(self, a):
        return self._node(RECIP, a)

    def sin(self, a):
        return self._node(SIN, a)

    def cos(self, a):
        return self._node(COS, a)

    def div(self, a, b):
        """a / b lowered as MUL(a, RECIP(b)) (optimize_graph.py:103-122)."""
        if self.nodes[b][0] == CONST:
            return self.mulc(a, 1.0 / self.nodes[b][3])
        return self.mul(a, self.recip(b))

    def sum(self, terms):
        acc = terms[0]
        for t in terms[1:]:
            acc = self.add(acc, t)
        return acc

    def output(self, node, rhs=0.0):
        self.out.append(int(node))
        self.rhs.append(float(rhs))
        return len(self.out) - 1

    # -- views
    @property
    def V(self):
        return len(self.nodes)

    @property
    def R(self):
        return len(self.out)

    def arrays(self):
        nodes = np.zeros(self.V, dtype=NODE_DTYPE)
        for k, (op, a, b, c) in enumerate(self.nodes):
            nodes[k] = (op, a, b, 0, c)
        return nodes, np.asarray(self.out, dtype=np.int32)

    def eval(self, X):
        """Row values at points X (N, n), FP64 numpy."""
        X = np.atleast_2d(np.asarray(X, dtype=np.float64))
        Y = [None] * self.V
        for i in range(self.n):
            Y[i] = X[:, i]
        with np.errstate(all="ignore"):
            for k in range(self.n, self.V):
                op, a, b, c = self.nodes[k]
                A = Y[a]
                if op == CONST:
                    y = np.full(X.shape[0], c)
                elif op == ADD:
                    y = A + Y[b]
                elif op == SUB:
                    y = A - Y[b]
                elif op == NEG:
                    y = -A
                elif op == MULC:
                    y = A * c
                elif op == ADDC:
                    y = A + c
                elif op == MUL:
                    y = A * Y[b]
                elif op == SQR:
                    y = A * A
                elif op == POW:
                    y = A ** c
                elif op == EXP:
                    y = np.exp(A)
                elif op == LOG:
                    y = np.log(A)
                elif op == RECIP:
                    y = 1.0 / A
                elif op == SIN:
                    y = np.sin(A)
                elif op == COS:
                    y = np.cos(A)
                else:
                    raise ValueError(op)
                Y[k] = y
        return np.stack([Y[o] for o in self.out], axis=1)

    def row_subprograms(self):
        """Per row: (local nodes, local->global input map).  A row's local
        list is the inputs it reads (increasing index) followed by its
        operation nodes in global order, so a backward pass visits the same
        nodes in the same order as on the global program."""
        subs = []
        for o in self.out:
            seen = set()
            stack = [o]
            while stack:
                v = stack.pop()
                if v in seen:
                    continue
                seen.add(v)
                op, a, b, c = self.nodes[v]
                if op == INPUT or op == CONST:
                    continue
                stack.append(a)
                if op in BINARY:
                    stack.append(b)
            ins = sorted(v for v in seen if v < self.n)
            ops = sorted(v for v in seen if v >= self.n)
            order = ins + ops
            local = {g: k for k, g in enumerate(order)}
            nodes = np.zeros(len(order), dtype=NODE_DTYPE)
            for k, g in enumerate(order):
                op, a, b, c = self.nodes[g]
                if op == INPUT:
                    nodes[k] = (INPUT, a, 0, 0, 0.0)
                else:
                    la = local[a] if op != CONST else 0
                    lb = local[b] if op in BINARY else 0
                    nodes[k] = (op, la, lb, 0, c)
            subs.append((nodes, np.asarray(ins, dtype=np.int32), local[o]))
        return subs


# ---------------------------------------------------------------- generators
def _quadratic(P, Q, c=None, d=0.0):
    """x^T Q x (+ c^T x + d) as scalar products: Q_ii x_i^2 and
    (Q_ij + Q_ji) x_i x_j for i < j."""
    n = Q.shape[0]
    terms = []
    for i in range(n):
        terms.append(P.mulc(P.sqr(i), Q[i, i]))
        for j in range(i + 1, n):
            terms.append(P.mulc(P.mul(i, j), Q[i, j] + Q[j, i]))
    if c is not None:
        for i in range(n):
            terms.append(P.mulc(i, c[i]))
    acc = P.sum(terms)
    if d:
        acc = P.addc(acc, d)
    return acc


def nlp_n8_m3_problem(seed=1):
    """n = 8, x in [-2, 2]^8.  f0 = x^T Q x + sum_i c_i exp(0.3 x_i) + 4
    bilinear x_a x_b (a != b); Q ~ N(0,1), c ~ U(-1,1).  Three constraints
    x^T A_k x + b_k^T x + d_k <= 1 with N(0,1) coefficients."""
    rng = np.random.default_rng(seed)
    n = 8
    P = Program(n, name="nlp_n8_m3")
    Q = rng.standard_normal((n, n))
    cc = rng.uniform(-1, 1, n)
    pairs = []
    while len(pairs) < 4:
        a, b = rng.integers(0, n, 2)
        if a != b:
            pairs.append((int(a), int(b)))
    terms = [_quadratic(P, Q)]
    for i in range(n):
        terms.append(P.mulc(P.exp(P.mulc(i, 0.3)), cc[i]))
    for a, b in pairs:
        terms.append(P.mul(a, b))
    P.output(P.sum(terms))
    for _ in range(3):
        A = rng.standard_normal((n, n))
        b = rng.standard_normal(n)
        d = rng.standard_normal()
        P.output(_quadratic(P, A, b, d), rhs=1.0)
    P.xl = -2.0 * np.ones(n)
    P.xu = 2.0 * np.ones(n)
    return P


def _random_expr(P, rng, depth, n):
    if depth == 0:
        if rng.random() < 0.8:
            return int(rng.integers(n))
        return P.const(float(rng.standard_normal()))
    k = rng.integers(7)
    if k == 0:
        return P.add(_random_expr(P, rng, depth - 1, n), _random_expr(P, rng, depth - 1, n))
    if k == 1:
        return P.mul(_random_expr(P, rng, depth - 1, n), _random_expr(P, rng, depth - 1, n))
    if k == 2:
        # a / (1 + b^2): a quotient whose denominator stays positive
        den = P.addc(P.sqr(_random_expr(P, rng, depth - 1, n)), 1.0)
        return P.div(_random_expr(P, rng, depth - 1, n), den)
    if k == 3:
        return P.pow(_random_expr(P, rng, depth - 1, n), int(rng.integers(2, 5)))
    if k == 4:
        return P.exp(P.mulc(_random_expr(P, rng, depth - 1, n), 0.3))
    if k == 5:
        return P.log(P.addc(P.sqr(_random_expr(P, rng, depth - 1, n)), 1.0))
    return P.sin(_random_expr(P, rng, depth - 1, n))


def random_nlp(n, m, seed, terms_per_row=3):
    """Rows are sums of random expressions of depth U{2..5} over
    {+, *, /, pow(2..4), exp, log(1 + x^2), sin}; leaves x_j or N(0,1)
    constants; box widths U[0.5, 4] centred at N(0,1)."""
    rng = np.random.default_rng(seed)
    P = Program(n, name=f"dag_n{n}_m{m}_s{seed}")
    for r in range(m + 1):
        terms = [_random_expr(P, rng, int(rng.integers(2, 6)), n) for _ in range(terms_per_row)]
        P.output(P.sum(terms), rhs=0.0 if r == 0 else float(abs(rng.standard_normal())))
    centre = rng.standard_normal(n)
    width = rng.uniform(0.5, 4.0, n)
    P.xl = centre - width / 2
    P.xu = centre + width / 2
    return P


def uneven_nlp_instances(count=256, seed=4):
    """Uneven batch: n = clip(round(lognormal(ln 40, 0.8)), 5, 200), m = n // 4."""
    rng = np.random.default_rng(seed)
    ns = np.clip(np.round(rng.lognormal(np.log(40.0), 0.8, size=count)), 5, 200).astype(int)
    return [random_nlp(int(n), int(n) // 4, seed=1000 * seed + k) for k, n in enumerate(ns)]
