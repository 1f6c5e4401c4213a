"""Known answers fixed by mathematics for the multiple-precision LDL^T and
the sampled arrow factor, with their own rational arithmetic (Python
Fractions; no code shared with the oracle or the library).

Each case states its expected bits from a closed form, so a mistake in a
restatement (a dropped term, a wrong index or transpose, a different
rounding sequence) fails it:

* exact_ldl_case: A = L D L^T with L unit lower of small integers (an
  envelope with a cut and ragged rows) and D = +-2^e, so every rounding of
  any elimination order is exact and the factor must be L, D^-1 exactly;
* rounding_order_case: an order-3 matrix whose ldl_dots entries, evaluated
  as RN(RN(a - RN(dot)) * RN(1/d)) (arrow.rs:177-182), differ from the
  once-rounded quotient RN((a - dot) / d);
* zero_pivot_case: a pivot that RN(dot) cancels to zero while the exact
  difference is -2^-(2P-2) (the exact retry of arrow.rs:185-197);
* rec_retry_case: order 256, where one rounding per level of ldl_rec
  cancels pivot 128 to an exact zero, so factor_loaded redoes the factor by
  exact dots (arrow.rs:819-846);
* analytic_sampled_case: the reference's analytic_schur block
  (sampled/tests.rs:407-434: identity R, basis [1, 1; 1, -1], weights
  (2, -3)), and its delta-perturbed form, through the leaf factor and one
  border row.
"""
from fractions import Fraction

import numpy as np

from sdpx_gpu import mp


# ---------------------------------------------------------------------------
# rational rounding and record encoding
# ---------------------------------------------------------------------------
def rn(x, p):
    """x (Fraction) rounded to nearest, ties to even, at p significant bits."""
    x = Fraction(x)
    if x == 0:
        return Fraction(0)
    s = -1 if x < 0 else 1
    x = abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    if Fraction(2) ** (e + 1) <= x:
        e += 1
    # 2^e <= x < 2^(e+1): keep p bits, ulp 2^(e - p + 1)
    scaled = x / Fraction(2) ** (e - p + 1)
    q, r = divmod(scaled.numerator, scaled.denominator)
    if 2 * r > scaled.denominator or (2 * r == scaled.denominator and q & 1):
        q += 1
    return s * Fraction(q) * Fraction(2) ** (e - p + 1)


def enc(values, nl):
    """Exactly representable Fractions -> records (raises if not exact)."""
    p = 64 * nl
    out = mp.zeros(len(values), nl)
    for idx, v in enumerate(values):
        v = Fraction(v)
        if v == 0:
            continue
        a = abs(v)
        e = a.numerator.bit_length() - a.denominator.bit_length()
        if Fraction(2) ** e > a:
            e -= 1
        if Fraction(2) ** (e + 1) <= a:
            e += 1
        m = a / Fraction(2) ** (e + 1 - p)  # in [2^(p-1), 2^p)
        assert m.denominator == 1, "value not representable"
        mi = m.numerator
        out[idx]["limbs"] = np.array([(mi >> (64 * i)) & ((1 << 64) - 1) for i in range(nl)], np.uint64)
        out[idx]["kind"] = -3 if v < 0 else 3
        out[idx]["exponent"] = e + 1
    return out


def dec(rec):
    """Records -> Fractions (regular and zero values only)."""
    nl = mp.limbs_of(rec)
    out = []
    for r in np.atleast_1d(rec):
        k = int(r["kind"])
        if abs(k) == 2:
            out.append(Fraction(0))
            continue
        assert abs(k) == 3, "non-finite value"
        m = sum(int(r["limbs"][i]) << (64 * i) for i in range(nl))
        v = Fraction(m) * Fraction(2) ** (int(r["exponent"]) - 64 * nl)
        out.append(-v if k < 0 else v)
    return out


def col_major(mat):
    """n x n list of Fractions [i][j] -> column-major flat list."""
    n = len(mat)
    return [mat[i][j] for j in range(n) for i in range(n)]


def lower_of(rec, n):
    """The strict lower triangle and diagonal of an n x n column-major
    record array as Fractions, [i][j] for j <= i."""
    v = dec(rec)
    return [[v[i + j * n] for j in range(i + 1)] for i in range(n)]


# ---------------------------------------------------------------------------
# cases
# ---------------------------------------------------------------------------
def exact_ldl_case(n):
    """(A rows [i][j] Fractions, L [i][j], D list, signs): A = L D L^T with
    integer L in {-1, 0, 1}, D = +-2^e (|e| <= 20); rows from 140 start at
    column 140 (a cut there), rows from 260 at 140 + i mod 11."""
    def first(i):
        if i < 140:
            return 0
        if i < 260:
            return 140
        return 140 + i % 11

    L = [[Fraction(0)] * n for _ in range(n)]
    for i in range(n):
        L[i][i] = Fraction(1)
        for j in range(min(first(i), i), i):
            L[i][j] = Fraction(((i * 7 + j * 3) % 3) - 1)
    D = [Fraction(2) ** (((k * 5) % 41) - 20) * (-1 if k % 4 == 3 else 1) for k in range(n)]
    signs = [-1 if k % 4 == 3 else 1 for k in range(n)]
    A = [[Fraction(0)] * n for _ in range(n)]
    for i in range(n):
        for j in range(i + 1):
            A[i][j] = sum(L[i][k] * L[j][k] * D[k] for k in range(j + 1))
            A[j][i] = A[i][j]
    return A, L, D, signs


def ldl_dots_rational(A, p):
    """ldl_dots (arrow.rs:156-253) in rational arithmetic at p bits for a
    dense (full-envelope) matrix: pivot d_j = RN(a_jj - RN(sum r w)) (exact
    difference rounded once when that is zero), dinv_j = RN(1/d_j), entry
    r_ij = RN(RN(a_ij - RN(sum r_ik w_jk)) * dinv_j), w_ij = RN(r_ij d_j)."""
    n = len(A)
    r = [[None] * n for _ in range(n)]
    w = [[None] * n for _ in range(n)]
    dinv = [None] * n
    for j in range(n):
        exact = sum(r[j][k] * w[j][k] for k in range(j))
        d = rn(A[j][j] - rn(exact, p), p)
        if d == 0 and j > 0:
            d = rn(A[j][j] - exact, p)
        assert d != 0
        dinv[j] = rn(1 / d, p)
        for i in range(j + 1, n):
            dot = rn(sum(r[i][k] * w[j][k] for k in range(j)), p)
            r[i][j] = rn(rn(A[i][j] - dot, p) * dinv[j], p)
            w[i][j] = rn(r[i][j] * d, p)
    return r, dinv


def rounding_order_case(p):
    """Order 3: (A, expected L [i][j] for j < i, expected dinv), where the
    sequence of roundings matters at 128, 256 and 512 bits: L_10 =
    RN(5 RN(1/3)) is not RN(5/3), and L_21 differs from the exactly
    eliminated quotient rounded once."""
    A = [[Fraction(3), Fraction(5), Fraction(1)],
         [Fraction(5), Fraction(13), Fraction(2)],
         [Fraction(1), Fraction(2), Fraction(50)]]
    r, dinv = ldl_dots_rational(A, p)
    # the naive once-rounded quotients (exact elimination, one rounding)
    d0 = A[0][0]
    naive10 = rn(A[1][0] / d0, p)
    d1_exact = A[1][1] - A[1][0] * A[1][0] / d0
    naive21 = rn((A[2][1] - A[2][0] * A[1][0] / d0) / d1_exact, p)
    assert r[1][0] != naive10 and r[2][1] != naive21
    return A, r, dinv


def zero_pivot_case(p):
    """Order 2 [[1, x], [x, RN(x^2)]], x = 1 + 2^-(p-1): RN(x x) cancels the
    pivot; the exact difference RN(x^2) - x^2 = -2^-(2p-2)."""
    x = 1 + Fraction(1, 2 ** (p - 1))
    a11 = rn(x * x, p)
    assert a11 - x * x == -Fraction(1, 2 ** (2 * p - 2))
    A = [[Fraction(1), x], [x, a11]]
    return A, x, -Fraction(2) ** (2 * p - 2)


def rec_retry_case(p, n=256, row=128):
    """Order 256: the identity with a_{128,0} = a_{0,128} = x and a_{128,128}
    = RN(x^2), x = 1 + 2^-(p-1).  ldl_rec cuts at 129, halves the leading
    129 at 64 and forms a_{128,128} - RN(x x) = 0 in the trailing update:
    pivot 128 fails.  Exact dots keep it: d_128 = -2^-(2p-2).  Expected
    factor: L = I + x e_128 e_0^T, D^-1 = 1 except -2^(2p-2) at 128."""
    x = 1 + Fraction(1, 2 ** (p - 1))
    A = [[Fraction(int(i == j)) for j in range(n)] for i in range(n)]
    A[row][0] = A[0][row] = x
    A[row][row] = rn(x * x, p)
    dinv = [Fraction(1)] * n
    dinv[row] = -Fraction(2) ** (2 * p - 2)
    return A, x, dinv


def analytic_sampled_case(p, delta_power=None):
    """The analytic_schur block (sampled/tests.rs:407-434) as one leaf of
    an arrow with t = 1 border row: h = K = 2, basis columns (1, 1) and
    (1, -1), weights (2, -3), R^-1 = I (or diag(1, 1 + delta), delta =
    2^-(p/4), when delta_power is set), coupling B = (4, 6)^T and border
    C = (-1).  Returns the inputs and the expected leaf Schur block H (the
    closed form of entry(): w_p w_q 1/2 (K_pq^2 + K_pq^2), Gram K = V^T V,
    V = R^-1 Q), the factor L, D^-1, and the border's Schur complement and
    its D^-1, all with the reference's rounding sequence."""
    delta = Fraction(0) if delta_power is None else Fraction(1, 2 ** delta_power)
    Q = [[Fraction(1), Fraction(1)], [Fraction(1), Fraction(-1)]]  # Q[row][col]
    R = [[Fraction(1), Fraction(0)], [Fraction(0), 1 + delta]]
    w = [Fraction(2), Fraction(-3)]
    V = [[sum(R[i][k] * Q[k][j] for k in range(2)) for j in range(2)] for i in range(2)]
    # Gram entries are exact sums rounded once (exact here: few bits)
    K = [[rn(sum(V[k][a] * V[k][b] for k in range(2)), p) for b in range(2)] for a in range(2)]
    H = [[None, None], [None, None]]
    for a in range(2):
        for b in range(2):
            c = rn(rn(w[a] * w[b], p) * Fraction(1, 2), p)
            s = rn(rn(K[a][b] * K[a][b], p) + rn(K[a][b] * K[a][b], p), p)
            H[a][b] = rn(c * s, p)
    r, dinv = ldl_dots_rational(H, p)
    L10 = r[1][0]
    # Y = L^-1 B: y_i = RN(b_i - sum_k L_ik y_k), one exact dot rounded once
    # (forward_suffix_rows, arrow.rs:1000-1013)
    b = [Fraction(4), Fraction(6)]
    y = [b[0], rn(b[1] - L10 * b[0], p)]
    z = [rn(y[k] * dinv[k], p) for k in range(2)]
    contrib = rn(y[0] * z[0] + y[1] * z[1], p)
    c_border = Fraction(-1)
    s = rn(c_border - contrib, p)
    border_dinv = rn(1 / s, p)
    return {"Q": Q, "R": R, "w": w, "b": b, "c": c_border, "H": H, "L10": L10, "dinv": dinv,
            "y": y, "s": s, "border_dinv": border_dinv}


def sampled_problem_of(case, nl):
    """The analytic case as a problem dict of workloads.sampled_problem's
    form (one leaf, t = 1)."""
    blk = {
        "h": 2, "k": 2,
        "rinv": enc(col_major(case["R"]), nl),
        "q": enc(col_major(case["Q"]), nl),
        "w": enc(case["w"], nl),
        "b": enc(case["b"], nl),  # K x t, column-major
    }
    return {"blocks": [blk], "t": 1, "border": enc([case["c"]], nl)}


def retry_border_problem(nl):
    """A sampled arrow whose border is the order-256 rec_retry_case matrix
    (given in the upper triangle, as the border block C is) and whose one
    leaf (the analytic block) has zero coupling, so the border's Schur
    complement is C itself.  Returns (problem, x, expected border D^-1)."""
    p = 64 * nl
    A, x, dinv = rec_retry_case(p)
    case = analytic_sampled_case(p)
    prob = sampled_problem_of(case, nl)
    t = len(A)
    blk = prob["blocks"][0]
    blk["b"] = enc([0] * (2 * t), nl)
    upper = [A[i][j] if i <= j else Fraction(0) for j in range(t) for i in range(t)]
    prob["t"] = t
    prob["border"] = enc(upper, nl)
    return prob, x, dinv
