"""MpFloat<N> interchange records as numpy structured arrays.

Layout (crates/arithmetic/src/lib.rs:333-340, 509-533): ``limbs`` (N x u64,
least significant first), ``kind`` (i32 MPFR kind code: 0 NaN, 1 Inf,
2 zero, 3 regular; negated for negative values), ``exponent`` (i32); a
regular value is M * 2**(exponent - 64N) with 2**(64N-1) <= M < 2**(64N).
Zero is all-zero limbs, kind 2, exponent 0 (lib.rs:366-373).

Generators here build synthetic workloads with numpy only (no big-integer
loops), so full-size inputs take seconds.
"""
import numpy as np

NAN_KIND, INF_KIND, ZERO_KIND, REGULAR_KIND = 0, 1, 2, 3


def dtype(n):
    return np.dtype([("limbs", "<u8", (n,)), ("kind", "<i4"), ("exponent", "<i4")])


def limbs_of(arr):
    return arr.dtype["limbs"].shape[0]


def zeros(count, n):
    out = np.zeros(count, dtype(n))
    out["kind"] = ZERO_KIND
    return out


def _bitlen64(x):
    """Bit length of each uint64 (0 for 0), vectorised."""
    x = x.astype(np.uint64, copy=True)
    n = np.zeros(x.shape, np.int64)
    for s in (32, 16, 8, 4, 2, 1):
        big = x >= (np.uint64(1) << np.uint64(s))
        n += big * s
        x = np.where(big, x >> np.uint64(s), x)
    return n + (x > 0)


def _shift_left(limbs, sh):
    """Shift multi-limb integers (rows of ``limbs``, LSB first) left by
    per-row ``sh`` bits (0 <= sh < 64 * L), dropping overflow.  Rows are
    grouped by their whole-limb shift, so each group moves by slicing."""
    L = limbs.shape[1]
    whole = sh // 64
    part = (sh % 64).astype(np.uint64)
    out = np.zeros_like(limbs)
    for w in np.unique(whole):
        w = int(w)
        sel = whole == w
        every = bool(sel.all())
        src = limbs if every else limbs[sel]
        p = (part if every else part[sel])[:, None]
        moved = np.zeros_like(src)
        moved[:, w:] = src[:, :L - w]
        prev = np.zeros_like(src)
        prev[:, w + 1:] = src[:, :L - w - 1]
        lo = np.where(p > 0, prev >> (np.uint64(64) - np.where(p > 0, p, np.uint64(1))), np.uint64(0))
        res = (moved << p) | lo
        if every:
            out[:] = res
        else:
            out[sel] = res
    return out


def from_integers(mag, negative, scale, n):
    """Values (-1)**negative * mag * 2**scale, exact; ``mag`` (V, L) uint64
    little-endian with bit length <= 64n.  Zero magnitudes give zero."""
    mag = np.ascontiguousarray(mag, dtype=np.uint64)
    V, L = mag.shape
    if L < n:
        mag = np.concatenate([mag, np.zeros((V, n - L), np.uint64)], axis=1)
        L = n
    nonzero = mag != 0
    top = L - 1 - np.argmax(nonzero[:, ::-1], axis=1)  # highest nonzero limb
    bl = np.where(nonzero.any(axis=1), 64 * top + _bitlen64(mag[np.arange(V), top]), 0)
    if np.any(bl > 64 * n):
        raise ValueError("magnitude wider than the precision")
    out = zeros(V, n)
    nz = bl > 0
    shifted = _shift_left(mag, np.where(nz, 64 * L - bl, 0))
    out["limbs"][nz] = shifted[nz][:, L - n:]
    out["kind"][nz] = np.where(np.asarray(negative)[nz], -REGULAR_KIND, REGULAR_KIND)
    out["exponent"][nz] = (np.asarray(scale, np.int64) + bl)[nz] if np.ndim(scale) else (scale + bl)[nz]
    return out


def random_fixed(count, n, frac_bits, rng):
    """Fixed-point values uniform in [-1, 1) with ``frac_bits`` fraction bits
    (frac_bits <= 64n), exactly representable at 64n bits."""
    words = (frac_bits + 63) // 64
    mag = rng.integers(0, 2**64, size=(count, words), dtype=np.uint64, endpoint=False)
    extra = 64 * words - frac_bits
    if extra:
        mag[:, -1] &= np.uint64((1 << (64 - extra)) - 1)
    neg = rng.integers(0, 2, size=count).astype(bool)
    return from_integers(mag, neg, -frac_bits, n)


def random_full(count, n, spread, rng, zero_fraction=0.05):
    """Full-mantissa values: random 64n-bit mantissas, random signs, binary
    exponents uniform in [-spread, spread], a fraction of exact zeros."""
    out = zeros(count, n)
    limbs = rng.integers(0, 2**64, size=(count, n), dtype=np.uint64)
    limbs[:, -1] |= np.uint64(1 << 63)
    out["limbs"] = limbs
    neg = rng.integers(0, 2, size=count).astype(bool)
    out["kind"] = np.where(neg, -REGULAR_KIND, REGULAR_KIND)
    out["exponent"] = rng.integers(-spread, spread + 1, size=count)
    z = rng.random(count) < zero_fraction
    out["limbs"][z] = 0
    out["kind"][z] = ZERO_KIND
    out["exponent"][z] = 0
    return out


def from_float64(x, n):
    """Exact conversion of binary64 values (finite) to 64n-bit records."""
    x = np.asarray(x, dtype=np.float64).ravel()
    if not np.all(np.isfinite(x)):
        raise ValueError("non-finite input")
    mant, ex = np.frexp(np.abs(x))  # |x| = mant * 2**ex, mant in [0.5, 1)
    # frexp normalises subnormals too, so every nonzero magnitude has a
    # 53-bit mantissa in [2^52, 2^53): the record's top limb is that
    # mantissa shifted to bit 63, the lower limbs are zero and the exponent
    # is ex (value = top * 2^(64(n-1)) * 2^(ex - 64n))
    m53 = (mant * 2.0**53).astype(np.uint64)  # exact
    out = zeros(x.size, n)
    nz = x != 0
    out["limbs"][:, n - 1] = np.where(nz, m53 << np.uint64(11), np.uint64(0))
    out["kind"] = np.where(nz, np.where(x < 0, -REGULAR_KIND, REGULAR_KIND), ZERO_KIND)
    out["exponent"] = np.where(nz, ex, 0)
    return out


def to_float64(rec):
    """Nearest binary64 of each value (lossy; for reporting only)."""
    n = limbs_of(rec)
    top = rec["limbs"][..., n - 1].astype(np.float64)
    nxt = rec["limbs"][..., n - 2].astype(np.float64) if n > 1 else 0.0
    m = top + nxt / 2.0**64
    e = rec["exponent"].astype(np.int64) - 64
    with np.errstate(over="ignore", under="ignore"):
        v = np.ldexp(m, e)
    kind = rec["kind"]
    v = np.where(np.abs(kind) == ZERO_KIND, 0.0, v)
    v = np.where(kind < 0, -v, v)
    v = np.where(np.abs(kind) == INF_KIND, np.where(kind < 0, -np.inf, np.inf), v)
    v = np.where(kind == NAN_KIND, np.nan, v)
    return v


def same_bits(x, y):
    """Bitwise equality with all zeros identified (the reference's test
    comparison ``x == y || both zero``, rns_blas_tests.rs:101-105)."""
    zx = np.abs(x["kind"]) == ZERO_KIND
    zy = np.abs(y["kind"]) == ZERO_KIND
    eq = (np.all(x["limbs"] == y["limbs"], axis=-1) & (x["kind"] == y["kind"])
          & (x["exponent"] == y["exponent"]))
    return eq | (zx & zy)


def _exact(rec):
    """(integer, binary scale) with value = integer * 2^scale, exactly."""
    n = limbs_of(np.asarray([rec]) if np.ndim(rec) == 0 else rec)
    kind = int(rec["kind"])
    if abs(kind) != REGULAR_KIND:
        return 0, 0
    m = 0
    for i, w in enumerate(rec["limbs"].tolist()):
        m |= int(w) << (64 * i)
    if kind < 0:
        m = -m
    return m, int(rec["exponent"]) - 64 * n


def sub_exact_log2(b, c):
    """log2 of max_i |b_i - c_i| computed exactly (regular or zero records of
    any equal width); -inf when every difference is zero."""
    best = -np.inf
    for x, y in zip(b, c):
        mx, ex = _exact(x)
        my, ey = _exact(y)
        e = min(ex, ey)
        d = (mx << (ex - e)) - (my << (ey - e))
        if d:
            best = max(best, (abs(d).bit_length() - 1) + e)
    return float(best)
