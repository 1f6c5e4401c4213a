"""Synthetic workloads of the Schur hot path (numpy only, seeded).

sampled_problem: the bootstrap-style sampled SDP of one IPM iteration:
blocks of side h ~ U{hmin..hmax}, K = 2h - 1 samples x_k ~ U[-1, 1], basis
Q[:, k] = (T_0(x_k), ..., T_{h-1}(x_k)) (Chebyshev), weights log-uniform in
[1e-6, 1], an NT inverse factor Rinv = I + G/(4 sqrt h) per block, a dense
coupling of every sample to the t free variables B[k, n] = T_n(x_k), and the
border block -eps I.  Values are lifted to full 64N-bit mantissas (random
low bits below the binary64 value) so every product carries full-width
operands.
"""
import numpy as np

from . import mp


def lift_full(x, n, rng):
    """Binary64 values -> 64n-bit records with random low mantissa bits."""
    rec = mp.from_float64(np.ravel(x), n)
    if n > 1:
        nz = np.abs(rec["kind"]) == mp.REGULAR_KIND
        low = rng.integers(0, 2**64, size=(rec.size, n - 1), dtype=np.uint64)
        rec["limbs"][nz, : n - 1] = low[nz]
    top_mask = np.uint64((1 << 11) - 1)  # 53-bit value occupies the top limb's high bits
    nz = np.abs(rec["kind"]) == mp.REGULAR_KIND
    rec["limbs"][nz, n - 1] |= rng.integers(0, 2**11, size=int(nz.sum()), dtype=np.uint64) & top_mask
    return rec


def chebyshev(x, count):
    x = np.clip(np.asarray(x, dtype=np.float64), -1.0, 1.0)
    return np.cos(np.outer(np.arange(count), np.arccos(x)))  # count x len(x)


def sampled_problem(n, nblocks=40, hmin=10, hmax=60, t=2000, eps=1e-8, seed=7, lift=True,
                    weight_range=(1e-6, 1.0), samples=lambda h: 2 * h - 1):
    """Returns dict with per-block records and float64 views:
    blocks: list of {h, k, rinv, q, w, b (records), and *_f64 arrays};
    border (t x t records, upper = -eps I), border_f64."""
    rng = np.random.default_rng(seed)
    blocks = []
    for _ in range(nblocks):
        h = int(rng.integers(hmin, hmax + 1))
        k = int(samples(h))
        x = rng.uniform(-1.0, 1.0, k)
        q = chebyshev(x, h)                          # h x k
        w = np.exp(rng.uniform(np.log(weight_range[0]), np.log(weight_range[1]), k))
        g = rng.standard_normal((h, h))
        rinv = np.eye(h) + g / (4.0 * np.sqrt(h))
        b = chebyshev(x, t).T                         # k x t
        conv = (lambda a: lift_full(a, n, rng)) if lift else (lambda a: mp.from_float64(np.ravel(a), n))
        blocks.append({
            "h": h, "k": k,
            "rinv": conv(rinv.T.ravel()), "q": conv(q.T.ravel()), "w": conv(w),
            "b": conv(b.T.ravel()),
            "rinv_f64": rinv, "q_f64": q, "w_f64": w, "b_f64": b,
        })
    border = mp.zeros(t * t, n)
    diag = mp.from_float64(np.full(t, -eps), n)
    border[np.arange(t) * (t + 1)] = diag
    return {"blocks": blocks, "t": t, "border": border, "eps": eps,
            "border_f64": -eps * np.eye(t)}
