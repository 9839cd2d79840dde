"""Algorithmic work of one log-likelihood evaluation (FP64 flops).

Standard dense counts: GEMM 2mnk, SYRK n^2 k, TRSM m^2 n, POTRF n^3/3, with r
the knots per coarse region, L = M-1 ancestor levels, n_F the observations of
finest region F and n_s the observations below a finest-but-one region.  The
counts follow the reference's per-region operation sequence
(src/class_approximation.cpp:108-159;
src/class_approximation-loop-regions-at-finest-level-in-prior.cpp:116-256;
src/class_approximation-posterior-inference.cpp:51-219):

  coarse prior, region at level l:  sum_j (2 r^3 j + r^3) + r^3 l + r^3/3
                                     = r^3 (l^2 + l + 1/3)
  finest conditioning (per obs):     sum_j (2 r^2 j + r^2) = r^2 L^2
  finest local (region F):           n_F^2 L r (SYRK) + n_F^3/3 + n_F^2
                                     + L r n_F^2 (H = L_F^{-1} B^T)
  posterior, region at level l:      r^3/3 (L L^T) + r^3/3 (POTRF) + l r^3 (P_j)
                                     + l^2 r^3 (A~ = A - P^T P, j <= k < l)
  level L-1 also forms A from the stacked finest rows: L^2 r^2 n_s
"""
import numpy as np


def level_counts(M, J):
    return [J ** l for l in range(M - 1)]


def prior_flops(l, r, nreg):
    return nreg * r ** 3 * (l * l + l + 1.0 / 3.0)


def posterior_flops(l, r, nreg, ns_total=0, L=None):
    f = nreg * r ** 3 * (2.0 / 3.0 + l + l * l)
    if L is not None and l == L - 1:
        f += L * L * r * r * ns_total
    return f


def finest_cross_flops(r, L, n):
    return r * r * L * L * n


def finest_local_flops(r, L, counts):
    c = np.asarray(counts, dtype=np.float64)
    return float((2 * L * r * c * c + c ** 3 / 3 + c * c).sum())


def groups(M, J, r, counts, working_counts=None):
    """dict: launch-group name -> algorithmic flops per evaluation."""
    L = M - 1
    n = float(np.sum(counts))
    out = {}
    nl = level_counts(M, J)
    for l in range(L):
        out[f"prior_l{l}"] = prior_flops(l, r, nl[l])
    out["finest_cross"] = finest_cross_flops(r, L, n)
    out["finest_local"] = finest_local_flops(r, L, counts)
    for l in range(L):
        nw = nl[l] if working_counts is None else working_counts[l]
        out[f"post_l{l}"] = posterior_flops(l, r, nw, ns_total=n, L=L)
    return out


def total(M, J, r, counts):
    return sum(groups(M, J, r, counts).values())
