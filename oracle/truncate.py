"""truncate!: decide how many weights of a non-increasing spectrum to keep.

Two restatements:
  ndtensors_truncate  the CPU NDTensors truncate! (ITensors 0.2) that the
                      reference's CPU path runs; this is what DMRG follows.
  refgpu_truncate     the reference's GPU kernel, src/tensor/cutruncate.jl:1-93,
                      including its divergences from the CPU version; it is
                      the one its golden vectors (test/test_cutruncate.jl:9-17)
                      were written for.
Both return (n_kept, truncerr, docut).
"""
import math

import numpy as np


def _docut(P, n):
    # cutruncate.jl:81-87 (identical rule in NDTensors)
    docut = 0.0
    if n < len(P):
        docut = (P[n - 1] + P[n]) / 2.0
        if abs(P[n - 1] - P[n]) < 1e-3 * P[n - 1]:
            docut += 1e-3 * P[n - 1]
    return docut


def ndtensors_truncate(P, maxdim=None, mindim=1, cutoff=0.0, absolute_cutoff=False,
                       relative_cutoff=True):
    P = [float(x) for x in P]
    origm = len(P)
    maxdim = origm if maxdim is None or maxdim <= 0 else min(int(maxdim), origm)
    mindim = max(int(mindim), 1)
    cutoff = max(float(cutoff), 0.0)
    if origm == 0:
        return 0, 0.0, 0.0
    if P[0] <= 0.0:
        return 1, 0.0, 0.0
    if origm == 1:
        return 1, 0.0, P[0] / 2.0
    rP = list(P)
    # zero out negative weights at the tail
    for i in range(origm - 1, -1, -1):
        if rP[i] >= 0.0:
            break
        rP[i] = 0.0
    n = origm
    truncerr = 0.0
    while n > maxdim:
        truncerr += rP[n - 1]
        n -= 1
    if absolute_cutoff:
        while n > mindim and rP[n - 1] <= cutoff:
            truncerr += rP[n - 1]
            n -= 1
    else:
        scale = 1.0
        if relative_cutoff:
            scale = sum(rP)
            if scale == 0.0:
                scale = 1.0
        while n > mindim and truncerr + rP[n - 1] <= cutoff * scale:
            truncerr += rP[n - 1]
            n -= 1
        truncerr /= scale
    n = max(n, 1)
    return n, truncerr, _docut(rP, n)


def refgpu_truncate(P, maxdim=None, mindim=1, cutoff=0.0, absolute_cutoff=False,
                    relative_cutoff=True):
    """Line-by-line semantics of src/tensor/cutruncate.jl:1-93."""
    P = np.asarray(P, dtype=np.float64)
    origm = len(P)
    maxdim = origm if maxdim is None or maxdim <= 0 else min(int(maxdim), origm)  # :3
    mindim = min(int(mindim), maxdim)                                            # :4
    maxP = P.max()                                                               # :10
    if maxP == 0.0:                                                              # :11-14
        return 1, 0.0, 0.0
    if origm == 1:                                                               # :15-18
        return 1, 0.0, maxP / 2.0
    rP = np.where(np.signbit(P), 0.0, P)                                         # :23
    n = origm
    truncerr = 0.0
    if n > maxdim:                                                               # :26-29
        truncerr = float(rP[: n - maxdim].sum())   # sums the LARGEST weights
        n = maxdim

    def first_below(sub):
        # :35-38 / :53-56. err = sub/|sub| is -1 below, +1 above; the
        # sign-bit flag makes below-cutoff entries -2, others 0; iamax picks
        # the first maximal |.| (1-based), or 1 if all are zero. An exact
        # zero of `sub` gives NaN and is treated as "not below".
        below = sub < 0.0
        return int(np.argmax(below)) if below.any() else 0

    if absolute_cutoff:                                                          # :32-41
        cut_ind = first_below(rP - cutoff)
        n = min(maxdim, origm - cut_ind)
        n = max(n, mindim)
        truncerr += float(rP[cut_ind:].sum())
    else:
        scale = 1.0
        if relative_cutoff:                                                      # :45-48
            scale = float(P.sum())
            scale = scale if scale > 0.0 else 1.0
        cut_ind = first_below(rP + truncerr - cutoff * scale)                    # :53-56
        if cut_ind > 0:                                                          # :57-65
            truncerr += float(rP[cut_ind:].sum())
        else:                                                                    # :66-74
            truncerr += float(rP[:maxdim].sum())
        n = min(maxdim, origm - cut_ind)
        n = max(n, mindim)
        truncerr = 0.0 if scale == 0.0 else truncerr / scale
    if n < 1:                                                                    # :78-80
        n = 1
    return n, truncerr, _docut(P, n)


def truncate(P, semantics="ndtensors", **kw):
    if semantics == "ndtensors":
        return ndtensors_truncate(P, **kw)
    if semantics == "refgpu":
        return refgpu_truncate(P, **kw)
    raise ValueError(semantics)


def truncation_bound_ok(P, n, truncerr, cutoff, relative=True):
    """Property: discarded weight / scale <= cutoff unless capped by maxdim."""
    scale = sum(P) if relative else 1.0
    disc = sum(P[n:]) / (scale if scale else 1.0)
    return math.isclose(disc, truncerr, rel_tol=1e-12, abs_tol=1e-300)
