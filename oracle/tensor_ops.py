"""Contraction, permutation and permuted addition of dense column-major tensors.

Tensors are numpy arrays whose logical shape is the index dimensions in order
(the flat data is column-major, as NDTensors `Dense` storage).
"""
import itertools

import numpy as np


def from_flat(flat, dims):
    """View a flat column-major buffer as a tensor (cudense.jl:19)."""
    return np.reshape(np.asarray(flat), tuple(dims), order="F")


def to_flat(t):
    return np.ravel(np.asarray(t), order="F")


def classify_labels(labelsA, labelsB, labelsC):
    """Split labels into contracted / A-only / B-only, in the order the
    reference numbers cuTENSOR modes: contracted (in A's order), then A-only,
    then B-only (src/tensor/cudense.jl:255-283)."""
    la, lb, lc = list(labelsA), list(labelsB), list(labelsC)
    if len(set(la)) != len(la) or len(set(lb)) != len(lb) or len(set(lc)) != len(lc):
        raise ValueError("repeated label within one tensor")
    contracted = [x for x in la if x in lb and x not in lc]
    a_only = [x for x in la if x not in lb]
    b_only = [x for x in lb if x not in la]
    for x in lc:
        if (x in la) == (x in lb):
            raise ValueError("output label %r must appear in exactly one input" % (x,))
    if sorted(a_only + b_only) != sorted(lc):
        raise ValueError("uncontracted labels do not match the output labels")
    return contracted, a_only, b_only


def mode_numbers(labelsA, labelsB, labelsC):
    """cuTENSOR integer modes for A, B, C exactly as cudense.jl:258-283 builds
    them (1-based positions in contracted ++ A_only ++ B_only)."""
    c, a, b = classify_labels(labelsA, labelsB, labelsC)
    order = c + a + b
    pos = {x: i + 1 for i, x in enumerate(order)}
    return ([pos[x] for x in labelsA], [pos[x] for x in labelsB],
            [pos[x] for x in labelsC])


def itensor_output_labels(labelsA, labelsB):
    """Output index order of ITensor `A*B`: A's unique indices then B's
    (NDTensors contract_labels, reached via contract!!, cudense.jl:83-110)."""
    return [x for x in labelsA if x not in labelsB] + [x for x in labelsB if x not in labelsA]


def _letters(labels, table):
    return "".join(table[x] for x in labels)


def contract(A, labelsA, B, labelsB, labelsC, alpha=1.0, beta=0.0, C=None):
    """C = alpha * A*B + beta * C by numpy einsum (the library cross-check)."""
    classify_labels(labelsA, labelsB, labelsC)
    allk = list(dict.fromkeys(list(labelsA) + list(labelsB)))
    if len(allk) > 52:
        raise ValueError("too many distinct labels")
    alphabet = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"
    table = {x: alphabet[i] for i, x in enumerate(allk)}
    spec = "%s,%s->%s" % (_letters(labelsA, table), _letters(labelsB, table),
                          _letters(labelsC, table))
    R = alpha * np.einsum(spec, np.asarray(A), np.asarray(B), optimize=True)
    if beta != 0.0:
        R = R + beta * np.asarray(C)
    return R


def contract_bruteforce(A, labelsA, B, labelsB, labelsC, dims_of):
    """Explicit loops over every index value — pins `contract` on tiny inputs.
    dims_of: label -> dimension."""
    c, _, _ = classify_labels(labelsA, labelsB, labelsC)
    shapeC = tuple(dims_of[x] for x in labelsC)
    R = np.zeros(shapeC, dtype=np.result_type(A, B))
    for out in itertools.product(*[range(n) for n in shapeC]):
        val = {x: out[i] for i, x in enumerate(labelsC)}
        s = 0.0
        for kk in itertools.product(*[range(dims_of[x]) for x in c]):
            for i, x in enumerate(c):
                val[x] = kk[i]
            ia = tuple(val[x] for x in labelsA)
            ib = tuple(val[x] for x in labelsB)
            s += A[ia] * B[ib]
        R[out] = s
    return R


def permute(A, perm, alpha=1.0):
    """B = alpha*permutedims(A, perm): B's mode j is A's mode perm[j]
    (cudense.jl:38-44,447-478; 0-based perm here)."""
    B = np.transpose(np.asarray(A), tuple(perm))
    return B if alpha == 1.0 else alpha * B


def perm_from_indices(indsA, indsB):
    """perm such that B = permutedims(A, perm) has index order indsB
    (getperm, used at cudense.jl:92,96)."""
    return [list(indsA).index(x) for x in indsB]


def add_permuted(A, perm, alpha, B, beta):
    """alpha*permutedims(A,perm) + beta*B (cudense.jl:333-360: opAC=ADD)."""
    return alpha * np.transpose(np.asarray(A), tuple(perm)) + beta * np.asarray(B)
