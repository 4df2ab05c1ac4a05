"""MPS/MPO algebra on device ITensors: the reference's cuMPS/cuMPO surface
(src/mps/cumps.jl, src/mps/cumpo.jl, exercised by test/test_cumpo.jl) and
gate application (examples/gate_evolution.jl).

An MPS or MPO is a list of site ITensors (link indices shared between
neighbours, site indices the rest), as in ITensors; every numeric step is
an ITensor operation of `itensor.py`, i.e. a libitg call: contraction,
permutation, QR/SVD, the block embedding of `add`.
"""
import numpy as np

from . import _lib
from . import ops
from .itensor import ITensor, Index, combiner, contract, qr, svd


class MPSList:
    """ITensors MPS/MPO: site tensors plus the orthogonality limits."""

    kind = "MPS"

    def __init__(self, tensors, llim=0, rlim=None):
        self.data = list(tensors)
        self.llim = llim
        self.rlim = len(self.data) + 1 if rlim is None else rlim

    def __len__(self):
        return len(self.data)

    def __getitem__(self, j):
        return self.data[j]

    def __setitem__(self, j, T):
        self.data[j] = T

    def copy(self):
        return type(self)(list(self.data), self.llim, self.rlim)

    def __repr__(self):
        return "\n".join([self.kind] + ["[%d] %r" % (j + 1, T.inds) for j, T in enumerate(self.data)]
                         + [""])

    def linkind(self, j):
        """Index shared by sites j and j+1 (None if there is none)."""
        if j < 0 or j + 1 >= len(self):
            return None
        common = [i for i in self.data[j].inds if i in self.data[j + 1].inds]
        return common[0] if common else None

    def siteinds(self, j):
        links = {self.linkind(j - 1), self.linkind(j)}
        return [i for i in self.data[j].inds if i not in links]

    def maxlinkdim(self):
        dims = [self.linkind(j).dim for j in range(len(self) - 1) if self.linkind(j) is not None]
        return max(dims) if dims else 1

    def dag(self):
        return type(self)([T.dag() for T in self.data], self.llim, self.rlim)

    def prime(self, n=1):
        return type(self)([T.prime(n) for T in self.data], self.llim, self.rlim)

    def cpu(self):
        return type(self)([T.cpu() for T in self.data], self.llim, self.rlim)

    def roc(self):
        return type(self)([T.roc() for T in self.data], self.llim, self.rlim)


class MPS(MPSList):
    kind = "MPS"


class MPO(MPSList):
    kind = "MPO"


def _normalize(T):
    return T * (1.0 / T.norm())


# ------------------------------------------------------------ constructors
def random_mps(sites, linkdims=1, seed=0, dtype="float64"):
    """randomCuMPS (src/mps/cumps.jl:44-53): Gaussian site tensors of the
    Philox stream with links of dimension `linkdims`, each normalised, then
    orthogonalised to the first site and normalised."""
    from .itensor import random_roc_itensor
    N = len(sites)
    links = [Index(linkdims, "Link,l=%d" % (j + 1)) for j in range(N - 1)]
    ts, off = [], 0
    for j in range(N):
        inds = ([links[j - 1]] if j > 0 else []) + [sites[j]] + ([links[j]] if j < N - 1 else [])
        T = random_roc_itensor(*inds, seed=seed, offset=off, dtype=dtype)
        off += int(np.prod([i.dim for i in inds]))
        ts.append(_normalize(T))
    psi = MPS(ts)
    orthogonalize(psi, 0, normalize=True)
    return psi


def random_mpo(sites, seed=0, dtype="float64"):
    """randomCuMPO(sites) (src/mps/cumpo.jl:13-21): MPO(sites)'s layout
    (links of dimension 1) with every site tensor a normalised Gaussian."""
    from .itensor import random_roc_itensor
    N = len(sites)
    links = [Index(1, "Link,l=%d" % (j + 1)) for j in range(N - 1)]
    ts, off = [], 0
    for j, s in enumerate(sites):
        inds = ([links[j - 1]] if j > 0 else []) + [s, s.prime()] + \
            ([links[j]] if j < N - 1 else [])
        T = random_roc_itensor(*inds, seed=seed, offset=off, dtype=dtype)
        off += s.dim * s.dim
        ts.append(_normalize(T))
    return MPO(ts)


def product_mps(sites, states, dtype="float64"):
    """productCuMPS(sites, states): a product state with links of dimension
    1; `states` are 0-based basis states or amplitude vectors per site."""
    from .itensor import itensor
    N = len(sites)
    links = [Index(1, "Link,l=%d" % (j + 1)) for j in range(N - 1)]
    ts = []
    for j, s in enumerate(sites):
        v = np.zeros(s.dim, dtype=dtype) if np.isscalar(states[j]) else np.asarray(states[j],
                                                                                   dtype=dtype)
        if np.isscalar(states[j]):
            v[int(states[j])] = 1.0
        inds = ([links[j - 1]] if j > 0 else []) + [s] + ([links[j]] if j < N - 1 else [])
        shape = [i.dim for i in inds]
        ts.append(itensor(np.reshape(v, shape, order="F"), *inds, dtype=dtype).roc())
    return MPS(ts)


# ------------------------------------------------------------------ inner
def _check_sites(A, B):
    if len(A) != len(B):
        raise _lib.DimensionMismatch("MPS/MPO lengths differ: %d and %d" % (len(A), len(B)))


def _sim_links(M, avoid):
    """M with every link index replaced by a fresh one if it collides with
    an index of `avoid` (ITensors' sim(linkinds) in inner)."""
    bad = set()
    for T in avoid.data:
        bad.update(T.inds)
    out = list(M.data)
    for j in range(len(M) - 1):
        l = M.linkind(j)
        if l is not None and l in bad:
            n = l.sim()
            for k in (j, j + 1):
                out[k] = ITensor([n if i == l else i for i in out[k].inds], out[k].data)
    return type(M)(out, M.llim, M.rlim)


def inner(*args):
    """inner(phi, psi) = <phi|psi>, inner(phi, K, psi) = <phi|K|psi>
    (complex conjugate on phi)."""
    if len(args) == 2:
        phi, psi = args
        _check_sites(phi, psi)
        pd = _sim_links(phi.dag(), psi)
        E = contract(pd[0], psi[0])
        for j in range(1, len(psi)):
            E = contract(contract(E, pd[j]), psi[j])
        return E.scalar()
    phi, K, psi = args
    _check_sites(phi, psi)
    _check_sites(phi, K)
    pd = _sim_links(_sim_links(phi.dag(), K), psi)
    # phi's site indices become K's site indices that psi does not hold
    # (ITensors' make_inds_match: primed sites, or another site set)
    for j in range(len(psi)):
        sx = [i for i in K.siteinds(j) if i not in psi[j].inds]
        sy = pd.siteinds(j)
        if len(sx) != 1 or len(sy) != 1 or sx[0].dim != sy[0].dim or \
                len([i for i in psi.siteinds(j) if i in K[j].inds]) != 1:
            raise _lib.DimensionMismatch("inner: site indices of site %d do not match" % j)
        pd[j] = ITensor([sx[0] if i == sy[0] else i for i in pd[j].inds], pd[j].data)
    E = contract(contract(pd[0], K[0]), psi[0])
    for j in range(1, len(psi)):
        E = contract(contract(contract(E, pd[j]), K[j]), psi[j])
    return E.scalar()


# ---------------------------------------------------------- orthogonalize
def _factor_left(T, left_inds, maxdim, cutoff, tags):
    """T = L * R with L over left_inds orthonormal: QR, or a truncated SVD
    when maxdim/cutoff ask for truncation."""
    if maxdim is None and not cutoff:
        Q, R = qr(T, left_inds, tags=tags)
        return Q, R
    U, S, V, _ = svd(T, left_inds, maxdim=maxdim, cutoff=cutoff or 0.0, lefttags=tags,
                     righttags=tags)
    return U, contract(S, V)


def orthogonalize(M, j, maxdim=None, cutoff=0.0, normalize=False):
    """orthogonalize!(M, j) (0-based j) in place: QR (or, with maxdim or
    cutoff, a truncated SVD) from both ends towards site j; with
    `normalize` the centre is scaled to norm 1."""
    N = len(M)
    for b in range(0, j):
        l = M.linkind(b)
        if l is None:
            continue
        left = [i for i in M[b].inds if i != l]
        A, R = _factor_left(M[b], left, maxdim, cutoff, "Link,l=%d" % (b + 1))
        M[b] = A
        M[b + 1] = contract(R, M[b + 1])
    for b in range(N - 1, j, -1):
        l = M.linkind(b - 1)
        if l is None:
            continue
        right = [i for i in M[b].inds if i != l]
        A, R = _factor_left(M[b], right, maxdim, cutoff, "Link,l=%d" % b)
        M[b] = A
        M[b - 1] = contract(M[b - 1], R)
    if normalize:
        M[j] = _normalize(M[j])
    M.llim, M.rlim = j - 1, j + 1
    return M


# --------------------------------------------------------------- contract
def contract_mpo(K, psi, maxdim=None, cutoff=0.0, method="naive"):
    """contract(K, psi): the MPO applied to an MPS (or an MPO to an MPO),
    site by site over the shared site indices, links fused by combiners,
    then truncated by an orthogonalisation and an SVD sweep (the "naive"
    algorithm of ITensors)."""
    if method not in ("naive", "densitymatrix"):
        raise _lib.ArgumentError("contract: unknown method %r" % (method,))
    _check_sites(K, psi)
    N = len(psi)
    ts = []
    is_mps = isinstance(psi, MPS)
    for j in range(N):
        P = psi[j]
        if is_mps:
            # the MPS site meets K's lowest-primed index of the same site
            (s,) = psi.siteinds(j)
            cands = [i for i in K.siteinds(j) if i.id == s.id]
            if not cands:
                raise _lib.DimensionMismatch("contract: site %d of the MPS is not K's" % j)
            p = min(i.plev for i in cands)
            if p != s.plev:
                P = ITensor([i.prime(p - s.plev) if i == s else i for i in P.inds], P.data)
        if not any(i in P.inds for i in K[j].inds):
            raise _lib.DimensionMismatch("contract: site %d shares no index" % j)
        T = contract(K[j], P)
        if is_mps:
            T = T.noprime()
        elif any(i.plev == 2 for i in T.inds):
            # prime(K) * L: the pair (s'', s) back to (s', s)
            T = ITensor([Index(i.dim, i.tags, 1, i.id) if i.plev == 2 else i for i in T.inds],
                        T.data)
        ts.append(T)
    # fuse the two link indices of every bond
    for j in range(N - 1):
        lk, lp = K.linkind(j), psi.linkind(j)
        pair = [i for i in (lk, lp) if i is not None]
        if is_mps:
            pair = [i.noprime() for i in pair]
        else:
            pair = [Index(i.dim, i.tags, 1, i.id) if i.plev == 2 else i for i in pair]
        if len(pair) < 2:
            continue
        Cm = combiner(*pair, tags="Link,l=%d" % (j + 1))
        ts[j] = contract(Cm, ts[j])
        ts[j + 1] = contract(Cm, ts[j + 1])
    out = type(psi)(ts)
    orthogonalize(out, N - 1)
    if maxdim is not None or cutoff:
        for b in range(N - 1, 0, -1):
            l = out.linkind(b - 1)
            right = [i for i in out[b].inds if i != l]
            A, R = _factor_left(out[b], right, maxdim, cutoff, "Link,l=%d" % b)
            out[b] = A
            out[b - 1] = contract(out[b - 1], R)
        out.llim, out.rlim = -1, 1
    return out


# -------------------------------------------------------------------- add
def add(A, B):
    """A + B for MPS or MPO: the direct sum of every link space (site j's
    tensor is block-diagonal over its two links, itg_embed)."""
    _check_sites(A, B)
    N = len(A)
    if A[0].is_complex() != B[0].is_complex():
        A = type(A)([T.complex() for T in A.data])
        B = type(B)([T.complex() for T in B.data])
    if N == 1:
        return type(A)([A[0] + B[0]])
    newl = [Index(A.linkind(j).dim + B.linkind(j).dim, "Link,l=%d" % (j + 1)) for j in range(N - 1)]
    ts = []
    for j in range(N):
        la = [A.linkind(j - 1), A.linkind(j)]
        lb = [B.linkind(j - 1), B.linkind(j)]
        sites = A.siteinds(j)
        if set(sites) != set(B.siteinds(j)):
            raise _lib.DimensionMismatch("add: site indices of site %d differ" % j)
        keep = [k for k in (0, 1) if la[k] is not None]
        order_a = [la[k] for k in keep] + sites
        order_b = [lb[k] for k in keep] + sites
        new_inds = [newl[j - 1] if k == 0 else newl[j] for k in keep] + sites
        dims = [i.dim for i in new_inds]
        dt = ops.dtype_name(A[j].data)
        out = ops.fill(ops.empty(int(np.prod(dims)), dt), 0.0)
        Pa = A[j].permute(*order_a)
        Pb = B[j].permute(*order_b)
        ops.embed(Pa.data, Pa.dims, out, dims, [0] * len(dims))
        offs = [la[k].dim for k in keep] + [0] * len(sites)
        ops.embed(Pb.data, Pb.dims, out, dims, offs)
        ts.append(ITensor(new_inds, out))
    return type(A)(ts)


# --------------------------------------------------------------- gates
QUBIT_OPS = {
    "Id": np.eye(2), "I": np.eye(2),
    "X": np.array([[0.0, 1.0], [1.0, 0.0]]),
    "Y": np.array([[0.0, -1j], [1j, 0.0]]),
    "Z": np.array([[1.0, 0.0], [0.0, -1.0]]),
    "H": np.array([[1.0, 1.0], [1.0, -1.0]]) / np.sqrt(2.0),
}
TWO_QUBIT_OPS = {
    "CNOT": np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=float),
    "SWAP": np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=float),
}


def op(name, *s):
    """A gate as an ITensor (s1', s2', ..., s1, s2, ...): the matrix's rows
    are the output basis state (site 1 most significant in TWO_QUBIT_OPS's
    matrices), as op_matrix builds it in examples/gate_evolution.jl."""
    from .itensor import itensor
    if len(s) == 1:
        m = QUBIT_OPS[name]
        dt = "complex128" if np.iscomplexobj(m) else "float64"
        return itensor(m, s[0].prime(), s[0], dtype=dt).roc()
    if len(s) == 2:
        m = TWO_QUBIT_OPS[name].reshape(2, 2, 2, 2)      # (o1, o2, i1, i2)
        return itensor(m, s[0].prime(), s[1].prime(), s[0], s[1]).roc()
    raise _lib.ArgumentError("op: one- and two-site gates only")


def ops_list(sites, specs):
    """ops(s, [("X", n), ...]) with 1-based site numbers, as in ITensors."""
    return [op(name, *[sites[n - 1] for n in ns]) for (name, *ns) in specs]


def apply(gates, psi, maxdim=None, cutoff=0.0):
    """apply(gates, psi): each one-site gate multiplies its site tensor; a
    two-site gate on neighbours j, j+1 acts on their product, which a
    (truncated) SVD splits back."""
    psi = psi.copy()
    N = len(psi)
    where = {}
    for j in range(N):
        for s in psi.siteinds(j):
            where[s] = j
    for G in gates:
        sites = [i for i in G.inds if i.plev == 0 and i in where]
        js = sorted(where[s] for s in sites)
        if len(js) == 1:
            j = js[0]
            psi[j] = contract(G, psi[j]).noprime()
        elif len(js) == 2 and js[1] == js[0] + 1:
            j = js[0]
            phi = contract(G, contract(psi[j], psi[j + 1])).noprime()
            left = [i for i in psi[j].inds if i != psi.linkind(j)]
            U, S, V, _ = svd(phi, left, maxdim=maxdim, cutoff=cutoff or 0.0,
                             lefttags="Link,l=%d" % (j + 1), righttags="Link,l=%d" % (j + 1))
            psi[j] = U
            psi[j + 1] = contract(S, V)
        else:
            raise _lib.ArgumentError("apply: gates act on one site or two neighbouring sites")
    return psi
