"""ITensor-level host mirror over libitg: Index, ITensor with `roc`/`cpu`
storage conversion, contraction by index identity, permutation, addition,
norm/scalar, and the order-2 factorisations used by ITensors (svd, eigen,
qr, factorize).

Mirrors the reference's ITensor-level GPU surface: `cuITensor`/`cu`/`cpu`
(src/cuitensor.jl:1-33), `randomCuITensor` (src/cuitensor.jl:35-49),
`permute`, `*`, `+`, `-` (src/tensor/cudense.jl:83-110,333-500) and
`svd`/`eigen`/`qr` (src/tensor/culinearalgebra.jl:33-121), with `roc` in place
of `cu`. Compute runs only on device storage: operations on host tensors
raise (there is no CPU path here).
"""
import itertools

import numpy as np

from . import _lib
from . import ops

_ids = itertools.count(1)


class Index:
    """Index identity = (id, plev); tags are labels only."""

    __slots__ = ("dim", "id", "tags", "plev")

    def __init__(self, dim, tags="", plev=0, id=None):
        self.dim = int(dim)
        self.id = next(_ids) if id is None else id
        self.tags = tags
        self.plev = plev

    def prime(self, n=1):
        return Index(self.dim, self.tags, self.plev + n, self.id)

    def noprime(self):
        return Index(self.dim, self.tags, 0, self.id)

    def sim(self):
        return Index(self.dim, self.tags, self.plev)

    def __eq__(self, o):
        return isinstance(o, Index) and o.id == self.id and o.plev == self.plev

    def __hash__(self):
        return hash((self.id, self.plev))

    def __repr__(self):
        return "Index(%d|%s|%s%s)" % (self.dim, self.id, self.tags, "'" * self.plev)


def _dtype_name(t):
    return str(t.dtype).split(".")[-1]


class ITensor:
    """Dense tensor: a flat column-major buffer over `inds` (first fastest),
    on the host (numpy) or on the device (torch tensor on the GPU)."""

    def __init__(self, inds, data):
        self.inds = tuple(inds)
        self.data = data

    # -------------------------------------------------------------- basics
    @property
    def dims(self):
        return tuple(i.dim for i in self.inds)

    @property
    def ondevice(self):
        return not isinstance(self.data, np.ndarray)

    def order(self):
        return len(self.inds)

    def __repr__(self):
        return "ITensor(%s, %s)" % (list(self.inds), "device" if self.ondevice else "host")

    def _need_device(self):
        if not self.ondevice:
            raise _lib.ArgumentError("compute on ITensor storage needs roc(T) first")

    # --------------------------------------------------------- conversion
    def roc(self):
        if self.ondevice:
            return self
        return ITensor(self.inds, ops.to_device(self.data, str(self.data.dtype)))

    def cpu(self):
        if not self.ondevice:
            return self
        return ITensor(self.inds, ops.to_host(self.data))

    def array(self, *inds):
        """Host array with axes in the order `inds` (default: own order)."""
        h = self.cpu()
        a = np.reshape(h.data, h.dims, order="F") if h.inds else np.reshape(h.data, ())
        if inds:
            a = np.transpose(a, [list(self.inds).index(i) for i in inds])
        return a

    # ----------------------------------------------------------- algebra
    def permute(self, *inds):
        self._need_device()
        if len(inds) != len(self.inds) or set(inds) != set(self.inds):
            raise _lib.DimensionMismatch("permute: index sets differ")
        perm = [list(self.inds).index(i) for i in inds]
        if perm == list(range(len(perm))):
            return ITensor(self.inds, ops.copy(self.data))
        B, _ = ops.permute(self.data, self.dims, perm)
        return ITensor(inds, B)

    def __mul__(self, other):
        if isinstance(other, (int, float, complex, np.number)):
            self._need_device()
            out = ops.copy(self.data)
            if isinstance(other, (complex, np.complexfloating)) and not out.is_complex():
                out = ops.convert(out, "complex128")
            ops.scal(out, other)
            return ITensor(self.inds, out)
        return contract(self, other)

    __rmul__ = lambda self, x: self.__mul__(x)

    def __truediv__(self, x):
        return self * (1.0 / float(x))

    def _aligned(self, other):
        if set(self.inds) != set(other.inds) or len(self.inds) != len(other.inds):
            raise _lib.DimensionMismatch("add: index sets differ")
        return [list(other.inds).index(i) for i in self.inds]

    def __add__(self, other):
        self._need_device()
        other._need_device()
        perm = self._aligned(other)
        out = ops.copy(self.data)
        ops.add_permuted(other.data, other.dims, perm, 1.0, out, 1.0)
        return ITensor(self.inds, out)

    def __sub__(self, other):
        self._need_device()
        other._need_device()
        perm = self._aligned(other)
        out = ops.copy(self.data)
        ops.add_permuted(other.data, other.dims, perm, -1.0, out, 1.0)
        return ITensor(self.inds, out)

    def norm(self):
        self._need_device()
        return ops.nrm2(self.data)

    def scalar(self):
        if self.inds:
            raise _lib.DimensionMismatch("scalar of a tensor with indices")
        v = self.cpu().data[0]
        return complex(v) if np.iscomplexobj(v) else float(v)

    def prime(self, n=1, inds=None):
        sel = set(inds) if inds is not None else None
        return ITensor([i.prime(n) if sel is None or i in sel else i for i in self.inds], self.data)

    def noprime(self):
        return ITensor([i.noprime() for i in self.inds], self.data)

    def dag(self):
        """Complex conjugate (the reference's dag on Dense storage; index
        arrows do not exist in this dense mirror). Real storage is shared."""
        if not self.is_complex():
            return self
        if not self.ondevice:
            return ITensor(self.inds, np.conj(self.data))
        return ITensor(self.inds, ops.conj(ops.copy(self.data)))

    def is_complex(self):
        return bool(np.iscomplexobj(self.data)) if not self.ondevice else self.data.is_complex()

    def complex(self):
        """complex(T): ComplexF64 copy of real storage."""
        if self.is_complex():
            return self
        if not self.ondevice:
            return ITensor(self.inds, self.data.astype(np.complex128))
        return ITensor(self.inds, ops.convert(self.data, "complex128"))


# ---------------------------------------------------------------- factories
def itensor(array, *inds, dtype="float64"):
    """Host ITensor from an array whose axes follow `inds`."""
    a = np.asarray(array, dtype=dtype)
    if a.size != int(np.prod([i.dim for i in inds])):
        raise _lib.DimensionMismatch("array size does not match the indices")
    return ITensor(inds, np.ravel(np.reshape(a, [i.dim for i in inds], order="F"), order="F"))


def roc_itensor(value_or_array, *inds, dtype="float64"):
    """cuITensor analogue: device ITensor from a scalar fill or an array
    (src/cuitensor.jl:1-26)."""
    n = int(np.prod([i.dim for i in inds])) if inds else 1
    if np.isscalar(value_or_array):
        t = ops.empty(n, dtype)
        ops.fill(t, float(value_or_array))
        return ITensor(inds, t)
    return itensor(value_or_array, *inds, dtype=dtype).roc()


def random_roc_itensor(*inds, seed=0, offset=0, dtype="float64"):
    """randomCuITensor analogue: Gaussian entries from the Philox stream."""
    n = int(np.prod([i.dim for i in inds])) if inds else 1
    return ITensor(inds, ops.randn(n, seed, offset, dtype))


def roc(x):
    return x.roc()


def cpu(x):
    return x.cpu()


# ---------------------------------------------------------------- Diag
class DiagITensor:
    """Diag storage (NDTensors.Diag on the device, src/tensor/cudiag.jl):
    only the diagonal of a tensor over `inds` is stored, as a device vector
    of length min(dims); a plain number makes it a uniform diagonal
    (NDTensors.Diag(2.0))."""

    def __init__(self, inds, data):
        self.inds = tuple(inds)
        self.data = data

    @property
    def dims(self):
        return tuple(i.dim for i in self.inds)

    @property
    def uniform(self):
        return not hasattr(self.data, "numel")

    ondevice = True

    def is_complex(self):
        if self.uniform:
            return isinstance(self.data, complex)
        return self.data.is_complex()

    def complex(self):
        if self.is_complex():
            return self
        if self.uniform:
            return DiagITensor(self.inds, complex(self.data))
        return DiagITensor(self.inds, ops.convert(self.data, "complex128"))

    def _need_device(self):
        pass

    def dense(self, dtype=None):
        """The dense tensor with this diagonal (NDTensors.dense)."""
        dt = dtype or ("complex128" if self.is_complex() else "float64")
        one = ops.fill(ops.empty(1, dt), 1.0)
        lab = list(range(1, len(self.inds) + 1))
        C, _ = ops.diag_contract(None if self.uniform else self.data, self.dims, lab, one, (), (),
                                 lab, uniform=self.data if self.uniform else None, dtype=dt)
        return ITensor(self.inds, C)

    def array(self, *inds):
        return self.dense().array(*inds)

    def __mul__(self, other):
        if isinstance(other, (int, float, complex, np.number)):
            if self.uniform:
                return DiagITensor(self.inds, self.data * other)
            out = ops.copy(self.data)
            if isinstance(other, (complex, np.complexfloating)) and not out.is_complex():
                out = ops.convert(out, "complex128")
            ops.scal(out, other)
            return DiagITensor(self.inds, out)
        return contract(self, other)

    __rmul__ = lambda self, x: self.__mul__(x)


def diag_itensor(values, *inds):
    """Diag ITensor with the given diagonal (host array, uploaded) or a
    uniform value (a number)."""
    if np.isscalar(values):
        return DiagITensor(inds, values)
    v = np.asarray(values)
    if v.size != min(i.dim for i in inds):
        raise _lib.DimensionMismatch("diagonal length must be the smallest dimension")
    return DiagITensor(inds, ops.to_device(v, "complex128" if np.iscomplexobj(v) else "float64"))


def _as_complex(T):
    return T.complex() if hasattr(T, "complex") else T


def _contract_diag(A, B):
    if A.is_complex() != B.is_complex():
        A, B = _as_complex(A), _as_complex(B)
    out_inds = [i for i in A.inds if i not in B.inds] + [i for i in B.inds if i not in A.inds]
    for i in A.inds:
        if i in B.inds and B.inds[B.inds.index(i)].dim != i.dim:
            raise _lib.DimensionMismatch("contract: index dimension mismatch")
    if isinstance(A, DiagITensor) and isinstance(B, DiagITensor):
        # Diag x Diag: the diagonals multiply (cudiag.jl:118-140)
        if A.uniform and B.uniform:
            data = A.data * B.data
        elif A.uniform or B.uniform:
            u, v = (A.data, B.data) if A.uniform else (B.data, A.data)
            data = ops.scal(ops.copy(v), u)
        else:
            if A.data.numel() != B.data.numel():
                raise _lib.DimensionMismatch("Diag x Diag: diagonal lengths differ")
            data = ops.hadamard(A.data, B.data)
        if out_inds:
            return DiagITensor(out_inds, data)
        # every index contracted: the trace of the product
        if not hasattr(data, "numel"):
            return ITensor([], ops.to_device(np.array([data * min(A.dims)]),
                                             "complex128" if isinstance(data, complex)
                                             else "float64"))
        ones = ops.fill(ops.empty(data.numel(), ops.dtype_name(data)), 1.0)
        s = ops.dot(ones, data)
        return ITensor([], ops.to_device(np.array([s]), ops.dtype_name(data)))
    D, T = (A, B) if isinstance(A, DiagITensor) else (B, A)
    T._need_device()
    labels = {}
    for i in list(A.inds) + list(B.inds):
        labels.setdefault(i, len(labels) + 1)
    C, _ = ops.diag_contract(None if D.uniform else D.data, D.dims, [labels[i] for i in D.inds],
                             T.data, T.dims, [labels[i] for i in T.inds],
                             [labels[i] for i in out_inds],
                             uniform=D.data if D.uniform else None)
    return ITensor(out_inds, C)


# ------------------------------------------------------------ Combiner
class Combiner:
    """combiner(i, j, ...): fuses its indices into one index c (first index
    fastest, NDTensors' combined-index order); contracting it with a tensor
    holding all of them combines, with one holding c splits. On the device a
    combine is one permutation (cucombiner.jl: Combiner x CuDense promotes
    to CuDense)."""

    def __init__(self, inds, tags="CMB,Link"):
        self.uncombined = tuple(inds)
        self.c = Index(int(np.prod([i.dim for i in inds])), tags)
        self.inds = (self.c,) + self.uncombined

    def combinedind(self):
        return self.c


def combiner(*inds, tags="CMB,Link"):
    return Combiner(inds, tags)


def _contract_combiner(A, B):
    Cm, T = (A, B) if isinstance(A, Combiner) else (B, A)
    T._need_device()
    if all(i in T.inds for i in Cm.uncombined):
        rest = [i for i in T.inds if i not in Cm.uncombined]
        P = T.permute(*(list(Cm.uncombined) + rest))
        return ITensor([Cm.c] + rest, P.data)
    if Cm.c in T.inds:
        rest = [i for i in T.inds if i != Cm.c]
        P = T.permute(*([Cm.c] + rest))
        return ITensor(list(Cm.uncombined) + rest, P.data)
    raise _lib.DimensionMismatch("combiner: the tensor holds neither its indices nor the "
                                 "combined index")


# ------------------------------------------------------------- contraction
def contract(A, B, mode="native"):
    """A*B: common indices summed; result indices = A's unique then B's
    unique (NDTensors contract_labels; contract!! dispatch of
    src/tensor/cudense.jl:83-110, one GETT kernel for every branch). Diag
    and Combiner operands take their own paths."""
    if isinstance(A, DiagITensor) or isinstance(B, DiagITensor):
        return _contract_diag(A, B)
    if isinstance(A, Combiner) or isinstance(B, Combiner):
        return _contract_combiner(A, B)
    if A.is_complex() != B.is_complex():
        A, B = A.complex(), B.complex()
    A._need_device()
    B._need_device()
    labels = {}
    for i in list(A.inds) + list(B.inds):
        labels.setdefault(i, len(labels) + 1)
    la = [labels[i] for i in A.inds]
    lb = [labels[i] for i in B.inds]
    out_inds = [i for i in A.inds if i not in B.inds] + [i for i in B.inds if i not in A.inds]
    for i in A.inds:
        if i in B.inds and B.inds[B.inds.index(i)].dim != i.dim:
            raise _lib.DimensionMismatch("contract: index dimension mismatch")
    C, _ = ops.contract(A.data, A.dims, la, B.data, B.dims, lb, [labels[i] for i in out_inds],
                        mode=mode)
    return ITensor(out_inds, C)


# ------------------------------------------------------------ factorisations
def _matricize(A, Linds):
    Linds = [i for i in A.inds if i in set(Linds)]
    Rinds = [i for i in A.inds if i not in set(Linds)]
    if not Linds or not Rinds:
        raise _lib.DimensionMismatch("factorisation needs left and right indices")
    P = A.permute(*(Linds + Rinds)) if list(A.inds) != Linds + Rinds else ITensor(A.inds, ops.copy(A.data))
    m = int(np.prod([i.dim for i in Linds]))
    n = int(np.prod([i.dim for i in Rinds]))
    return P, Linds, Rinds, m, n


class Spectrum:
    def __init__(self, eigs, truncerr):
        self.eigs = eigs
        self.truncerr = truncerr


def svd(A, Linds, maxdim=None, mindim=1, cutoff=0.0, absolute_cutoff=False,
        relative_cutoff=True, lefttags="Link,u", righttags="Link,v"):
    """U, S, V, spec with A ~ U*S*V (culinearalgebra.jl:33-72); S is a dense
    k x k diagonal ITensor as in the reference."""
    P, Li, Ri, m, n = _matricize(A, Linds)
    U, S, V, k, te, _ = ops.svd_trunc(P.data, m, n, maxdim=maxdim, mindim=mindim, cutoff=cutoff,
                                      absolute_cutoff=absolute_cutoff,
                                      relative_cutoff=relative_cutoff)
    u, v = Index(k, lefttags), Index(k, righttags)
    s = ops.to_host(S)[:k]
    dt = _dtype_name(A.data)
    Ut = ITensor(Li + [u], U[: m * k])
    Vt = ITensor(Ri + [v], V[: n * k])
    St = itensor(np.diag(s), u, v, dtype=dt).roc()
    return Ut, St, Vt, Spectrum(s ** 2, te)


def eigen(A, Linds, Rinds, maxdim=None, cutoff=None, absolute_cutoff=False,
          relative_cutoff=True, tags="Link,eigen"):
    """Hermitian eigen: D (k), U (Rinds', r) with A ~ U D U^T, descending
    (culinearalgebra.jl:74-108). Returns (D, U, spec)."""
    P, Li, Ri, m, n = _matricize(A, Linds)
    if m != n:
        raise _lib.DimensionMismatch("eigen needs a square matricisation")
    w, U, k, te, _ = ops.eigh_trunc(P.data, m, maxdim=maxdim,
                                    cutoff=-1.0 if cutoff is None else cutoff,
                                    absolute_cutoff=absolute_cutoff,
                                    relative_cutoff=relative_cutoff)
    r = Index(k, tags)
    wh = ops.to_host(w)[:k]
    return wh, ITensor(Li + [r], U[: m * k]), Spectrum(wh, te)


def qr(A, Linds, tags="Link,qr"):
    """Thin QR (culinearalgebra.jl:110-121)."""
    P, Li, Ri, m, n = _matricize(A, Linds)
    Q, R = ops.qr(P.data, m, n)
    q = Index(min(m, n), tags)
    return ITensor(Li + [q], Q), ITensor([q] + Ri, R)


def factorize(A, Linds, ortho="left", which_decomp=None, maxdim=None, mindim=1, cutoff=None,
              tags="Link,fact"):
    """ITensors factorize: "eigen" (density matrix) when cutoff >= 1e-12,
    else "svd"; returns (L, R, spec) with L or R orthonormal per `ortho`."""
    if which_decomp is None:
        which_decomp = "svd" if cutoff is None or cutoff < 1e-12 else "eigen"
    Li = [i for i in A.inds if i in set(Linds)]
    Ri = [i for i in A.inds if i not in set(Linds)]
    if which_decomp == "svd":
        U, S, V, spec = svd(A, Li, maxdim=maxdim, mindim=mindim,
                            cutoff=0.0 if cutoff is None else cutoff, lefttags=tags,
                            righttags=tags)
        if ortho == "left":
            return U, contract(S, V), spec
        return contract(U, S), V, spec
    if which_decomp == "eigen":
        # rho = A dag(A') with the other side summed: rho(l, l') =
        # sum_r A(l, r) conj(A(l', r)); rho = U D U^H, the other factor dag(U) A
        if ortho == "left":
            Ap = A.prime(1, Li)
            rho = contract(A, Ap.dag())     # (Li, Li') with the right indices summed
            D, U, spec = eigen(rho.noprime_pairs(Li), Li, [i.prime() for i in Li],
                               maxdim=maxdim, cutoff=cutoff, tags=tags)
            return U, contract(U.dag(), A), spec
        Ap = A.prime(1, Ri)
        rho = contract(A, Ap.dag())
        D, U, spec = eigen(rho.noprime_pairs(Ri), Ri, [i.prime() for i in Ri], maxdim=maxdim,
                           cutoff=cutoff, tags=tags)
        return contract(A, U.dag()), U, spec
    raise _lib.ArgumentError("factorize: unknown which_decomp %r" % (which_decomp,))


def _noprime_pairs(self, inds):
    """rho(i..', i..) -> indices (i.., i..') with the unprimed set first, as
    eigen(rho, Linds, Linds') expects."""
    order = [i for i in self.inds if i in set(inds)] + [i for i in self.inds if i not in set(inds)]
    return self.permute(*order)


ITensor.noprime_pairs = _noprime_pairs


# ------------------------------------------------------------------ davidson
def davidson(M, v, maxiter=2, errgoal=1e-14):
    """davidson(M, v; maxiter) of ITensors' iterative solvers for a Hermitian
    map M: ITensor -> ITensor with v's indices (test/test_cuiterativesolvers.jl:
    ITensorMap), on real or complex storage. Returns (lambda, v)."""
    from .mps import davidson as _dav
    v._need_device()
    inds = v.inds

    def mv(x):
        y = M(ITensor(inds, x))
        if list(y.inds) != list(inds):
            y = y.permute(*inds)
        return y.data

    lam, phi = _dav(mv, v.data, maxiter=maxiter, errgoal=errgoal)
    return lam, ITensor(inds, phi)
