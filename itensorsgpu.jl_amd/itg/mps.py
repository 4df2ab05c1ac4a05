"""MPS/MPO on device storage, ProjMPO, davidson and two-site dmrg.

Host orchestration in the shape of ITensors 0.2's `dmrg(H, psi0, sweeps)`
(reached by the reference through test/dmrg.jl:27,75 and
examples/dmrg.jl:25), with every FLOP in libitg:
  - sweepnext: bonds 1..N-1 left to right then N-1..1 (2(N-1) per sweep)
  - ProjMPO product: the fused two-site Heff apply (itg_heff2_apply)
  - davidson: Krylov/Davidson on device vectors (FP64 dots, norms and
    orthogonalisation), host-side projected eigenproblem, optional
    diagonal or learned preconditioner with a revert guard
  - replacebond!: factorize with "eigen" (density matrix) when
    cutoff >= 1e-12 else "svd", then normalisation of the centre
  - environments: L.A.W.A^dag / A.W.A^dag.R contractions
Tensor index orders: MPS site (l, s, r) with dimension-1 boundary links;
MPO site (wl, wr, s', s) with wl = 1 on the first and wr = 1 on the last site;
environments L (a', w, a) and R (b', w, b).
`mode` selects the arithmetic of the Heff applies ("native" = FP64 MFMA,
"fp32", "bf16x3"); storage, environments (whose errors would accumulate
along the chain), Davidson vector algebra, the density matrix, its
eigendecomposition and the reported energy (an FP64 Rayleigh quotient at the
end of each sweep) stay FP64.
"""
import math
import time

import numpy as np

from . import _lib
from . import ops
from .itensor import ITensor, Index


# ---------------------------------------------------------------- operators
def spin_ops(S):
    d = int(round(2 * S + 1))
    m = S - np.arange(d)
    Sz = np.diag(m)
    Sp = np.zeros((d, d))
    for i in range(1, d):
        Sp[i - 1, i] = math.sqrt(S * (S + 1) - m[i] * (m[i] + 1))
    Sm = Sp.T.copy()
    return {"Sz": Sz, "S+": Sp, "S-": Sm, "Sx": 0.5 * (Sp + Sm), "Sy": -0.5j * (Sp - Sm),
            "Id": np.eye(d)}


class OpSum:
    """AutoMPO-style sum of one- and two-site terms (the `ampo` of
    test/dmrg.jl:9-14,64-68), compiled to an MPO by a finite-state
    construction: the channel (i, op) carries op_i from site i to every
    partner j > i of a term (i, op, j, op'), whose coefficient and op' are
    applied at j. Bond dimension at a cut = 2 + number of open channels
    (5 for the Heisenberg chain)."""

    def __init__(self):
        self.one = []   # (coef, op, i)
        self.two = []   # (coef, opA, i, opB, j)

    def add(self, coef, op1, i, op2=None, j=None):
        coef = complex(coef) if np.iscomplexobj(coef) and complex(coef).imag != 0 else float(
            np.real(coef))
        if op2 is None:
            self.one.append((coef, op1, int(i)))
        else:
            i, j = int(i), int(j)
            if i == j:
                raise _lib.ArgumentError("two-site term on one site")
            if i > j:
                op1, op2, i, j = op2, op1, j, i
            self.two.append((coef, op1, i, op2, j))
        return self

    def is_complex(self):
        """Complex coefficients or an Sy operator: the MPO is ComplexF64."""
        terms = [(c, o) for (c, o, _) in self.one] + \
            [(c, a) for (c, a, _, _, _) in self.two] + [(c, b) for (c, _, _, b, _) in self.two]
        return any(isinstance(c, complex) or o == "Sy" for c, o in terms)

    def to_mpo_arrays(self, N, S=0.5):
        ops_ = spin_ops(S)
        d = ops_["Id"].shape[0]
        # channels at cut k (between sites k and k+1)
        chans = []
        for k in range(N - 1):
            cs = []
            for (_, opA, i, _, j) in self.two:
                if i <= k < j and (i, opA) not in cs:
                    cs.append((i, opA))
            chans.append(cs)
        Ws = []
        for k in range(N):
            left = ["done"] + (chans[k - 1] if k > 0 else []) + ["start"] if k > 0 else ["start"]
            right = ["done"] + (chans[k] if k < N - 1 else []) + ["start"] if k < N - 1 else ["done"]
            W = np.zeros((len(left), len(right), d, d),
                         dtype=np.complex128 if self.is_complex() else np.float64)
            li = {c: a for a, c in enumerate(left)}
            ri = {c: a for a, c in enumerate(right)}
            I = ops_["Id"]
            if "start" in li and "start" in ri:
                W[li["start"], ri["start"]] += I
            if "done" in li and "done" in ri:
                W[li["done"], ri["done"]] += I
            for (c, op, i) in self.one:
                if i == k:
                    W[li["start"], ri["done"]] += c * ops_[op]
            for (c, opA, i, opB, j) in self.two:
                if i == k:
                    W[li["start"], ri[(i, opA)]] = ops_[opA]
                elif j == k:
                    W[li[(i, opA)], ri["done"]] += c * ops_[opB]
            # channels passing through site k
            for ch in (chans[k - 1] if k > 0 else []):
                if k < N - 1 and ch in ri:
                    W[li[ch], ri[ch]] = I
            Ws.append(W)
        return Ws


def heisenberg_opsum(bonds, J=1.0):
    os_ = OpSum()
    for (i, j) in bonds:
        os_.add(J, "Sz", i, "Sz", j)
        os_.add(0.5 * J, "S+", i, "S-", j)
        os_.add(0.5 * J, "S-", i, "S+", j)
    return os_


def tfim_opsum(N, J=1.0, h=0.5):
    os_ = OpSum()
    for j in range(N):
        if j < N - 1:
            os_.add(-J, "Sz", j, "Sz", j + 1)
        os_.add(-h, "Sx", j)
    return os_


def cylinder_bonds(Lx, Ly, snake=True):
    def site(x, y):
        if snake and x % 2 == 1:
            y = Ly - 1 - y
        return x * Ly + y

    bonds = set()
    for x in range(Lx):
        for y in range(Ly):
            s = site(x, y)
            if x + 1 < Lx:
                bonds.add(tuple(sorted((s, site(x + 1, y)))))
            t = site(x, (y + 1) % Ly)
            if t != s:
                bonds.add(tuple(sorted((s, t))))
    return sorted(bonds)


# ------------------------------------------------------------- containers
class MPO:
    """Device MPO: per site a flat buffer with dims (wl, wr, d, d); complex
    arrays give a ComplexF64 MPO."""

    def __init__(self, arrays, dtype=None):
        if dtype is None:
            dtype = "complex128" if any(np.iscomplexobj(W) for W in arrays) else "float64"
        self.dims = [tuple(int(x) for x in W.shape) for W in arrays]
        self.data = [ops.to_device(np.ravel(W, order="F"), dtype) for W in arrays]

    def __len__(self):
        return len(self.data)

    def is_complex(self):
        return self.data[0].is_complex()

    def complex(self):
        """The same operator on ComplexF64 storage."""
        if self.is_complex():
            return self
        out = MPO.__new__(MPO)
        out.dims = list(self.dims)
        out.data = [ops.convert(t, "complex128") for t in self.data]
        return out

    @staticmethod
    def from_opsum(os_, N, S=0.5, dtype=None):
        return MPO(os_.to_mpo_arrays(N, S), dtype)


class MPS:
    """Device MPS: per site a flat buffer with dims (l, s, r)."""

    def __init__(self, data, dims, llim=0, rlim=None):
        self.data = list(data)
        self.dims = [tuple(x) for x in dims]
        self.llim = llim
        self.rlim = len(data) + 1 if rlim is None else rlim

    def __len__(self):
        return len(self.data)

    def maxlinkdim(self):
        return max(dm[2] for dm in self.dims)

    def is_complex(self):
        return self.data[0].is_complex()

    def complex(self):
        """complex(psi): the same state on ComplexF64 storage."""
        if self.is_complex():
            return self
        return MPS([ops.convert(t, "complex128") for t in self.data], self.dims, self.llim,
                   self.rlim)

    def to_host(self):
        return [np.reshape(ops.to_host(t), dm, order="F") for t, dm in zip(self.data, self.dims)]

    @staticmethod
    def from_host(arrays, dtype="float64"):
        return MPS([ops.to_device(np.ravel(a, order="F"), dtype) for a in arrays],
                   [a.shape for a in arrays])

    def itensors(self, sites=None):
        """ITensor view with fresh link indices (and given site indices)."""
        N = len(self)
        links = [Index(self.dims[0][0], "Link,l=0")] + \
                [Index(self.dims[j][2], "Link,l=%d" % (j + 1)) for j in range(N)]
        sites = sites or [Index(dm[1], "Site,n=%d" % j) for j, dm in enumerate(self.dims)]
        return [ITensor((links[j], sites[j], links[j + 1]), self.data[j]) for j in range(N)]


def link_dims(N, d, chi):
    return [1] + [int(min(chi, d ** min(j, N - j))) for j in range(1, N)] + [1]


def random_mps(N, d, chi, seed, dtype="float64"):
    """Gaussian site tensors drawn in site order from one Philox stream
    (itg_randn), right-orthogonalised by device QR and normalised: the
    orthogonality centre is site 0 (randomCuMPS + orthogonalize!, cumps.jl:44-53)."""
    dims = link_dims(N, d, chi)
    data, shp, off = [], [], 0
    for j in range(N):
        n = dims[j] * d * dims[j + 1]
        t = ops.randn(n, seed, off, dtype)
        ops.scal(t, 1.0 / ops.nrm2(t))          # normalize!(M[i]), cumps.jl:48
        data.append(t)
        shp.append((dims[j], d, dims[j + 1]))
        off += n
    psi = MPS(data, shp)
    right_orthogonalize(psi)
    nrm = ops.nrm2(psi.data[0])
    ops.scal(psi.data[0], 1.0 / nrm)
    return psi


def right_orthogonalize(psi):
    """orthogonalize!(psi, 1): LQ of every site from the right end."""
    for j in range(len(psi) - 1, 0, -1):
        dl, d, dr = psi.dims[j]
        Mt, _ = ops.permute(psi.data[j], (dl, d * dr), (1, 0))       # (d dr) x dl
        Q, R = ops.qr(Mt, d * dr, dl)                                 # Q: (d dr) x k
        k = min(d * dr, dl)
        A, _ = ops.permute(Q, (d * dr, k), (1, 0))                    # k x (d dr)
        psi.data[j] = A
        psi.dims[j] = (k, d, dr)
        pl, pd, _ = psi.dims[j - 1]
        out = ops.empty(pl * pd * k, str(A.dtype).split(".")[-1])
        ops.gemm("N", "T", pl * pd, k, dl, 1.0, psi.data[j - 1], pl * pd, R, k, 0.0, out, pl * pd)
        # keep the running product O(1): QR is invariant under positive
        # scaling, and the centre is normalised at the end
        ops.scal(out, 1.0 / ops.nrm2(out))
        psi.data[j - 1] = out
        psi.dims[j - 1] = (pl, pd, k)
    psi.llim, psi.rlim = 0, 2


# ------------------------------------------------------------------ timing
class PhaseTimer:
    """Accumulates device time per phase with events recorded on the stream
    libitg runs on, plus the algorithmic FLOPs of each timed call.
    `factorize_kinds` (libitg kernel classes) are additionally event-timed
    during the factorisation of bond `factorize_bond` when the sweep passes
    it left to right (the library's own per-launch timing, added to whatever
    classes are being timed already)."""

    def __init__(self, factorize_bond=-1, factorize_kinds=()):
        self.pairs = {}
        self.flops = {}
        self.calls = {}
        self.factorize_bond = factorize_bond
        self.factorize_kinds = tuple(factorize_kinds)
        self.base_kinds = ()

    def start(self):
        import torch
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        return ev

    def start_at(self, ev):
        """An existing start event, for a second phase ending at the next stop."""
        return ev

    def stop(self, name, ev0, flops=0.0):
        import torch
        ev1 = torch.cuda.Event(enable_timing=True)
        ev1.record()
        self.pairs.setdefault(name, []).append((ev0, ev1))
        self.flops[name] = self.flops.get(name, 0.0) + flops
        self.calls[name] = self.calls.get(name, 0) + 1

    def report(self):
        import torch
        torch.cuda.synchronize()
        out = {}
        for name, prs in self.pairs.items():
            ms = sum(a.elapsed_time(b) for a, b in prs)
            out[name] = {"ms": ms, "calls": self.calls[name], "flops": self.flops[name]}
        return out


def heff_flops(chiL, wl, wr, d, chiR):
    """Algorithmic FLOPs of one two-site Heff apply (C1 + W pass + C4)."""
    return (2.0 * chiL * wl * d * d * chiR * chiL + 2.0 * chiL * chiR * (wl * d * d) * (d * d * wr)
            + 2.0 * chiL * d * d * chiR * wr * chiR)


# ------------------------------------------------------------ environments
def ones_env(dtype="float64"):
    t = ops.empty(1, dtype)
    ops.fill(t, 1.0)
    return t, (1, 1, 1)


def _bra(A):
    """dag(A) on the storage: a conjugated copy for complex, A itself for real."""
    return ops.conj(ops.copy(A)) if A.is_complex() else A


def update_left(L, Ld, A, Ad, W, Wd, mode="native"):
    """L_new(r', wr, r) = L(l', wl, l) A(l, s, r) W(wl, wr, s', s) conj(A(l', s', r'))."""
    # T2's order (l', s', wr, r) makes the last contraction, like the first,
    # a plain GEMM: (l' s') contiguous, (wr r) one stride
    T1, d1 = ops.contract(L, Ld, [1, 2, 3], A, Ad, [3, 4, 5], [1, 2, 4, 5], mode=mode)
    T2, d2 = ops.contract(T1, d1, [1, 2, 4, 5], W, Wd, [2, 6, 7, 4], [1, 7, 6, 5])
    return ops.contract(T2, d2, [1, 7, 6, 5], _bra(A), Ad, [1, 7, 8], [8, 6, 5], mode=mode)


def update_right(R, Rd, A, Ad, W, Wd, mode="native"):
    """R_new(l', wl, l) = R(r', wr, r) A(l, s, r) W(wl, wr, s', s) conj(A(l', s', r'))."""
    # T2's order (s', r', wl, l) makes the last contraction a plain GEMM
    T1, d1 = ops.contract(R, Rd, [1, 2, 3], A, Ad, [4, 5, 3], [1, 2, 4, 5], mode=mode)
    T2, d2 = ops.contract(T1, d1, [1, 2, 4, 5], W, Wd, [6, 2, 7, 5], [7, 1, 6, 4])
    return ops.contract(T2, d2, [7, 1, 6, 4], _bra(A), Ad, [8, 7, 1], [8, 6, 4], mode=mode)


def heff_apply_generic(L, Ld, W1, W1d, W2, W2d, R, Rd, x, xdims):
    """The two-site Heff product as four FP64 contractions in the order
    L.psi -> .W1 -> .W2 -> .R, on any storage (the complex path):
    out(a', s1', s2', b') = L(a', wl, a) W1(wl, wm, s1', s1) W2(wm, wr, s2', s2)
    R(b', wr, b) psi(a, s1, s2, b)."""
    T1, d1 = ops.contract(L, Ld, [1, 2, 3], x, xdims, [3, 4, 5, 6], [1, 2, 4, 5, 6])
    T2, d2 = ops.contract(T1, d1, [1, 2, 4, 5, 6], W1, W1d, [2, 7, 8, 4], [1, 7, 8, 5, 6])
    T3, d3 = ops.contract(T2, d2, [1, 7, 8, 5, 6], W2, W2d, [7, 9, 10, 5], [1, 8, 10, 9, 6])
    y, _ = ops.contract(T3, d3, [1, 8, 10, 9, 6], R, Rd, [11, 9, 6], [1, 8, 10, 11])
    return y


class ProjMPO:
    """Left/right environments of H around a two-site window."""

    def __init__(self, H, mode="native", timer=None, env_mode="native"):
        self.H = H
        self.mode = mode            # Heff apply arithmetic
        self.env_mode = env_mode    # environment update arithmetic
        self.timer = timer
        N = len(H)
        self.L = [None] * (N + 1)
        self.R = [None] * (N + 1)
        self._prep = None        # (bond, mode, dims, buffer) of itg_heff2_prepare
        # grow-only buffers shared by every bond: one each for the prepared
        # operator, the Heff work and the Krylov basis, sized for the largest
        # shape seen so far (bonds of a truncating sweep all differ in shape,
        # so per-shape buffers would pile up)
        self._bufs = {}

    def buffer(self, name, nelem, dtype):
        """Grow-only buffer `name` of at least nelem elements of dtype."""
        b = self._bufs.get(name)
        if b is None or b.numel() < nelem or str(b.dtype).split(".")[-1] != dtype:
            self._bufs.pop(name, None)
            b = ops.empty(nelem, dtype)
            self._bufs[name] = b
        return b

    def heff_work(self, dims):
        if self.H.is_complex():
            return None     # the complex product allocates its intermediates
        return self.buffer("work", _lib.lib().itg_heff2_work_elems(*dims), "float32")

    def build_right(self, psi, upto=2):
        N = len(psi)
        dt = ops.dtype_name(self.H.data[0])
        self.R[N] = ones_env(dt)
        for j in range(N - 1, upto - 1, -1):
            R, Rd = self.R[j + 1]
            self.R[j] = update_right(R, Rd, psi.data[j], psi.dims[j], self.H.data[j],
                                     self.H.dims[j], self.env_mode)
        self.L[0] = ones_env(dt)

    def prepare(self, b, xdims):
        """Fused W1.W2 operator (and, in bf16x3, the split L and R) for the
        Krylov loop at bond b; product() uses it until the next prepare.
        Complex operators have no prepared form."""
        if self.H.is_complex():
            self._prep = None
            return
        L, Ld = self.L[b]
        R, Rd = self.R[b + 2]
        wl, wm = self.H.dims[b][0], self.H.dims[b][1]
        wr = self.H.dims[b + 1][1]
        chiL, d, _, chiR = xdims
        dims = (chiL, wl, wm, wr, d, chiR)
        buf = self.buffer("prep", _lib.lib().itg_heff2_prep_elems(*dims), "float32")
        buf = ops.heff2_prepare(L, self.H.data[b], self.H.data[b + 1], R, *dims, mode=self.mode,
                                prep=buf)
        self._prep = (b, self.mode, dims, buf, L, R)

    def eigsolve(self, b, phi, xdims, maxiter=2, errgoal=1e-14, diag=None, stats=None, net=None):
        """Native device Davidson (itg_davidson2) on the prepared bond b;
        phi is overwritten with the Ritz vector. The work buffers are the
        ProjMPO's grow-only ones, so bonds of one shape replay the split-BF16
        graphs captured for it."""
        L, Ld = self.L[b]
        R, Rd = self.R[b + 2]
        p = self._prep
        if p is None or p[0] != b or p[1] != self.mode or p[4] is not L or p[5] is not R:
            self.prepare(b, xdims)
            p = self._prep
        dims = p[2]
        work = self.heff_work(dims)
        kwork = self.buffer("kwork", _lib.lib().itg_davidson2_work_elems(dims[0], dims[4],
                                                                           dims[5], maxiter),
                            "float64")
        ev = self.timer.start() if self.timer is not None else None
        nrev = 0
        if net is not None:
            # the guarded learned correction inside the device loop
            mwork = self.buffer("mwork", _lib.lib().itg_davidson2_mlp_work_elems(
                dims[0], dims[4], dims[5], net.h), "float32")
            lam, nmv, nrev = ops.davidson2_mlp(L, R, p[3], phi, *dims, diag=diag,
                                               params=net.params, hidden=net.h, skip=net.skip,
                                               floor=net.floor, maxiter=maxiter, errgoal=errgoal,
                                               mode=self.mode, work=work, kwork=kwork,
                                               mwork=mwork,
                                               net_steps=getattr(net, "net_steps", 0))
        else:
            lam, nmv = ops.davidson2(L, R, p[3], phi, *dims, maxiter=maxiter, errgoal=errgoal,
                                     diag=diag, mode=self.mode, work=work, kwork=kwork)
        if ev is not None:
            self.timer.stop("davidson", ev, nmv * heff_flops(dims[0], dims[1], dims[3], dims[4],
                                                             dims[5]))
        if stats is not None:
            stats["matvecs"] = stats.get("matvecs", 0) + nmv
            stats["reverted"] = stats.get("reverted", 0) + nrev
        return lam, phi

    def product(self, b, x, xdims, out=None, work=None):
        L, Ld = self.L[b]
        R, Rd = self.R[b + 2]
        wl, wm = self.H.dims[b][0], self.H.dims[b][1]
        wr = self.H.dims[b + 1][1]
        chiL, d, _, chiR = xdims
        dims = (chiL, wl, wm, wr, d, chiR)
        ev = self.timer.start() if self.timer is not None else None
        p = self._prep
        if self.H.is_complex():
            if self.mode != "native":
                raise _lib.ArgumentError("complex Heff runs in FP64 (native) mode only")
            y = heff_apply_generic(L, Ld, self.H.data[b], self.H.dims[b], self.H.data[b + 1],
                                   self.H.dims[b + 1], R, Rd, x, xdims)
        elif (p is not None and p[0] == b and p[1] == self.mode and p[2] == dims and p[4] is L
                and p[5] is R):
            y = ops.heff2_apply_prepared(L, R, p[3], x, *dims, mode=self.mode, out=out, work=work)
        else:
            y = ops.heff2_apply(L, self.H.data[b], self.H.data[b + 1], R, x, *dims,
                                mode=self.mode, out=out, work=work)
        if ev is not None:
            self.timer.stop("heff", ev, heff_flops(chiL, wl, wr, d, chiR))
        return y

    def diag(self, b, xdims):
        if self.H.is_complex():
            raise _lib.ArgumentError("the Heff diagonal is built for real operators only")
        L, Ld = self.L[b]
        R, Rd = self.R[b + 2]
        wl, wm = self.H.dims[b][0], self.H.dims[b][1]
        wr = self.H.dims[b + 1][1]
        chiL, d, _, chiR = xdims
        return ops.heff2_diag(L, self.H.data[b], self.H.data[b + 1], R, chiL, wl, wm, wr, d, chiR)


# ----------------------------------------------------------------- davidson
def davidson(matvec, x0, maxiter=2, miniter=1, errgoal=1e-14, northo=2, precond=None,
             stats=None, fallback=None):
    """ITensors-style Davidson on device vectors. Returns (lambda, phi).
    precond(q, lam) -> correction (device) or None for the plain residual.
    With `fallback` (a preconditioner callable, or "plain"), a step taken
    with `precond` that does not lower the residual norm is reverted (its
    basis vector dropped, the previous Ritz pair restored) and taken again
    with the fallback; stats["reverted"] counts such steps. The subspace,
    its projected matrix and the Ritz values stay FP64 in every mode."""
    nrm = ops.nrm2(x0)
    phi = ops.copy(x0)
    ops.scal(phi, 1.0 / nrm)
    cplx = phi.is_complex()     # Hermitian problem on complex storage
    V = [phi]
    AV = [matvec(phi)]
    lam = ops.dot(V[0], AV[0]).real
    q = ops.copy(AV[0])
    ops.axpby(-lam, V[0], 1.0, q)
    M = np.array([[lam]], dtype=complex if cplx else float)
    last = float("nan")
    maxiter = min(maxiter, phi.numel() - 1)
    nmv, nrev = 1, 0
    qn = ops.nrm2(q)
    for ni in range(1, maxiter + 1):
        conv = (qn < errgoal and abs(lam - last) < errgoal) or qn < max(1e-12, errgoal * 1e-3)
        if qn < 1e-20 or (conv and ni > miniter):
            break
        last = lam
        pres = [precond] if fallback is None else [precond, fallback]
        for attempt, pre in enumerate(pres):
            t = None
            if pre is not None and pre != "plain":
                t = pre(q, lam)
            if t is None:
                t = ops.copy(q)
            for _ in range(northo):
                for v in V:
                    ops.axpby(-ops.dot(v, t), v, 1.0, t)
            tn = ops.nrm2(t)
            if tn < 1e-14 * max(1.0, qn):
                if attempt + 1 < len(pres):
                    nrev += 1
                    continue
                t = None
                break
            ops.scal(t, 1.0 / tn)
            V.append(t)
            AV.append(matvec(t))
            nmv += 1
            m = len(V)
            Mn = np.zeros((m, m), dtype=M.dtype)
            Mn[: m - 1, : m - 1] = M
            for k in range(m):
                Mn[k, m - 1] = ops.dot(V[k], AV[m - 1])
                Mn[m - 1, k] = np.conj(Mn[k, m - 1])
            ev, U = np.linalg.eigh(Mn)
            u = U[:, 0]
            if cplx:
                # phase so that the first coefficient is real and >= 0
                if abs(u[0]) > 0:
                    u = u * (np.conj(u[0]) / abs(u[0]))
            elif u[0] < 0:
                u = -u
            nphi = ops.copy(V[0])
            ops.scal(nphi, u[0])
            nq = ops.copy(AV[0])
            ops.scal(nq, u[0])
            for k in range(1, m):
                ops.axpby(u[k], V[k], 1.0, nphi)
                ops.axpby(u[k], AV[k], 1.0, nq)
            ops.axpby(-float(ev[0]), nphi, 1.0, nq)
            nqn = ops.nrm2(nq)
            if attempt + 1 < len(pres) and not nqn < qn:
                # the guarded step did not lower the residual: drop it
                V.pop()
                AV.pop()
                nrev += 1
                continue
            M, lam, phi, q, qn = Mn, float(ev[0]), nphi, nq, nqn
            break
        if t is None:
            break
    ops.scal(phi, 1.0 / ops.nrm2(phi))
    if stats is not None:
        stats["matvecs"] = stats.get("matvecs", 0) + nmv
        stats["reverted"] = stats.get("reverted", 0) + nrev
        stats["residual"] = qn
    return lam, phi


def diag_preconditioner(dg):
    """Davidson correction t = q / (diag(H) - lambda), regularised."""

    def pre(q, lam):
        return ops.precond_diag(q, dg, lam, 1e-8)

    return pre


# --------------------------------------------------------------- factorize
def noiseterm(PH, b, phi, pdims, ortho, mode="native"):
    """The density-matrix perturbation of ITensors' two-site dmrg with
    noise (`drho = noise * noiseterm(PH, phi, ortho)`, used through
    `noise!` at examples/dmrg.jl:24 and test/dmrg.jl:24): the half of the
    projected operator on the side being kept, applied to phi with its MPO
    link left open, then traced over everything but that side's indices.
      left:  nt(a', s1', w, s2, b) = L_b(a', wl, a) W_b(wl, w, s1', s1) phi(a, s1, s2, b)
             drho[(a s1), (a' s1')] = sum_{w s2 b} nt nt       (m x m, m = chiL d)
      right: nt(a, s1, w, s2', b') = phi(a, s1, s2, b) R(b', wr, b) W_{b+1}(w, wr, s2', s2)
             drho[(s2 b), (s2' b')] = sum_{a s1 w} nt nt       (n x n, n = d chiR)
    The large contraction and the Gram product run in `mode` (the term is
    scaled by the noise, 1e-5 and below); the result is exactly symmetric."""
    chiL, d1, d2, chiR = pdims
    dt = str(phi.dtype).split(".")[-1]
    if ortho == "left":
        L, Ld = PH.L[b]
        W, Wd = PH.H.data[b], PH.H.dims[b]
        T, dT = ops.contract(L, Ld, [1, 2, 3], phi, pdims, [3, 4, 5, 6], [1, 2, 4, 5, 6],
                             mode=mode)
        nt, dn = ops.contract(T, dT, [1, 2, 4, 5, 6], W, Wd, [2, 7, 8, 4], [1, 8, 7, 5, 6])
        m, K = chiL * d1, Wd[1] * d2 * chiR
        if nt.is_complex():
            # drho(i, i') = sum_K nt(i, K) conj(nt(i', K))
            drho, _ = ops.contract(nt, (m, K), [1, 2], _bra(nt), (m, K), [3, 2], [1, 3])
        else:
            drho = ops.empty(m * m, dt)
            ops.gemm("N", "T", m, m, K, 1.0, nt, m, nt, m, 0.0, drho, m, mode=mode)
    else:
        R, Rd = PH.R[b + 2]
        W, Wd = PH.H.data[b + 1], PH.H.dims[b + 1]
        T, dT = ops.contract(phi, pdims, [1, 2, 3, 4], R, Rd, [5, 6, 4], [1, 2, 3, 5, 6],
                             mode=mode)
        nt, dn = ops.contract(T, dT, [1, 2, 3, 5, 6], W, Wd, [7, 6, 8, 3], [1, 2, 7, 8, 5])
        n, K = d2 * chiR, chiL * d1 * Wd[0]
        if nt.is_complex():
            # drho(j, j') = sum_K conj(nt(K, j)) nt(K, j')
            drho, _ = ops.contract(_bra(nt), (K, n), [1, 2], nt, (K, n), [1, 3], [2, 3])
        else:
            drho = ops.empty(n * n, dt)
            ops.gemm("T", "N", n, n, K, 1.0, nt, K, nt, K, 0.0, drho, n, mode=mode)
        m = n
    # (X + X^T)/2: the split-BF16 product sums its three partial products
    # in a different order for (i, j) and (j, i), so only this makes the
    # term exactly symmetric in every mode ((X + X^H)/2: exactly Hermitian)
    sym, _ = ops.permute(drho, (m, m), (1, 0), alpha=0.5)
    ops.conj(sym)
    return ops.add_permuted(drho, (m, m), (0, 1), 0.5, sym, 1.0)


def _replacebond_complex(phi, m, n, ortho, maxdim, mindim, cutoff, drho, noise, use_eigen):
    """Complex phi (m x n). Density-matrix route: left, rho = phi phi^H =
    U D U^H, A = U, B = U^H phi; right, rho = phi^H phi = V D V^H, B = V^H,
    A = phi V. SVD route (phi = U S V^H): left A = U, B = S V^H; right
    A = U S, B = V^H. The non-orthogonal factor is normalised, as in the
    real routes."""
    cut = 0.0 if cutoff is None else cutoff
    pb = _bra(phi)
    if not use_eigen:
        U, S, V, k, terr, _ = ops.svd_trunc(ops.copy(phi), m, n, maxdim=maxdim, mindim=mindim,
                                            cutoff=cut)
        s = ops.to_host(S)[:k]
        Sd = ops.to_device(np.ravel(np.diag(s).astype(np.complex128), order="F"), "complex128")
        Vh, _ = ops.permute(_bra(V[: n * k]), (n, k), (1, 0))          # k x n
        if ortho == "left":
            A = U[: m * k].clone()
            B, _ = ops.contract(Sd, (k, k), [1, 2], Vh, (k, n), [2, 3], [1, 3])
            ops.scal(B, 1.0 / ops.nrm2(B))
        else:
            B = Vh
            A, _ = ops.contract(U[: m * k], (m, k), [1, 2], Sd, (k, k), [2, 3], [1, 3])
            ops.scal(A, 1.0 / ops.nrm2(A))
        return A, B, k, terr
    if ortho == "left":
        rho, _ = ops.contract(phi, (m, n), [1, 2], pb, (m, n), [3, 2], [1, 3])
        if drho is not None:
            ops.axpby(noise, drho, 1.0, rho)
        w, U, k, terr, _ = ops.eigh_trunc(rho, m, maxdim=maxdim, cutoff=cut)
        A = U[: m * k].clone()
        B, _ = ops.contract(_bra(A), (m, k), [1, 2], phi, (m, n), [1, 3], [2, 3])
        ops.scal(B, 1.0 / ops.nrm2(B))
    else:
        rho, _ = ops.contract(pb, (m, n), [1, 2], phi, (m, n), [1, 3], [2, 3])
        if drho is not None:
            ops.axpby(noise, drho, 1.0, rho)
        w, U, k, terr, _ = ops.eigh_trunc(rho, n, maxdim=maxdim, cutoff=cut)
        V = U[: n * k].clone()
        B, _ = ops.permute(_bra(V), (n, k), (1, 0))
        A, _ = ops.contract(phi, (m, n), [1, 2], V, (n, k), [2, 3], [1, 3])
        ops.scal(A, 1.0 / ops.nrm2(A))
    return A, B, k, terr


def replacebond(psi, b, phi, pdims, ortho, maxdim, mindim, cutoff, drho=None, noise=0.0):
    """replacebond!(psi, b, phi; ortho, normalize=true) via factorize. With
    a density-matrix perturbation `drho` (scaled by `noise`) the eigen route
    is taken on rho + noise*drho, as factorize does for
    eigen_perturbation."""
    l, d1, d2, r = pdims
    m, n = l * d1, d2 * r
    use_eigen = (cutoff is not None and cutoff >= 1e-12) or drho is not None
    dt = str(phi.dtype).split(".")[-1]
    if phi.is_complex():
        A, B, k, terr = _replacebond_complex(phi, m, n, ortho, maxdim, mindim, cutoff, drho,
                                             noise, use_eigen)
    elif use_eigen:
        if ortho == "left":
            rho = ops.empty(m * m, dt)
            ops.gemm("N", "T", m, m, n, 1.0, phi, m, phi, m, 0.0, rho, m)
            if drho is not None:
                ops.axpby(noise, drho, 1.0, rho)
            w, U, k, terr, _ = ops.eigh_trunc(rho, m, maxdim=maxdim, cutoff=cutoff)
            # own copy of the kept columns: the n x n eigenvector buffer is freed
            A = U[: m * k].clone()
            B = ops.empty(k * n, dt)
            ops.gemm("T", "N", k, n, m, 1.0, U, m, phi, m, 0.0, B, k)
            ops.scal(B, 1.0 / ops.nrm2(B))
        else:
            rho = ops.empty(n * n, dt)
            ops.gemm("T", "N", n, n, m, 1.0, phi, m, phi, m, 0.0, rho, n)
            if drho is not None:
                ops.axpby(noise, drho, 1.0, rho)
            w, U, k, terr, _ = ops.eigh_trunc(rho, n, maxdim=maxdim, cutoff=cutoff)
            B, _ = ops.permute(U[: n * k], (n, k), (1, 0))
            A = ops.empty(m * k, dt)
            ops.gemm("N", "N", m, k, n, 1.0, phi, m, U, n, 0.0, A, m)
            ops.scal(A, 1.0 / ops.nrm2(A))
    else:
        work = ops.copy(phi)
        U, S, V, k, terr, _ = ops.svd_trunc(work, m, n, maxdim=maxdim, mindim=mindim,
                                            cutoff=0.0 if cutoff is None else cutoff)
        s = ops.to_host(S)[:k]
        Sd = ops.to_device(np.ravel(np.diag(s), order="F"), dt)
        if ortho == "left":
            A = U[: m * k].clone()
            B = ops.empty(k * n, dt)
            ops.gemm("N", "T", k, n, k, 1.0, Sd, k, V, n, 0.0, B, k)
            ops.scal(B, 1.0 / ops.nrm2(B))
        else:
            B, _ = ops.permute(V[: n * k], (n, k), (1, 0))
            A = ops.empty(m * k, dt)
            ops.gemm("N", "N", m, k, k, 1.0, U, m, Sd, k, 0.0, A, m)
            ops.scal(A, 1.0 / ops.nrm2(A))
    psi.data[b], psi.dims[b] = A, (l, d1, k)
    psi.data[b + 1], psi.dims[b + 1] = B, (k, d2, r)
    if ortho == "left":
        psi.llim, psi.rlim = b + 1, b + 3
    else:
        psi.llim, psi.rlim = b - 1, b + 1
    return terr, k


# --------------------------------------------------------------------- dmrg
class Sweeps:
    """Sweep schedule (Sweeps/maxdim!/mindim!/cutoff!/noise!, test/dmrg.jl:19-24);
    the last given value repeats."""

    def __init__(self, nsweep, maxdim=(), mindim=(1,), cutoff=(0.0,), noise=(0.0,)):
        self.nsweep = nsweep

        def fill(v):
            v = list(v) if v else [None]
            return [v[min(i, len(v) - 1)] for i in range(nsweep)]

        self.maxdim = fill(maxdim)
        self.mindim = fill(mindim)
        self.cutoff = fill(cutoff)
        self.noise = fill(noise)

    def as_dict(self):
        return {"nsweep": self.nsweep, "maxdim": self.maxdim, "mindim": self.mindim,
                "cutoff": self.cutoff, "noise": self.noise}

    @staticmethod
    def from_dict(d):
        return Sweeps(d["nsweep"], d["maxdim"], d["mindim"], d["cutoff"], d["noise"])

    def tail(self, k):
        """The schedule from sweep k (0-based) on."""
        return Sweeps(self.nsweep - k, self.maxdim[k:], self.mindim[k:], self.cutoff[k:],
                      self.noise[k:])


def sweepnext(N):
    for b in range(N - 1):
        yield b, 1
    for b in range(N - 2, -1, -1):
        yield b, 2


def dmrg(H, psi, sweeps, mode="native", krylovdim=3, precond=None, errgoal=1e-14,
         record=None, sweep_callback=None, env_mode=None, ph=None, timer=None, native=True,
         checkpoint=None, first_sweep=0, history=None):
    """Two-site DMRG; psi (right-orthogonal, centre at site 0) is updated in
    place. Returns the final energy. `precond`: None, "diag" or a callable
    factory (PH, b, xdims) -> precond(q, lam). Passing the ProjMPO `ph` of a
    previous call continues from its environments (a completed sweep leaves
    psi right-orthogonal and the right environments current). With
    `native` (default) the per-bond Davidson runs in libitg (itg_davidson2)
    unless a learned preconditioner is given. With `checkpoint` (a path) the
    MPS and the sweep state (sweeps done, energy, per-sweep records, schedule)
    are written after every sweep; dmrg_resume continues from such a file.
    `first_sweep` numbers the sweeps of a resumed schedule and `history` holds
    the records of the sweeps before it."""
    N = len(psi)
    history = list(history or [])
    # ComplexF64: operator and state on complex storage, FP64 Heff as
    # contractions, the host Davidson (complex inner products) and the
    # Hermitian density-matrix route
    if psi.is_complex() or H.is_complex():
        if mode != "native":
            raise _lib.ArgumentError("complex dmrg runs in FP64 (native) mode only")
        H = H.complex()
        if not psi.is_complex():
            psi.data = [ops.convert(t, "complex128") for t in psi.data]
        native = False
    # environments accumulate along the chain (L_b is b successive updates),
    # so their errors add up: they stay FP64 unless asked otherwise
    env_mode = "native" if env_mode is None else env_mode
    if ph is None:
        PH = ProjMPO(H, mode, timer, env_mode)
        PH.build_right(psi)
    else:
        PH = ph
        PH.mode, PH.env_mode, PH.timer = mode, env_mode, timer
    dmrg.last_projmpo = PH
    energy = 0.0
    for sw in range(sweeps.nsweep):
        t0 = time.perf_counter()
        maxerr, stats = 0.0, {}
        noise = float(sweeps.noise[sw] or 0.0)
        for b, ha in sweepnext(N):
            A, Ad = psi.data[b], psi.dims[b]
            B, Bd = psi.data[b + 1], psi.dims[b + 1]
            phi, pdims = ops.contract(A, Ad, [1, 2, 3], B, Bd, [3, 4, 5], [1, 2, 4, 5])
            PH.prepare(b, pdims)
            # a learned preconditioner that the device loop can run itself
            net = getattr(precond, "device_net", None) if callable(precond) else None
            if native and (not callable(precond) or net is not None):
                # device Davidson (itg_davidson2, or itg_davidson2_mlp with
                # the guarded learned step); phi is a fresh contraction
                # output, overwritten with the Ritz vector
                dg = PH.diag(b, pdims) if (precond == "diag" or net is not None) else None
                energy, phi = PH.eigsolve(b, phi, pdims, maxiter=krylovdim - 1,
                                          errgoal=errgoal, diag=dg, stats=stats, net=net)
            else:
                work = PH.heff_work((pdims[0], H.dims[b][0], H.dims[b][1], H.dims[b + 1][1],
                                     pdims[1], pdims[3]))

                def mv(x, b=b, pdims=pdims, work=work):
                    return PH.product(b, x, pdims, work=work)

                pre = None
                if precond == "diag":
                    pre = diag_preconditioner(PH.diag(b, pdims))
                elif callable(precond):
                    pre = precond(PH, b, pdims)
                energy, phi = davidson(mv, phi, maxiter=krylovdim - 1, errgoal=errgoal,
                                       precond=pre, stats=stats,
                                       fallback=getattr(pre, "fallback", None))
            ortho = "left" if ha == 1 else "right"
            ev = timer.start() if timer is not None else None
            kt = timer is not None and b == timer.factorize_bond and ha == 1 and \
                timer.factorize_kinds
            if kt:
                _lib.ctx().timing_set(tuple(timer.base_kinds) + timer.factorize_kinds)
            drho = noiseterm(PH, b, phi, pdims, ortho, mode) if noise > 0.0 else None
            terr, k = replacebond(psi, b, phi, pdims, ortho, sweeps.maxdim[sw],
                                  sweeps.mindim[sw], sweeps.cutoff[sw], drho, noise)
            if kt:
                _lib.ctx().timing_set(timer.base_kinds)
                timer.stop("factorize_timed_bond", timer.start_at(ev))
            if ev is not None:
                timer.stop("factorize", ev)
            maxerr = max(maxerr, terr)
            ev = timer.start() if timer is not None else None
            if ha == 1:
                L, Ld = PH.L[b]
                PH.L[b + 1] = update_left(L, Ld, psi.data[b], psi.dims[b], H.data[b], H.dims[b],
                                          env_mode)
            else:
                R, Rd = PH.R[b + 2]
                PH.R[b + 1] = update_right(R, Rd, psi.data[b + 1], psi.dims[b + 1],
                                           H.data[b + 1], H.dims[b + 1], env_mode)
            if ev is not None:
                timer.stop("environment", ev)
        lam = energy
        if mode != "native":
            # FP64 Rayleigh quotient of the current MPS at the first bond:
            # L[0] is the boundary and R[2] was rebuilt in FP64 from the
            # current tensors, so this is <psi|H|psi>/<psi|psi> exactly
            phi, pdims = ops.contract(psi.data[0], psi.dims[0], [1, 2, 3], psi.data[1],
                                      psi.dims[1], [3, 4, 5], [1, 2, 4, 5])
            saved = PH.mode
            PH.mode = "native"
            y = PH.product(0, phi, pdims)
            PH.mode = saved
            energy = ops.dot(phi, y) / ops.dot(phi, phi)
        _lib.ctx().sync()
        rec = {"sweep": first_sweep + sw + 1, "energy": energy, "energy_davidson": lam,
               "maxerr": maxerr, "maxlinkdim": psi.maxlinkdim(),
               "seconds": time.perf_counter() - t0,
               "matvecs": stats.get("matvecs", 0), "reverted": stats.get("reverted", 0)}
        if record is not None:
            record.append(rec)
        history.append(rec)
        if checkpoint is not None:
            from .checkpoint import save_mps
            full = Sweeps(first_sweep + sweeps.nsweep, [None] * first_sweep + sweeps.maxdim,
                          [None] * first_sweep + sweeps.mindim,
                          [None] * first_sweep + sweeps.cutoff,
                          [None] * first_sweep + sweeps.noise).as_dict()
            save_mps(checkpoint, psi, {"sweeps_done": first_sweep + sw + 1, "energy": energy,
                                       "mode": mode, "krylovdim": krylovdim,
                                       "schedule": full, "record": history})
        if sweep_callback is not None:
            sweep_callback(rec)
    return energy


def dmrg_resume(H, path, sweeps, **kwargs):
    """Continue a dmrg run from the checkpoint at `path` (written by
    dmrg(..., checkpoint=path) after a completed sweep): the MPS is read
    back, its environments rebuilt, and the sweeps of `sweeps` after those
    already done are run, checkpointing to the same path. A completed sweep
    leaves the MPS right-orthogonal with its centre at site 0, the state
    dmrg starts from, so the continued run repeats the uninterrupted one.
    Returns (energy, psi)."""
    from .checkpoint import load_mps
    psi, meta = load_mps(path)
    done = int(meta["sweeps_done"])
    if done >= sweeps.nsweep:
        return float(meta["energy"]), psi
    kwargs.setdefault("krylovdim", meta.get("krylovdim", 3))
    kwargs.setdefault("mode", meta.get("mode", "native"))
    E = dmrg(H, psi, sweeps.tail(done), checkpoint=path, first_sweep=done,
             history=meta.get("record"), **kwargs)
    return E, psi


def expect_energy(H, psi):
    """<psi|H|psi>/<psi|psi> by a full left-to-right transfer on the device
    (real part: H is Hermitian)."""
    cplx = psi.is_complex() or H.is_complex()
    if cplx:
        psi, H = psi.complex(), H.complex()
    dt = "complex128" if cplx else "float64"
    L, Ld = ones_env(dt)
    Nn, Nd = ones_env(dt)
    for j in range(len(psi)):
        L, Ld = update_left(L, Ld, psi.data[j], psi.dims[j], H.data[j], H.dims[j])
        # norm transfer: N(r', r) = N(l', l) A(l, s, r) conj(A(l', s, r'))
        T, td = ops.contract(Nn, Nd, [1, 2, 3], psi.data[j], psi.dims[j], [3, 4, 5], [1, 2, 4, 5])
        Nn, Nd = ops.contract(T, td, [1, 2, 4, 5], _bra(psi.data[j]), psi.dims[j], [1, 4, 6],
                              [6, 2, 5])
    return float(np.real(ops.to_host(L)[0] / ops.to_host(Nn)[0]))
