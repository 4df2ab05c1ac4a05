"""Learned Davidson preconditioner: a 3-layer GELU MLP mapping local blocks
of the residual to a correction, trained and applied by libitg kernels.

No counterpart in the reference (ITensors' `davidson` has no
preconditioner). The network sees a two-site residual r (chiL, d, d, chiR)
as chiL*chiR rows, one per bond pair (a, b), each with 3*d*d features of its
d*d physical block: r/rms(r), delta/rms(delta) and (r/delta)/rms(r/delta),
delta = diag(Heff) - lambda (itg_mlp_features). It predicts the d*d block of
the normalised exact correction t* = psi0 - (x.psi0) x, where psi0 is the
ground state of the same Heff and x the current Davidson vector
(harvest_pairs). All features are scale-free and local, so one network
serves every chi. Inside Davidson the learned step is guarded: if it does
not lower the residual norm it is reverted and taken with the diagonal
preconditioner (mps.davidson's `fallback`).
"""
import numpy as np

from . import _lib
from . import ops
from .mps import davidson, diag_preconditioner


class MLPPreconditioner:
    """Parameters, Adam state and the inference path of one network."""

    def __init__(self, d, hidden=256, seed=0, floor=1e-8, chunk_rows=1 << 20):
        self.d = int(d)
        self.i, self.h, self.o = 3 * self.d * self.d, int(hidden), self.d * self.d
        self.floor = float(floor)
        self.chunk_rows = int(chunk_rows)
        self.nparam = ops.mlp_param_count(self.i, self.h, self.o)
        rng = np.random.default_rng(seed)
        parts = []
        for fan_in, fan_out in ((self.i, self.h), (self.h, self.h), (self.h, self.o)):
            parts.append(rng.standard_normal(fan_in * fan_out) / np.sqrt(fan_in))
            parts.append(np.zeros(fan_out))
        host = np.concatenate(parts).astype(np.float32)
        assert host.size == self.nparam
        self.params = ops.to_device(host, "float32")
        self.m = ops.fill(ops.empty(self.nparam, "float32"), 0.0)
        self.v = ops.fill(ops.empty(self.nparam, "float32"), 0.0)
        self.grad = ops.empty(self.nparam, "float32")
        self.step = 0
        self._bufs = {}

    # ------------------------------------------------------------ storage
    def host_params(self):
        return ops.to_host(self.params)

    def set_params(self, host):
        host = np.asarray(host, dtype=np.float32).ravel()
        if host.size != self.nparam:
            raise _lib.DimensionMismatch("expected %d parameters, got %d" % (self.nparam, host.size))
        self.params = ops.to_device(host, "float32")
        return self

    def save(self, path):
        np.savez(path, params=self.host_params(), d=self.d, hidden=self.h)

    @staticmethod
    def load(path, **kw):
        z = np.load(path, allow_pickle=False)
        net = MLPPreconditioner(int(z["d"]), int(z["hidden"]), **kw)
        return net.set_params(z["params"])

    def _buf(self, name, n, dtype="float32"):
        b = self._bufs.get(name)
        if b is None or b.numel() < n:
            b = ops.empty(n, dtype)
            self._bufs[name] = b
        return b[:n]

    # ----------------------------------------------------------- training
    def loss(self, X, T, rows):
        Y, acts = ops.mlp_forward(self.params, X, rows, self.i, self.h, self.o,
                                  Y=self._buf("Y", rows * self.o),
                                  acts=self._buf("acts", 4 * rows * self.h))
        _, loss = ops.mlp_backward(self.params, X, T, acts, Y, rows, self.i, self.h, self.o,
                                   G=self.grad)
        return loss

    def train(self, X, T, rows, steps=500, batch=4096, lr=1e-3, seed=0, log=None):
        """Adam on mini-batches of contiguous rows at random offsets of the
        (already interleaved) training set; returns the per-step losses."""
        rng = np.random.default_rng(seed)
        batch = min(int(batch), int(rows))
        Yb = self._buf("Y", batch * self.o)
        Ab = self._buf("acts", 4 * batch * self.h)
        losses = []
        for it in range(int(steps)):
            r0 = int(rng.integers(0, rows - batch + 1))
            Xb = X[r0 * self.i:(r0 + batch) * self.i]
            Tb = T[r0 * self.o:(r0 + batch) * self.o]
            ops.mlp_forward(self.params, Xb, batch, self.i, self.h, self.o, Y=Yb, acts=Ab)
            _, loss = ops.mlp_backward(self.params, Xb, Tb, Ab, Yb, batch, self.i, self.h, self.o,
                                       G=self.grad)
            self.step += 1
            ops.mlp_adam(self.params, self.grad, self.m, self.v, self.step, lr=lr)
            losses.append(loss)
            if log is not None and (it % 100 == 0 or it == steps - 1):
                log(it, loss)
        return losses

    # ---------------------------------------------------------- inference
    def correction(self, q, lam, diag, xdims):
        """t = network(features of q) scattered back to (chiL, d, d, chiR)."""
        chiL, d, d2, chiR = xdims
        if d != self.d or d2 != self.d:
            raise _lib.DimensionMismatch("network built for d=%d, got %s" % (self.d, xdims))
        rows = chiL * chiR
        t = ops.empty(q.numel(), "float64")
        for r0 in range(0, rows, self.chunk_rows):
            n = min(self.chunk_rows, rows - r0)
            X = ops.mlp_features(q, diag, lam, chiL, d, chiR, self.floor, r0, 1, n,
                                 X=self._buf("X", n * self.i))
            Y, _ = ops.mlp_forward(self.params, X, n, self.i, self.h, self.o,
                                   Y=self._buf("Y", n * self.o),
                                   acts=self._buf("acts", 4 * n * self.h))
            ops.mlp_scatter(Y, chiL, d, chiR, r0, n, t)
        return t

    def factory(self, guard=True, min_residual=0.0):
        """dmrg `precond` factory: (ProjMPO, b, xdims) -> precond(q, lam)
        carrying the diagonal preconditioner as its revert fallback. Residuals
        with norm below `min_residual` take the diagonal step directly (the
        network is trained on states far from convergence)."""

        def make(PH, b, xdims):
            dg = PH.diag(b, xdims)
            dpre = diag_preconditioner(dg)

            def pre(q, lam):
                if min_residual > 0.0 and ops.nrm2(q) < min_residual:
                    return dpre(q, lam)
                return self.correction(q, lam, dg, xdims)

            if guard:
                pre.fallback = dpre
            return pre

        return make


# ------------------------------------------------------------- harvesting
def exact_correction(matvec, x, dg, iters=40, errgoal=1e-10):
    """(lambda0, psi0, t*) for the operator `matvec` around the unit vector
    x: psi0 from a long diagonal-preconditioned Davidson, sign-fixed so that
    x.psi0 > 0, and t* = psi0 - (x.psi0) x."""
    lam0, psi0 = davidson(matvec, x, maxiter=iters, errgoal=errgoal,
                          precond=diag_preconditioner(dg))
    c = ops.dot(x, psi0)
    if c < 0:
        ops.scal(psi0, -1.0)
        c = -c
    t = ops.copy(psi0)
    ops.axpby(-c, x, 1.0, t)
    return lam0, psi0, t


def harvest_pairs(matvec, phi, dg, xdims, rows_per_bond, seed=0, floor=1e-8, iters=40):
    """Training rows from one Heff: features of the residual at x = phi/|phi|
    and the matching blocks of t*/rms(t*). Rows are a strided subsample
    (row0 + k*step) of the chiL*chiR bond pairs. Returns (X, T, nrows, info)."""
    chiL, d, _, chiR = xdims
    x = ops.copy(phi)
    ops.scal(x, 1.0 / ops.nrm2(x))
    lam0, _, t = exact_correction(matvec, x, dg, iters=iters)
    Ax = matvec(x)
    lam = ops.dot(x, Ax)
    r = ops.copy(Ax)
    ops.axpby(-lam, x, 1.0, r)
    rows = chiL * chiR
    step = max(1, rows // max(1, int(rows_per_bond)))
    row0 = int(np.random.default_rng(seed).integers(0, step))
    n = (rows - row0 + step - 1) // step
    X = ops.mlp_features(r, dg, lam, chiL, d, chiR, floor, row0, step, n)
    tn = ops.nrm2(t)
    scale = np.sqrt(t.numel()) / tn if tn > 0 else 1.0
    T = ops.mlp_gather(t, scale, chiL, d, chiR, row0, step, n)
    return X, T, n, {"lambda": lam, "lambda0": lam0, "residual": ops.nrm2(r)}


def concat(chunks, dtype="float32"):
    """Concatenate device buffers into one (stream-ordered copies)."""
    total = sum(int(c.numel()) for c in chunks)
    out = ops.empty(total, dtype)
    off = 0
    c = _lib.ctx()
    for ch in chunks:
        n = int(ch.numel())
        if n:
            c.check(_lib.lib().itg_memcpy_d2d(c.ptr, _lib.ptr(out[off:off + n]), _lib.ptr(ch),
                                              n * ch.element_size()))
        off += n
    return out


def evaluate(matvec, phi, dg, xdims, net, maxiter=2, lam0=None):
    """Davidson from the same start with no, diagonal and learned (guarded)
    preconditioners. Returns {name: (lambda, residual, reverted)} and lam0."""
    x = ops.copy(phi)
    ops.scal(x, 1.0 / ops.nrm2(x))
    if lam0 is None:
        lam0, _, _ = exact_correction(matvec, x, dg)
    dpre = diag_preconditioner(dg)
    out = {}
    for name, pre, fb in (("plain", None, None), ("diag", dpre, None),
                          ("mlp", (lambda q, lam: net.correction(q, lam, dg, xdims)), dpre)):
        st = {}
        lam, _ = davidson(matvec, x, maxiter=maxiter, precond=pre, fallback=fb, stats=st)
        out[name] = (lam, st["residual"], st.get("reverted", 0))
    return out, lam0


def harvest_sweep(H, psi, maxdim, rows_per_bond=4096, bonds=None, cutoff=1e-12, seed=0,
                  iters=40, sweeps=1):
    """Training rows from the chain's own effective Hamiltonians: `sweeps`
    full two-site sweeps (diagonal-preconditioned Davidson, 2 steps,
    replacebond at maxdim) that harvest pairs at the listed bonds of every
    left-to-right pass before each update, so the rows cover the states of a
    run from its random start towards convergence. psi (right-canonical,
    centre at site 0) is left in the same form. Returns (X, T, rows, d)."""
    from . import mps
    N = len(psi)
    PH = mps.ProjMPO(H)
    PH.build_right(psi)
    bonds = set(range(1, N - 2)) if bonds is None else set(bonds)
    Xs, Ts, rows, d = [], [], 0, None
    for sw in range(sweeps):
        for b, ha in mps.sweepnext(N):
            A, Ad, B, Bd = psi.data[b], psi.dims[b], psi.data[b + 1], psi.dims[b + 1]
            phi, pdims = ops.contract(A, Ad, [1, 2, 3], B, Bd, [3, 4, 5], [1, 2, 4, 5])
            d = pdims[1]
            PH.prepare(b, pdims)

            def mv(x, b=b, pdims=pdims):
                return PH.product(b, x, pdims)

            dg = PH.diag(b, pdims)
            if ha == 1 and b in bonds:
                X, T, n, _ = harvest_pairs(mv, phi, dg, pdims, rows_per_bond,
                                           seed=seed + 1000 * sw + b, iters=iters)
                Xs.append(X)
                Ts.append(T)
                rows += n
            lam, phi = mps.davidson(mv, phi, maxiter=2, precond=mps.diag_preconditioner(dg))
            ortho = "left" if ha == 1 else "right"
            mps.replacebond(psi, b, phi, pdims, ortho, maxdim, 1, cutoff)
            if ha == 1:
                L, Ld = PH.L[b]
                PH.L[b + 1] = mps.update_left(L, Ld, psi.data[b], psi.dims[b], H.data[b],
                                              H.dims[b])
            else:
                R, Rd = PH.R[b + 2]
                PH.R[b + 1] = mps.update_right(R, Rd, psi.data[b + 1], psi.dims[b + 1],
                                               H.data[b + 1], H.dims[b + 1])
    return concat(Xs), concat(Ts), rows, d
