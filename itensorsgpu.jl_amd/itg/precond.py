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
serves every chi.

The network's output is added to a fixed skip step, by default the
diagonal correction (r/delta)/rms(r/delta) ("diag"; "plain" uses r/rms(r),
"none" nothing), and its last layer starts at zero: the untrained network
is exactly the skip step. Two objectives train it:
- train_sweep / rq_step: the Ritz value of the one-step Davidson subspace
  span{x, t}, minimised directly over the chain's own Heff operators (its
  gradient in t is 2 (b/|t_perp|) (H u - mu u), u = the new Ritz vector);
- harvest_sweep + train: mean-squared error against the blocks of t*
  (less the skip step).
Inside Davidson the learned step is guarded: if it does not lower the
residual norm it is reverted and taken with the diagonal preconditioner
(mps.davidson's `fallback`).
"""
import numpy as np

from . import _lib
from . import ops
from .mps import davidson, diag_preconditioner


class MLPPreconditioner:
    """Parameters, Adam state and the inference path of one network."""

    SKIPS = ("none", "plain", "diag")

    def __init__(self, d, hidden=256, seed=0, floor=1e-8, chunk_rows=1 << 22, skip="diag"):
        if skip not in self.SKIPS:
            raise ValueError("skip must be one of %s" % (self.SKIPS,))
        self.d = int(d)
        self.skip = skip
        self.i, self.h, self.o = 3 * self.d * self.d, int(hidden), self.d * self.d
        self.floor = float(floor)
        self.chunk_rows = int(chunk_rows)
        self.nparam = ops.mlp_param_count(self.i, self.h, self.o)
        rng = np.random.default_rng(seed)
        parts = []
        for fan_in, fan_out in ((self.i, self.h), (self.h, self.h), (self.h, self.o)):
            w = rng.standard_normal(fan_in * fan_out) / np.sqrt(fan_in)
            if fan_out == self.o and skip != "none":
                w[:] = 0.0
            parts.append(w)
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
        np.savez(path, params=self.host_params(), d=self.d, hidden=self.h, skip=self.skip)

    @staticmethod
    def load(path, **kw):
        z = np.load(path, allow_pickle=False)
        kw.setdefault("skip", str(z["skip"]) if "skip" in z else "none")
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
        """t = network(features of q) scattered back to (chiL, d, d, chiR),
        plus the skip step."""
        chiL, d, d2, chiR = xdims
        if d != self.d or d2 != self.d:
            raise _lib.DimensionMismatch("network built for d=%d, got %s" % (self.d, xdims))
        rows = chiL * chiR
        t = ops.empty(q.numel(), "float64")
        for r0 in range(0, rows, self.chunk_rows):
            n = min(self.chunk_rows, rows - r0)
            X = ops.mlp_features(q, diag, lam, chiL, d, chiR, self.floor, r0, 1, n,
                                 X=self._buf("X", n * self.i))
            Y = ops.mlp_infer(self.params, X, n, self.i, self.h, self.o,
                              Y=self._buf("Y", n * self.o),
                              work=self._buf("work", 2 * n * self.h))
            ops.mlp_scatter(Y, chiL, d, chiR, r0, n, t)
        sv, sc = self.skip_step(q, lam, diag)
        if sv is not None:
            ops.axpby(sc, sv, 1.0, t)
        return t

    def skip_step(self, q, lam, diag):
        """(vector, scale) of the skip step: scale * vector has rms 1."""
        if self.skip == "none":
            return None, 0.0
        sv = q if self.skip == "plain" else ops.precond_diag(q, diag, lam, self.floor)
        sn = ops.nrm2(sv)
        return sv, (np.sqrt(sv.numel()) / sn if sn > 0 else 0.0)

    def rq_grad(self, matvec, x, lam, q, diag, xdims, X=None, backward=True):
        """Ritz value mu of span{x, t}, t = correction(q), for a unit vector
        x with lam = x.Hx and q = Hx - lam x; with `backward`, also the
        parameter gradient into self.grad. Returns (mu, gscale): self.grad
        is dmu/dparams / gscale (gscale > 0; 0 when the step is degenerate
        and self.grad is not written). X: the feature rows of q, if at hand.

        With u = ax x + c t the unit Ritz vector (coefficients stationary),
        dmu/dt = 2c (Hu - mu u); gathered into Y's layout it is dmu/dY. The
        backward pass differentiates mean((Y - T)^2), whose gradient in Y is
        2 (Y - T) / (rows out): with T = Y - G/|G| it is G / (|G| rows out / 2),
        so self.grad = dmu/dparams / (|G| rows out / 2)."""
        chiL, d, _, chiR = xdims
        rows = chiL * chiR
        if X is None:
            X = ops.mlp_features(q, diag, lam, chiL, d, chiR, self.floor, 0, 1, rows,
                                 X=self._buf("Xrq", rows * self.i))
        Y, acts = ops.mlp_forward(self.params, X, rows, self.i, self.h, self.o,
                                  Y=self._buf("Y", rows * self.o),
                                  acts=self._buf("acts", 4 * rows * self.h))
        t = self._buf("t", q.numel(), "float64")
        ops.mlp_scatter(Y, chiL, d, chiR, 0, rows, t)
        sv, sc = self.skip_step(q, lam, diag)
        if sv is not None:
            ops.axpby(sc, sv, 1.0, t)
        Ht = matvec(t)
        xt, tt, qt, tHt = ops.dot(x, t), ops.dot(t, t), ops.dot(q, t), ops.dot(t, Ht)
        n2 = tt - xt * xt
        if not n2 > 1e-24 * max(tt, 1e-300):
            return lam, 0.0
        tn = np.sqrt(n2)
        # the 2x2 projected problem on {x, e}, e = (t - (x.t) x)/|t_perp|
        xHt = qt + lam * xt
        ehe = (tHt - 2.0 * xt * xHt + xt * xt * lam) / n2
        xhe = qt / tn
        ev, U = np.linalg.eigh(np.array([[lam, xhe], [xhe, ehe]]))
        mu, a, b = float(ev[0]), float(U[0, 0]), float(U[1, 0])
        if not backward:
            return mu, 0.0
        c = b / tn
        ax = a - c * xt
        # g = dmu/dt = 2c (Hu - mu u), Hu = ax (q + lam x) + c Ht: the factor
        # 2c carries the sign of the eigenvector eigh happened to return
        g = self._buf("g", q.numel(), "float64")
        f = 2.0 * c
        ops.axpby(f * c, Ht, 0.0, g)
        ops.axpby(f * ax, q, 1.0, g)
        ops.axpby(f * ax * (lam - mu), x, 1.0, g)
        ops.axpby(-f * mu * c, t, 1.0, g)
        G = ops.mlp_gather(g, 1.0, chiL, d, chiR, 0, 1, rows, T=self._buf("G", rows * self.o))
        gn = ops.nrm2(G)
        if not gn > 0:
            return mu, 0.0
        Tt = self._buf("Trq", rows * self.o)
        ops.axpby(1.0, Y, 0.0, Tt)
        ops.axpby(-1.0 / gn, G, 1.0, Tt)
        ops.mlp_backward(self.params, X, Tt, acts, Y, rows, self.i, self.h, self.o, G=self.grad)
        return mu, gn * rows * self.o / 2.0

    def rq_step(self, matvec, x, lam, q, diag, xdims, X=None, lr=1e-3):
        """One Adam step down the Ritz value of span{x, t} (rq_grad; the
        gradient is normalised per step). Returns mu - lam (the energy the
        step gains, <= 0)."""
        mu, gscale = self.rq_grad(matvec, x, lam, q, diag, xdims, X=X)
        if gscale > 0:
            self.step += 1
            ops.mlp_adam(self.params, self.grad, self.m, self.v, self.step, lr=lr)
        return mu - lam

    # Davidson steps per bond that try the network on the device path (0:
    # every step); the first step carries most of the gain
    net_steps = 0

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

        # the device Davidson runs the same guarded step itself
        # (itg_davidson2_mlp) when nothing else is asked of the factory
        if guard and min_residual <= 0.0:
            make.device_net = self
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


def harvest_pairs(matvec, phi, dg, xdims, rows_per_bond, seed=0, floor=1e-8, iters=40,
                  skip="diag"):
    """Training rows from one Heff: features of the residual at x = phi/|phi|
    and the matching blocks of t*/rms(t*), less the skip step (the target of
    a network built with that skip). Rows are a strided subsample
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
    if skip != "none":
        sv = r if skip == "plain" else ops.precond_diag(r, dg, lam, floor)
        sn = ops.nrm2(sv)
        if sn > 0:
            ops.axpby(-1.0, ops.mlp_gather(sv, np.sqrt(sv.numel()) / sn, chiL, d, chiR, row0,
                                           step, n), 1.0, T)
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
                  iters=40, sweeps=1, skip="diag"):
    """Training rows from the chain's own effective Hamiltonians: `sweeps`
    full two-site sweeps (diagonal-preconditioned Davidson, 2 steps,
    replacebond at maxdim) that harvest pairs at the listed bonds of every
    left-to-right pass before each update, so the rows cover the states of a
    run from its random start towards convergence. psi (right-canonical,
    centre at site 0) is left in the same form. `skip` as in harvest_pairs.
    Returns (X, T, rows, d)."""
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
                                           seed=seed + 1000 * sw + b, iters=iters,
                                           skip=skip)
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


def train_sweep(net, H, psi, maxdim, steps_per_bond=10, cutoff=1e-12, sweeps=3, lr=1e-3,
                both=True, log=None):
    """Train `net` on the chain's own effective Hamiltonians by the one-step
    Ritz value (MLPPreconditioner.rq_step): `sweeps` full two-site sweeps
    (diagonal-preconditioned Davidson, 2 steps, replacebond at maxdim) that
    take `steps_per_bond` Adam steps at every bond of the left-to-right
    passes (and of the right-to-left ones with `both`) before its update,
    so the network sees states from the random start towards convergence.
    psi (right-canonical, centre at site 0) is left in the same form.
    Returns the per-bond (first, last) energy gains of the learned step."""
    from . import mps
    N = len(psi)
    PH = mps.ProjMPO(H)
    PH.build_right(psi)
    gains = []
    for sw in range(sweeps):
        for b, ha in mps.sweepnext(N):
            A, Ad, B, Bd = psi.data[b], psi.dims[b], psi.data[b + 1], psi.dims[b + 1]
            phi, pdims = ops.contract(A, Ad, [1, 2, 3], B, Bd, [3, 4, 5], [1, 2, 4, 5])
            PH.prepare(b, pdims)

            def mv(x, b=b, pdims=pdims):
                return PH.product(b, x, pdims)

            dg = PH.diag(b, pdims)
            if (ha == 1 or both) and pdims[1] == net.d and pdims[2] == net.d:
                x = ops.copy(phi)
                ops.scal(x, 1.0 / ops.nrm2(x))
                q = mv(x)
                lam = ops.dot(x, q)
                ops.axpby(-lam, x, 1.0, q)
                rows = pdims[0] * pdims[3]
                X = ops.mlp_features(q, dg, lam, pdims[0], net.d, pdims[3], net.floor, 0, 1, rows)
                first = last = 0.0
                for k in range(int(steps_per_bond)):
                    last = net.rq_step(mv, x, lam, q, dg, pdims, X=X, lr=lr)
                    if k == 0:
                        first = last
                gains.append((first, last))
                if log is not None:
                    log(sw, b, ha, first, last)
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
    return gains
