"""Learned Davidson preconditioner on the GPU: MLP forward/backward/Adam and
residual feature rows against the CPU oracle (oracle/mlp.py), training on
the device, the revert guard, and what the trained network gains over the
diagonal preconditioner on held-out effective Hamiltonians."""
import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def itg():
    import itg as m
    m._lib.ctx()
    return m


def _relerr(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("rows,i,h,o", [(300, 12, 256, 4), (37, 27, 64, 9), (1, 3, 16, 1)])
def test_mlp_forward_backward_vs_oracle(itg, rows, i, h, o):
    from itg import ops
    rng = np.random.default_rng(rows + i)
    p = (rng.standard_normal(omlp.param_count(i, h, o)) * 0.3).astype(np.float32)
    X = rng.standard_normal((rows, i)).astype(np.float32)
    T = rng.standard_normal((rows, o)).astype(np.float32)
    P, Xd, Td = (ops.to_device(a.ravel(), "float32") for a in (p, X, T))
    Y, acts = ops.mlp_forward(P, Xd, rows, i, h, o)
    G, loss = ops.mlp_backward(P, Xd, Td, acts, Y, rows, i, h, o)
    Yo, (Z1, A1, Z2, A2) = omlp.forward(p, X, i, h, o)
    lo, go = omlp.loss_and_grad(p, X, T, i, h, o)
    # FP32 MFMA products with FP32 accumulation over at most 300 terms
    assert _relerr(ops.to_host(Y).reshape(rows, o), Yo) < 2e-5
    a = ops.to_host(acts).reshape(4, rows, h)
    for got, want in zip(a, (Z1, A1, Z2, A2)):
        assert _relerr(got, want) < 2e-5
    assert abs(loss - lo) < 1e-5 * lo
    g = ops.to_host(G)
    # per block (weights and biases separately), relative to the block's scale
    sizes = [h * i, h, h * h, h, o * h, o]
    s = 0
    for n in sizes:
        assert _relerr(g[s:s + n], go[s:s + n]) < 2e-4
        s += n


def test_mlp_adam_vs_oracle(itg):
    from itg import ops
    rng = np.random.default_rng(1)
    n = 1000
    p, g = rng.standard_normal(n), rng.standard_normal(n)
    m, v = rng.standard_normal(n) * 0.1, rng.random(n) * 0.1
    P, G, M, V = (ops.to_device(a, "float32") for a in (p, g, m, v))
    ops.mlp_adam(P, G, M, V, 7, lr=3e-3)
    # the kernel's constants are FP32: give the FP64 oracle the same rounded
    # beta1/beta2 (1 - 0.999 differs by 1.3e-5 relative otherwise)
    f = lambda x: np.asarray(x, dtype=np.float32).astype(np.float64)
    po, mo, vo = omlp.adam(f(p), f(g), f(m), f(v), 7, lr=float(f(3e-3)), b1=float(f(0.9)),
                           b2=float(f(0.999)))
    np.testing.assert_allclose(ops.to_host(P), po, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(ops.to_host(M), mo, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(ops.to_host(V), vo, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("chiL,d,chiR,row0,step", [(5, 2, 7, 0, 1), (9, 3, 4, 2, 5),
                                                   (64, 2, 48, 3, 17)])
def test_features_gather_scatter_vs_oracle(itg, chiL, d, chiR, row0, step):
    from itg import ops
    rng = np.random.default_rng(chiL * chiR)
    n = chiL * d * d * chiR
    r, dg = rng.standard_normal(n), rng.standard_normal(n) + 0.5
    dg[3] = -0.25                              # exactly lambda: clipped to the floor
    lam = -0.25
    rd, dd = ops.to_device(r), ops.to_device(dg)
    X = ops.mlp_features(rd, dd, lam, chiL, d, chiR, 1e-8, row0, step)
    Xo = omlp.features(r, dg, lam, chiL, d, chiR, 1e-8, row0, step)
    np.testing.assert_allclose(ops.to_host(X).reshape(Xo.shape), Xo, rtol=2e-7, atol=1e-30)
    T = ops.mlp_gather(rd, 0.5, chiL, d, chiR, row0, step)
    np.testing.assert_array_equal(ops.to_host(T).reshape(-1, d * d),
                                  (omlp.gather(r, 0.5, chiL, d, chiR, row0, step))
                                  .astype(np.float32))
    full = ops.mlp_gather(rd, 1.0, chiL, d, chiR)
    back = ops.mlp_scatter(full, chiL, d, chiR, 0, chiL * chiR, ops.fill(ops.empty(n), 0.0))
    np.testing.assert_array_equal(ops.to_host(back), r.astype(np.float32).astype(np.float64))
    with pytest.raises(itg._lib.DimensionMismatch):
        ops.mlp_features(rd, dd, lam, chiL, d, chiR, 1e-8, 0, 1, chiL * chiR + 1)


@pytest.mark.parametrize("rows,h,tol", [(1000, 256, 1e-2), (64, 256, 1e-2), (37, 64, 2e-5)])
def test_mlp_infer_matches_forward(itg, rows, h, tol):
    """The inference path against the FP32 forward: h = 256 runs the fused
    kernel (hidden product in bf16, ~3e-3 relative), other widths the
    forward's FP32 products without saved activations."""
    from itg import ops
    rng = np.random.default_rng(rows)
    i, o = 12, 4
    n = ops.mlp_param_count(i, h, o)
    P = ops.to_device((rng.standard_normal(n) * 0.2).astype(np.float32), "float32")
    X = ops.to_device(rng.standard_normal(rows * i).astype(np.float32), "float32")
    Yf, _ = ops.mlp_forward(P, X, rows, i, h, o)
    Yi = ops.mlp_infer(P, X, rows, i, h, o)
    a, b = ops.to_host(Yi), ops.to_host(Yf)
    assert np.linalg.norm(a - b) <= tol * np.linalg.norm(b)


@pytest.mark.parametrize("rows,i,o", [(1000, 12, 4), (130, 27, 9), (64, 5, 1)])
def test_mlp_infer_fused_vs_oracle_emulation(itg, rows, i, o):
    """The fused h = 256 inference kernel against the oracle's restatement of
    its arithmetic (bf16-rounded layer-1 activations and W2, FP32 elsewhere;
    oracle.mlp.forward_bf16_hidden), in the Frobenius norm: at least ten times
    closer to it than to the exact forward. (Elementwise the two can differ
    by one bf16 step where a layer-1 activation lies within FP32 rounding of
    a bf16 rounding boundary: the kernel's FP32 GELU and the oracle's FP64
    one may round it to neighbouring bf16 values.)"""
    from itg import ops
    rng = np.random.default_rng(rows + i)
    h = 256
    p = (rng.standard_normal(omlp.param_count(i, h, o)) * 0.2).astype(np.float32)
    X = rng.standard_normal((rows, i)).astype(np.float32)
    Yi = ops.mlp_infer(ops.to_device(p, "float32"), ops.to_device(X.ravel(), "float32"),
                       rows, i, h, o)
    got = ops.to_host(Yi).reshape(rows, o)
    emu = omlp.forward_bf16_hidden(p, X, i, h, o)
    exact = omlp.forward(p, X, i, h, o)[0]
    fro = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    e_emu, e_exact = fro(got, emu), fro(got, exact)
    print("mlp_infer rows=%d in=%d out=%d: vs emulation %.2e, vs exact forward %.2e "
          "(max-norm %.2e, %.2e)" % (rows, i, o, e_emu, e_exact, _relerr(got, emu),
                                     _relerr(got, exact)))
    assert e_emu < 5e-5
    assert e_exact > 10 * e_emu


def test_rq_grad_finite_differences(itg):
    """The Ritz-value gradient of the learned step (MLPPreconditioner.rq_grad,
    which train_sweep descends) against central differences of the Ritz
    value itself along random parameter directions, on a Heisenberg bond's
    own Heff: the directional derivatives agree in sign and to ~1%."""
    from itg import ops
    from itg.precond import MLPPreconditioner
    (mv, phi, dg, pdims), = _chain_operators(itg, 10, 8, 4, {4})
    x = ops.copy(phi)
    ops.scal(x, 1.0 / ops.nrm2(x))
    Hx = ops.copy(mv(x))
    lam = ops.dot(x, Hx)
    q = ops.copy(Hx)
    ops.axpby(-lam, x, 1.0, q)
    rng = np.random.default_rng(5)
    net = MLPPreconditioner(2, hidden=32, seed=6)
    # a network away from its zero-initialised last layer
    net.set_params(rng.standard_normal(net.nparam).astype(np.float32) * 0.3)
    p0 = net.host_params().astype(np.float64)
    mu0, gscale = net.rq_grad(mv, x, lam, q, dg, pdims)
    assert gscale > 0 and mu0 < lam
    grad = ops.to_host(net.grad).astype(np.float64) * gscale
    for trial in range(4):
        delta = rng.standard_normal(p0.size)
        eps = 1e-3 * np.linalg.norm(p0) / np.linalg.norm(delta)
        pp = (p0 + eps * delta).astype(np.float32)
        pm = (p0 - eps * delta).astype(np.float32)
        net.set_params(pp)
        mup = net.rq_grad(mv, x, lam, q, dg, pdims, backward=False)[0]
        net.set_params(pm)
        mum = net.rq_grad(mv, x, lam, q, dg, pdims, backward=False)[0]
        fd = mup - mum
        an = float(grad @ (pp.astype(np.float64) - pm.astype(np.float64)))
        print("rq_grad trial %d: finite difference %.6e, gradient %.6e" % (trial, fd, an))
        assert fd * an > 0
        assert abs(fd - an) <= 2e-2 * abs(an)
    net.set_params(p0.astype(np.float32))


def test_mlp_training_lowers_loss(itg):
    from itg import ops
    from itg.precond import MLPPreconditioner
    rng = np.random.default_rng(2)
    net = MLPPreconditioner(2, hidden=128, seed=3)
    rows = 16384
    X = rng.standard_normal((rows, net.i)).astype(np.float32)
    # a smooth nonlinear target of the kind the correction map is
    T = (X[:, :4] / (1.0 + X[:, 4:8] ** 2) + 0.3 * X[:, 8:12]).astype(np.float32)
    Xd, Td = ops.to_device(X.ravel(), "float32"), ops.to_device(T.ravel(), "float32")
    l0 = net.loss(Xd, Td, rows)
    net.train(Xd, Td, rows, steps=400, batch=2048, lr=2e-3)
    l1 = net.loss(Xd, Td, rows)
    assert l1 < 0.1 * l0, (l0, l1)


# ------------------------------------------------ use inside Davidson
def _chain_operators(itg, N, chi, seed, bonds):
    """Heff operators of a Heisenberg chain at the listed bonds, taken during
    a left-to-right half sweep from a random chi MPS (states are the
    pre-update two-site tensors, as Davidson first sees them)."""
    from itg import mps, ops
    H = mps.MPO.from_opsum(mps.heisenberg_opsum([(j, j + 1) for j in range(N - 1)]), N)
    psi = mps.random_mps(N, 2, chi, seed)
    PH = mps.ProjMPO(H)
    PH.build_right(psi)
    out = []
    for b in range(N - 1):
        A, Ad, B, Bd = psi.data[b], psi.dims[b], psi.data[b + 1], psi.dims[b + 1]
        phi, pdims = ops.contract(A, Ad, [1, 2, 3], B, Bd, [3, 4, 5], [1, 2, 4, 5])

        def mv(x, b=b, pdims=pdims):
            return PH.product(b, x, pdims)

        dg = PH.diag(b, pdims)
        if b in bonds:
            out.append((mv, ops.copy(phi), dg, pdims))
        lam, phi = mps.davidson(mv, phi, maxiter=2, precond=mps.diag_preconditioner(dg))
        mps.replacebond(psi, b, phi, pdims, "left", chi, 1, 1e-12)
        L, Ld = PH.L[b]
        PH.L[b + 1] = mps.update_left(L, Ld, psi.data[b], psi.dims[b], H.data[b], H.dims[b])
    return out


def test_guard_reverts_a_useless_network(itg):
    """Zero weights (no residual skip) give a zero correction: every learned
    step is reverted and the result is exactly the diagonal-preconditioned
    Davidson."""
    from itg import mps, ops
    from itg.precond import MLPPreconditioner
    net = MLPPreconditioner(2, hidden=32, skip="none")
    net.set_params(np.zeros(net.nparam, dtype=np.float32))
    (mv, phi, dg, pdims), = _chain_operators(itg, 12, 16, 5, {5})
    dpre = mps.diag_preconditioner(dg)
    st_d, st_m = {}, {}
    lam_d, _ = mps.davidson(mv, phi, maxiter=3, precond=dpre, stats=st_d)
    lam_m, _ = mps.davidson(mv, phi, maxiter=3, stats=st_m, fallback=dpre,
                            precond=lambda q, lam: net.correction(q, lam, dg, pdims))
    assert st_m["reverted"] == 3
    assert st_m["matvecs"] == st_d["matvecs"]
    assert lam_m == lam_d


def test_guard_keeps_residual_decreasing_with_garbage_weights(itg):
    """Large random weights: every learned step the guard accepts lowers the
    residual. The calls show the steps: the learned correction is asked for
    at each iteration's residual q; a fallback call at the same q means the
    learned step was reverted, a learned call at the next iteration means it
    was kept, and then the new residual must be smaller. (The fallback step
    itself is not guarded, so the final residual is not bounded by the
    first one.)"""
    from itg import mps, ops
    from itg.precond import MLPPreconditioner
    net = MLPPreconditioner(2, hidden=32, seed=9)
    net.set_params(np.random.default_rng(9).standard_normal(net.nparam).astype(np.float32) * 30)
    (mv, phi, dg, pdims), = _chain_operators(itg, 12, 16, 6, {6})
    calls = []
    diag = mps.diag_preconditioner(dg)

    def learned(q, lam):
        calls.append(("learned", ops.nrm2(q)))
        return net.correction(q, lam, dg, pdims)

    def fallback(q, lam):
        calls.append(("fallback", ops.nrm2(q)))
        return diag(q, lam)

    st = {}
    mps.davidson(mv, phi, maxiter=6, stats=st, fallback=fallback, precond=learned)
    kept = 0
    for (k0, r0), (k1, r1) in zip(calls, calls[1:]):
        if k0 == "learned" and k1 == "learned":
            assert r1 < r0
            kept += 1
    if calls and calls[-1][0] == "learned":
        assert st["residual"] < calls[-1][1]
    # each iteration asks the learned step once; reverted ones also the fallback
    assert st["reverted"] == sum(1 for k, _ in calls if k == "fallback")
    assert st["reverted"] + kept + (1 if calls and calls[-1][0] == "learned" else 0) >= 1


def test_learned_preconditioner_gain(itg, capsys):
    """Heisenberg N=100: train the network by the one-step Ritz value over
    three sweeps of one random chi=64 chain, then run DMRG at maxdim 256 from
    another random start with the diagonal preconditioner and with the
    guarded network (diagonal fallback). Training raises the energy the
    learned step gains on the bonds it trains on; after two sweeps the
    learned form is far closer to the converged energy than the diagonal one
    (here 8e-8 against 4.5e-4; 6e-7 against 1.2e-4 and 4e-8 against 2.2e-5
    from other starts at maxdim 256 and 512); and with the guard both runs
    converge to the same energy. Both legs run the same (host-driven)
    Davidson: from a random start the first-sweep energy moves by ~0.06 with
    the floating-point order of the Davidson alone (the device Davidson gives
    -44.07 where the host one gives -44.01 with the same diagonal
    preconditioner), so only like-for-like runs are compared."""
    from itg.mps import MPO, Sweeps, dmrg, heisenberg_opsum, random_mps
    from itg.precond import MLPPreconditioner, train_sweep
    N, chi = 100, 256
    H = MPO.from_opsum(heisenberg_opsum([(j, j + 1) for j in range(N - 1)]), N)
    net = MLPPreconditioner(2, hidden=256, seed=1)
    gains = np.array(train_sweep(net, H, random_mps(N, 2, 64, seed=21), 64, steps_per_bond=10,
                                 sweeps=3, lr=3e-4))
    nb = len(gains) // 3
    first, last = gains[:nb, 0].sum(), gains[:nb, 1].sum()
    res = {}
    for name, pre in (("diag", "diag"), ("mlp", net.factory())):
        rec = []
        dmrg(H, random_mps(N, 2, chi, seed=22), Sweeps(3, maxdim=[chi], cutoff=[1e-12]),
             precond=pre, record=rec, native=False)
        res[name] = rec
    with capsys.disabled():
        print("\nlearned preconditioner: first-sweep step gain %.4f -> %.4f over %d bonds"
              % (first, last, nb))
        for k in res:
            print("  %-4s" % k, ["%.10f (rev %d)" % (r["energy"], r.get("reverted", 0))
                                 for r in res[k]])
    # the trained step gains more energy than the diagonal step it starts as
    assert last < 1.02 * first < 0
    e_conv = min(r["energy"] for k in res for r in res[k])
    err2 = {k: res[k][1]["energy"] - e_conv for k in res}
    assert err2["mlp"] < 0.1 * err2["diag"]
    # and, with the guard, the same converged energy to 1e-6 per site
    assert abs(res["mlp"][-1]["energy"] - res["diag"][-1]["energy"]) < 1e-6 * N


def test_learned_step_against_diag_per_bond(itg, capsys):
    """Single Davidson runs (2 steps) on held-out bond operators (the first
    half sweep from a random chi=64 state, far from convergence): energy
    error with no, diagonal and learned (guarded) preconditioners. Each run
    starts from the same vector on the same operator, so the comparison is
    deterministic. The untrained network is exactly the diagonal step; after
    training by the one-step Ritz value on another chain, summed over eight
    bonds it ends below the diagonal form it falls back to (0.84-0.90 of it
    in the runs so far), and well below it on some bonds."""
    from itg.mps import MPO, heisenberg_opsum, random_mps
    from itg.precond import MLPPreconditioner, evaluate, train_sweep
    held_out = _chain_operators(itg, 24, 64, 12, set(range(4, 20, 2)))
    net = MLPPreconditioner(2, hidden=256, seed=1)
    mv, phi, dg, pdims = held_out[0]
    res, _ = evaluate(mv, phi, dg, pdims, net, maxiter=2)
    # (a diagonal step that does not lower the residual is "reverted" to the
    # same diagonal step, so only the Ritz value is compared)
    assert abs(res["mlp"][0] - res["diag"][0]) < 1e-10 * abs(res["diag"][0])
    H = MPO.from_opsum(heisenberg_opsum([(j, j + 1) for j in range(23)]), 24)
    train_sweep(net, H, random_mps(24, 2, 48, seed=11), 48, steps_per_bond=10, sweeps=3, lr=3e-4)
    lines, tot, ratios = [], {"plain": 0.0, "diag": 0.0, "mlp": 0.0}, []
    for mv, phi, dg, pdims in held_out:
        res, lam0 = evaluate(mv, phi, dg, pdims, net, maxiter=2)
        e = {k: v[0] - lam0 for k, v in res.items()}
        assert all(np.isfinite(v) and v > -1e-9 for v in e.values())
        for k in tot:
            tot[k] += e[k]
        ratios.append(e["mlp"] / e["diag"])
        lines.append("%s plain %.3e diag %.3e mlp %.3e (reverted %d)" % (
            pdims, e["plain"], e["diag"], e["mlp"], res["mlp"][2]))
    with capsys.disabled():
        for ln in lines:
            print("  dE after 2 steps:", ln)
        print("  summed dE: plain %.3e diag %.3e mlp %.3e" % (tot["plain"], tot["diag"],
                                                             tot["mlp"]))
    # the gain varies by bond (the guard hands some bonds back to the
    # diagonal step entirely): summed over eight bonds, and on the best one
    assert tot["mlp"] < 0.97 * tot["diag"]
    assert min(ratios) < 0.9
