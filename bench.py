#!/usr/bin/env python3
"""Two-site DMRG sweep throughput on one MI355X.

Workload (the headline configuration): S=1/2 Heisenberg chain N=100, J=1,
MPO bond dimension 5, two-site Davidson with 3 Krylov vectors, sweeps at
maxdim 2048. A step is one full sweep (2(N-1) bond updates). Default start:
a random right-canonical MPS with link dims min(2^j, 2^(N-j), 2048) (seed 1)
and cutoff 1e-12: random states have flat spectra, so every bulk bond update
runs at the full chi=2048 shapes (Heff on 2048x2x2x2048, eigensolver
n=4096). --config3 instead starts from link dims 10 with the maxdim ramp
64..1024. Both forms run from the same start: contractions in split-BF16
(3 products, FP32 accumulation; the headline `value`) and in FP64; Davidson
vectors, environments, density matrices, eigensolver and energies are FP64
in both.

The whole run keeps to a wall budget (--budget seconds from process start):
each form does min(--warmup, 1) untimed sweeps and then up to --steps timed
sweeps, stopping early when its share of the budget is spent; the line's
`steps` and `warmup` are the numbers actually run. After the two forms come,
while time is left: the converged precision check (|dE|/N at maxdim 1024),
the CPU baseline (the oracle's BLAS form on one bulk bond at the same
shapes), the learned-preconditioner form and one eigendecomposition with its
kernels timed.

  python bench.py [--steps K] [--warmup W] [--N 100] [--chi 2048] [--config3]
  python bench.py --micro      # Heff apply of the chi=1024 microbench only
"""
import argparse
import json
import math
import os
import sys
import time

T_START = time.perf_counter()
ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "itensorsgpu.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "s/sweep + MFMA util, Heisenberg N=100 χ=2048, mixed vs FP64, |ΔE|/N≤1e-9"
PEAK = {"bf16x3": 2516.6, "fp32": 157.3, "native": 78.6}   # dense TFLOP/s, MI355X
HBM_PEAK = 8000.0                                             # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ramp_for(chi):
    return [m for m in (64, 128, 256, 512, 1024) if m < chi]


class Budget:
    """Wall-clock budget of the whole process, counted from its start."""

    def __init__(self, seconds):
        self.total = float(seconds)

    def used(self):
        return time.perf_counter() - T_START

    def left(self):
        return self.total - self.used()


class OutOfTime(Exception):
    pass


def run_mode(mode, N, chi, steps, warmup, krylov, budget, share, config3=False,
             timed_kinds=(), factorize_kinds=()):
    """`warmup` untimed sweeps, then up to `steps` timed sweeps at maxdim chi,
    stopping early (after at least one) when the next sweep would end past
    `share` seconds of timed work or past the process budget.

    Default workload: every timed sweep starts from a fresh random
    right-canonical MPS with link dims min(2^j, 2^(N-j), chi) (seeds 1, 2, ...,
    the same in every form), whose flat spectra keep every bulk bond at the
    full chi x chi shapes under cutoff 1e-12; the new state and its right
    environments are made before the sweep's clock starts. (Sweeping one
    state on converges it: at cutoff 1e-12 the chain's ground state needs
    chi ~ 200, so later sweeps would time a far smaller workload.)
    config3: random MPS with link dims 10 and the ramp 64..1024, then the
    timed sweeps continue that one state at maxdim chi.
    The timed region of each sweep is bracketed by device synchronisation."""
    from itg import _lib
    from itg.mps import MPO, PhaseTimer, ProjMPO, Sweeps, dmrg, heisenberg_opsum, random_mps
    import torch
    H = MPO.from_opsum(heisenberg_opsum([(j, j + 1) for j in range(N - 1)]), N)
    t0 = time.perf_counter()
    cutoff = 1e-12
    ph = None
    if config3:
        psi = random_mps(N, 2, 10, seed=1)
        ramp = ramp_for(chi)
        dmrg(H, psi, Sweeps(len(ramp), maxdim=ramp, cutoff=[cutoff]), mode=mode,
             krylovdim=krylov, sweep_callback=lambda r: log("[%s] ramp" % mode, r))
        ph = dmrg.last_projmpo
    wrec = []
    for w in range(warmup):
        if not config3:
            psi = random_mps(N, 2, chi, seed=1000 + w)
            if ph is not None:
                ph.build_right(psi)
        dmrg(H, psi, Sweeps(1, maxdim=[chi], cutoff=[cutoff]), mode=mode, krylovdim=krylov,
             ph=ph, record=wrec, sweep_callback=lambda r: log("[%s] warmup" % mode, r))
        ph = dmrg.last_projmpo
    if ph is None:
        ph = ProjMPO(H, mode)
    setup_s = time.perf_counter() - t0
    log("[%s] set-up done (%.1f s)" % (mode, setup_s))
    # the middle bond (full chi x chi shape) has its eigensolver kernels
    # event-timed inside the first timed sweep
    timer = PhaseTimer(factorize_bond=N // 2 if factorize_kinds else -1,
                       factorize_kinds=factorize_kinds)
    timer.base_kinds = tuple(timed_kinds)
    rec = []
    torch.cuda.synchronize()
    _lib.ctx().sync()
    _lib.ctx().timing(timed_kinds)
    timed = 0.0
    fresh_s = 0.0
    done = 0
    while done < steps:
        if not config3:
            tf = time.perf_counter()
            _lib.ctx().timing_set(())
            psi = random_mps(N, 2, chi, seed=1 + done)
            ph.build_right(psi)
            _lib.ctx().timing_set(timed_kinds)
            torch.cuda.synchronize()
            _lib.ctx().sync()
            fresh_s += time.perf_counter() - tf
        t1 = time.perf_counter()
        dmrg(H, psi, Sweeps(1, maxdim=[chi], cutoff=[cutoff]), mode=mode, krylovdim=krylov,
             ph=ph, timer=timer, record=rec,
             sweep_callback=lambda r: log("[%s] timed" % mode, r))
        torch.cuda.synchronize()
        _lib.ctx().sync()
        timed += time.perf_counter() - t1
        ph = dmrg.last_projmpo
        timer.factorize_kinds = ()
        done += 1
        per = timed / done
        if timed + per > share or budget.left() < per + (per + fresh_s / done) + 30.0:
            break
    kinds = tuple(timed_kinds) + tuple(factorize_kinds)
    kernels = {_lib.KERNEL_NAMES[k]: _lib.ctx().timing_report(k) for k in kinds}
    _lib.ctx().timing(())
    E = rec[-1]["energy"]
    return {"s_per_sweep": timed / done, "steps": done, "warmup": warmup,
            "timed_s": timed, "energy": E, "maxlinkdim": psi.maxlinkdim(),
            "setup_s": setup_s, "fresh_state_s": fresh_s, "phases": timer.report(),
            "sweeps": rec,
            "warmup_sweeps": [{k: r[k] for k in ("energy", "seconds", "maxerr", "matvecs")}
                              for r in wrec],
            "kernels": {k: v for k, v in kernels.items() if v["launches"]},
            "kernel_sets": {"sweep": [_lib.KERNEL_NAMES[k] for k in timed_kinds],
                            "bond_%d_factorize" % (N // 2):
                                [_lib.KERNEL_NAMES[k] for k in factorize_kinds]}}


def eigh_only(n, reps=2):
    """Wall time of `reps` eigendecompositions of one random n x n density
    matrix (keeping n/2, cutoff 1e-12) after a warm-up, for kernel traces
    of the eigensolver alone."""
    import torch
    from itg import ops
    A = ops.randn(n * n, 3)
    rho = ops.empty(n * n)
    ops.gemm("N", "T", n, n, n, 1.0 / n, A, n, A, n, 0.0, rho, n)
    ops.eigh_trunc(ops.copy(rho), n, maxdim=n // 2, cutoff=1e-12)
    ms = []
    for _ in range(reps):
        r2 = ops.copy(rho)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ops.eigh_trunc(r2, n, maxdim=n // 2, cutoff=1e-12)
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    log("eigh n=%d: %s ms" % (n, ["%.1f" % x for x in ms]))
    return {"n": n, "ms": ms}


def eigensolver_kernels(n):
    """One density-matrix eigendecomposition at the sweep's size (n = d*chi,
    maxdim n/2): wall time of a plain call (`ms`), then a call with its
    tridiagonalisation and GEMM kernels event-timed (per-launch events add
    wall time of their own: `ms_with_kernel_events`)."""
    import torch
    from itg import _lib, ops
    A = ops.randn(n * n, 3)
    rho = ops.empty(n * n)
    ops.gemm("N", "T", n, n, n, 1.0 / n, A, n, A, n, 0.0, rho, n)
    ops.eigh_trunc(ops.copy(rho), n, maxdim=n // 2, cutoff=1e-12)      # warm-up
    r1 = ops.copy(rho)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ops.eigh_trunc(r1, n, maxdim=n // 2, cutoff=1e-12)
    torch.cuda.synchronize()
    plain = time.perf_counter() - t0
    kinds = (_lib.K_TRD_SYMV, _lib.K_TRD_POST, _lib.K_GETT)
    r2 = ops.copy(rho)
    torch.cuda.synchronize()
    _lib.ctx().timing(kinds)
    t0 = time.perf_counter()
    ops.eigh_trunc(r2, n, maxdim=n // 2, cutoff=1e-12)
    _lib.ctx().sync()
    dt = time.perf_counter() - t0
    out = {"n": n, "ms": plain * 1e3, "ms_with_kernel_events": dt * 1e3,
           "kernels": {_lib.KERNEL_NAMES[k]: _lib.ctx().timing_report(k) for k in kinds}}
    _lib.ctx().timing(())
    return out


def precision_check(N, chi, nsweep, krylov, modes, budget, reserve):
    """Every arithmetic form from the same start (random MPS, link dims 10,
    seed 2) through nsweep sweeps with maxdim ramping to chi, cutoff 1e-12:
    converged FP64 energies <psi|H|psi>, compared per site. Stops (and
    reports which forms finished) when the budget would drop below
    `reserve` seconds."""
    from itg.mps import MPO, Sweeps, dmrg, heisenberg_opsum, random_mps
    H = MPO.from_opsum(heisenberg_opsum([(j, j + 1) for j in range(N - 1)]), N)
    ramp = [m for m in (16, 32, 64, 128, 256, 512) if m < chi] + [chi]
    out, recs = {}, {}

    def guard(r, m):
        log("[%s] precision check" % m, r)
        if budget.left() < reserve:
            raise OutOfTime()

    for m in modes:
        t0 = time.perf_counter()
        psi = random_mps(N, 2, 10, seed=2)
        rec = []
        try:
            out[m] = dmrg(H, psi, Sweeps(nsweep, maxdim=ramp, cutoff=[1e-12]), mode=m,
                          krylovdim=krylov, record=rec, sweep_callback=lambda r: guard(r, m))
        except OutOfTime:
            log("[%s] precision check stopped after %d sweeps: out of time" % (m, len(rec)))
            return None, {"stopped": m, "sweeps_done": len(rec)}
        recs[m] = [{"energy": r["energy"], "maxlinkdim": r["maxlinkdim"],
                    "maxerr": r["maxerr"], "seconds": r["seconds"]} for r in rec]
        log("[%s] precision check: E = %.15f after %d sweeps (%.1f s)"
            % (m, out[m], nsweep, time.perf_counter() - t0))
    return out, recs


def kernel_sample(chi, reps=3, eigh_n=0):
    """Short fixed workload for hardware-counter passes: an in-place scale of
    2^27 doubles (exactly 1 GiB read and 1 GiB written: the calibration of
    FETCH_SIZE/WRITE_SIZE), then `reps` prepared split-BF16 Heff applies at
    the bulk-bond shape (chi, w=5, d=2), and with eigh_n > 0 one
    density-matrix eigendecomposition of that size (maxdim n/2). Counter
    passes serialise every dispatch, so the Heff sample stays at a few dozen
    launches; the eigendecomposition adds two per tridiagonalised column."""
    import torch
    from itg import ops
    x = ops.randn(1 << 27, 1)
    for _ in range(3):
        ops.scal(x, 0.5)
    torch.cuda.synchronize()
    del x
    w, d = 5, 2
    L = ops.randn(chi * w * chi, 11)
    R = ops.randn(chi * w * chi, 12)
    W1 = ops.randn(w * w * d * d, 13)
    W2 = ops.randn(w * w * d * d, 14)
    psi = ops.randn(chi * d * d * chi, 15)
    y = ops.empty(chi * d * d * chi)
    work = ops.heff2_work(chi, w, w, w, d, chi)
    dims = (chi, w, w, w, d, chi)
    prep = ops.heff2_prepare(L, W1, W2, R, *dims, mode="bf16x3")
    for _ in range(reps):
        ops.heff2_apply_prepared(L, R, prep, psi, *dims, mode="bf16x3", out=y, work=work)
    torch.cuda.synchronize()
    if eigh_n:
        del L, R, psi, y, work, prep
        A = ops.randn(eigh_n * eigh_n, 3)
        rho = ops.empty(eigh_n * eigh_n)
        ops.gemm("N", "T", eigh_n, eigh_n, eigh_n, 1.0 / eigh_n, A, eigh_n, A, eigh_n, 0.0, rho,
                 eigh_n)
        ops.eigh_trunc(rho, eigh_n, maxdim=eigh_n // 2, cutoff=1e-12)
        torch.cuda.synchronize()
    log("kernel sample done")


def precond_ab(N, chi, sweeps, krylov, train_chi=64, steps_per_bond=10):
    """The learned preconditioner against the method without it: train the
    network by the one-step Ritz value over three sweeps of a different
    random chain (train_chi, steps_per_bond Adam steps per bond), then run
    `sweeps` sweeps at maxdim chi from one random start with no
    preconditioner, the diagonal one, and the guarded network (diagonal
    fallback). All three legs run the host-driven Davidson, so they differ
    only in the preconditioner. Reports energies per sweep, s/sweep, matvecs
    and reverted steps."""
    import numpy as np
    from itg import _lib
    from itg.mps import MPO, Sweeps, dmrg, heisenberg_opsum, random_mps
    from itg.precond import MLPPreconditioner, train_sweep
    H = MPO.from_opsum(heisenberg_opsum([(j, j + 1) for j in range(N - 1)]), N)
    t0 = time.perf_counter()
    net = MLPPreconditioner(2, hidden=256, seed=1)
    gains = np.array(train_sweep(net, H, random_mps(N, 2, train_chi, seed=4), train_chi,
                                 steps_per_bond=steps_per_bond, sweeps=3, lr=3e-4))
    _lib.ctx().sync()
    train_s = time.perf_counter() - t0
    nb = len(gains) // 3
    g0, g1 = float(gains[:nb, 0].sum()), float(gains[:nb, 1].sum())
    log("precond A/B: %d Adam steps, first-sweep step gain %.4f -> %.4f (%.1f s)"
        % (net.step, g0, g1, train_s))
    out = {"train": {"adam_steps": net.step, "first_sweep_gain_start": g0,
                     "first_sweep_gain_end": g1, "seconds": train_s, "chi": train_chi,
                     "steps_per_bond": steps_per_bond, "objective": "one-step Ritz value"}}
    for name, pre in (("none", None), ("diag", "diag"), ("mlp", net.factory())):
        rec = []
        psi = random_mps(N, 2, chi, seed=3)
        dmrg(H, psi, Sweeps(sweeps, maxdim=[chi], cutoff=[1e-12]), krylovdim=krylov,
             precond=pre, record=rec, native=False,
             sweep_callback=lambda r: log("[%s]" % name, r))
        _lib.ctx().sync()
        out[name] = [{"energy": r["energy"], "seconds": r["seconds"], "matvecs": r["matvecs"],
                      "reverted": r.get("reverted", 0)} for r in rec]
    return out


def learned_form(N, chi, krylov, train_chi=64, steps_per_bond=10):
    """The timed sweep in split-BF16 with the learned preconditioner: train
    the guarded network by the one-step Ritz value over three sweeps of a
    random chain at maxdim train_chi (seed 4), then time one sweep at maxdim
    chi from the same random MPS (seed 1) as the plain split-BF16 sweep. The
    network runs inside the device Davidson (itg_davidson2_mlp: 2 steps per
    bond, the first one learned and kept only if it lowers the residual, else
    the diagonal step; the second the diagonal step); its training is
    reported, not timed with the sweep."""
    import torch
    from itg import _lib
    from itg.mps import MPO, PhaseTimer, ProjMPO, Sweeps, dmrg, heisenberg_opsum, random_mps
    from itg.precond import MLPPreconditioner, train_sweep
    H = MPO.from_opsum(heisenberg_opsum([(j, j + 1) for j in range(N - 1)]), N)
    t0 = time.perf_counter()
    net = MLPPreconditioner(2, hidden=256, seed=1)
    train_sweep(net, H, random_mps(N, 2, train_chi, seed=4), train_chi,
                steps_per_bond=steps_per_bond, sweeps=3, lr=3e-4)
    net.net_steps = 1      # the network on the first Davidson step of a bond
    _lib.ctx().sync()
    train_s = time.perf_counter() - t0
    log("[learned] trained: %d Adam steps (%.1f s)" % (net.step, train_s))
    psi = random_mps(N, 2, chi, seed=1)
    ph = ProjMPO(H, "bf16x3")
    ph.build_right(psi)
    timer = PhaseTimer()
    rec = []
    torch.cuda.synchronize()
    _lib.ctx().sync()
    t1 = time.perf_counter()
    E = dmrg(H, psi, Sweeps(1, maxdim=[chi], cutoff=[1e-12]), mode="bf16x3", krylovdim=krylov,
             ph=ph, timer=timer, record=rec, precond=net.factory(),
             sweep_callback=lambda r: log("[learned] timed", r))
    torch.cuda.synchronize()
    _lib.ctx().sync()
    t2 = time.perf_counter()
    return {"s_per_sweep": t2 - t1, "energy": E, "matvecs": rec[0]["matvecs"],
            "matvecs_per_bond": rec[0]["matvecs"] / (2.0 * (N - 1)),
            "maxerr": rec[0]["maxerr"],
            "reverted": rec[0]["reverted"], "phases": timer.report(),
            "train": {"seconds": train_s, "adam_steps": net.step, "chi": train_chi,
                      "objective": "one-step Ritz value"}}


def diag_form(N, chi, krylov):
    """One split-BF16 sweep from the same random MPS (seed 1) with the
    diagonal-preconditioned device Davidson: the classical preconditioner
    beside the plain and learned forms."""
    import torch
    from itg import _lib
    from itg.mps import MPO, PhaseTimer, ProjMPO, Sweeps, dmrg, heisenberg_opsum, random_mps
    H = MPO.from_opsum(heisenberg_opsum([(j, j + 1) for j in range(N - 1)]), N)
    psi = random_mps(N, 2, chi, seed=1)
    ph = ProjMPO(H, "bf16x3")
    ph.build_right(psi)
    rec = []
    torch.cuda.synchronize()
    _lib.ctx().sync()
    t1 = time.perf_counter()
    E = dmrg(H, psi, Sweeps(1, maxdim=[chi], cutoff=[1e-12]), mode="bf16x3", krylovdim=krylov,
             ph=ph, record=rec, precond="diag",
             sweep_callback=lambda r: log("[diag] timed", r))
    torch.cuda.synchronize()
    _lib.ctx().sync()
    t2 = time.perf_counter()
    return {"s_per_sweep": t2 - t1, "energy": E, "matvecs": rec[0]["matvecs"],
            "matvecs_per_bond": rec[0]["matvecs"] / (2.0 * (N - 1)),
            "maxerr": rec[0]["maxerr"]}


def cut_dims(N, d, chi):
    return [1] + [min(chi, d ** min(j, N - j)) for j in range(1, N)] + [1]


def sweep_cost_units(N, d, chi, w):
    """Sum over the 2(N-1) bond updates of a sweep of the Heff FLOPs,
    in units of one full-size (chi x chi) bond."""
    from itg.mps import heff_flops
    c = cut_dims(N, d, chi)
    full = heff_flops(chi, w, w, d, chi)
    tot = sum(heff_flops(c[b], w, w, d, c[b + 2]) for b in range(N - 1))
    return 2.0 * tot / full


def cpu_baseline(chi, w=5, d=2, krylov=3, seed=0, eig_n=0):
    """The oracle's BLAS form (numpy FP64, OpenBLAS/LAPACK) on one bulk bond
    update at bond dimension chi, at the sweep's own shapes: one Heff apply
    (timed, counted `krylov` times), the density matrix P P^T, its
    eigendecomposition (LAPACK syevd at n = d chi, ~10 s on 16 threads; with
    0 < eig_n < d chi, the leading eig_n block instead, scaled by
    (d chi / eig_n)^3) and one environment update. Returns (seconds of one
    bond update, description of the parts)."""
    import numpy as np
    from oracle import dmrg as odmrg
    rng = np.random.default_rng(seed)
    L = rng.standard_normal((chi, w, chi))
    R = rng.standard_normal((chi, w, chi))
    W = rng.standard_normal((w, w, d, d))
    psi = rng.standard_normal((chi, d, d, chi))
    parts = {}
    t0 = time.perf_counter()
    x = odmrg.heff_apply_blas(L, W, W, R, psi)
    parts["heff_apply"] = time.perf_counter() - t0
    log("cpu baseline: Heff apply at chi=%d: %.1f s" % (chi, parts["heff_apply"]))
    t0 = time.perf_counter()
    P = np.reshape(x / np.linalg.norm(x), (chi * d, d * chi), order="F")
    rho = P @ P.T
    parts["density_matrix"] = time.perf_counter() - t0
    log("cpu baseline: density matrix n=%d: %.1f s" % (chi * d, parts["density_matrix"]))
    n = min(eig_n, chi * d) if eig_n > 0 else chi * d
    t0 = time.perf_counter()
    ev, U = np.linalg.eigh(rho[:n, :n])
    parts["eigh_sample"] = time.perf_counter() - t0
    parts["eigh_scaled"] = parts["eigh_sample"] * (chi * d / n) ** 3
    log("cpu baseline: eigh n=%d: %.1f s (scaled to n=%d: %.1f s)"
        % (n, parts["eigh_sample"], chi * d, parts["eigh_scaled"]))
    Q, _ = np.linalg.qr(rng.standard_normal((chi * d, chi)))
    A = np.reshape(Q, (chi, d, chi), order="F")
    t0 = time.perf_counter()
    odmrg.update_left_blas(L, A, W)
    parts["environment"] = time.perf_counter() - t0
    log("cpu baseline: environment update: %.1f s" % parts["environment"])
    bond = (krylov * parts["heff_apply"] + parts["density_matrix"] + parts["eigh_scaled"] +
            parts["environment"])
    return bond, parts


def blas_threads():
    """Threads the host BLAS runs with (threadpoolctl), else the CPUs this
    process may use."""
    try:
        from threadpoolctl import threadpool_info
        n = max([i.get("num_threads", 0) for i in threadpool_info()
                 if i.get("user_api") == "blas"] or [0])
        if n:
            return n
    except Exception:  # noqa: BLE001
        pass
    return len(os.sched_getaffinity(0))


def micro(reps=10, chi=1024):
    """Prepared Heff apply at chi (w=5, d=2) in each mode."""
    import numpy as np
    import torch
    from itg import ops
    from itg.mps import heff_flops
    w, d = 5, 2
    out = {}
    L = ops.randn(chi * w * chi, 11)
    R = ops.randn(chi * w * chi, 12)
    W1 = ops.randn(w * w * d * d, 13)
    W2 = ops.randn(w * w * d * d, 14)
    x = ops.randn(chi * d * d * chi, 15)
    y = ops.empty(chi * d * d * chi)
    work = ops.heff2_work(chi, w, w, w, d, chi)
    dims = (chi, w, w, w, d, chi)
    for mode in ("native", "fp32", "bf16x3"):
        prep = ops.heff2_prepare(L, W1, W2, R, *dims, mode=mode)
        for _ in range(2):
            ops.heff2_apply_prepared(L, R, prep, x, *dims, mode=mode, out=y, work=work)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            ops.heff2_apply_prepared(L, R, prep, x, *dims, mode=mode, out=y, work=work)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        fl = heff_flops(chi, w, w, d, chi)
        out[mode] = {"ms": ms, "tflops": fl / ms / 1e9, "frac": fl / ms / 1e9 / PEAK[mode]}
    return out


def roofline_from(kd, bound, peak, unit, launch_set):
    """roofline object of one kernel class from its event-timed launches."""
    n = max(kd["launches"], 1)
    if bound == "mfma":
        achieved = kd["flops"] / kd["ms"] / 1e9          # TFLOP/s
    else:
        achieved = kd["bytes"] / kd["ms"] / 1e6          # GB/s
    return {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "traffic": None, "launches": kd["launches"],
            "avg_ms": kd["ms"] / n, "flops_per_launch": kd["flops"] / n,
            "bytes_per_launch": kd["bytes"] / n, "launch_set": launch_set}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--budget", type=float, default=450.0,
                    help="wall seconds for the whole run, counted from process start")
    ap.add_argument("--N", type=int, default=100)
    ap.add_argument("--chi", type=int, default=2048)
    ap.add_argument("--krylov", type=int, default=3)
    ap.add_argument("--modes", default="bf16x3,native")
    ap.add_argument("--micro", action="store_true")
    ap.add_argument("--precond-ab", action="store_true",
                    help="learned vs diagonal vs no preconditioner, N=--N, maxdim=--ab-chi")
    ap.add_argument("--ab-chi", type=int, default=256)
    ap.add_argument("--ab-sweeps", type=int, default=3)
    ap.add_argument("--kernel-sample", action="store_true",
                    help="short fixed workload for rocprofv3 --pmc passes (see kernel_sample)")
    ap.add_argument("--sample-eigh", type=int, default=0,
                    help="with --kernel-sample: one eigendecomposition of this size after it")
    ap.add_argument("--config3", action="store_true",
                    help="config-3 start: link dims 10, ramp 64..1024, cutoff 1e-12")
    ap.add_argument("--cpu-chi", type=int, default=0,
                    help="bond dimension of the CPU baseline's sample (0 = --chi)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pc-chi", type=int, default=1024,
                    help="maxdim of the converged precision check (0 = skip)")
    ap.add_argument("--pc-sweeps", type=int, default=10)
    ap.add_argument("--eigh", type=int, default=0,
                    help="only time two density-matrix eigendecompositions of this size (n)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the split-BF16 + learned-preconditioner sweep")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the sweep's Heff/eigensolver path runs on a "
              "single MI355X here; multi-GPU sharding is a design note, not built", file=sys.stderr)
        sys.exit(2)
    import torch
    from itg import _lib
    info = _lib.ctx().info()
    log("device", info)
    if args.micro:
        print(json.dumps({"micro": {c: micro(chi=c) for c in (1024, 2048)}, "device": info}))
        return
    if args.kernel_sample:
        kernel_sample(args.chi, eigh_n=args.sample_eigh)
        return
    if args.eigh:
        print(json.dumps({"eigh": eigh_only(args.eigh), "device": info}))
        return
    if args.precond_ab:
        print(json.dumps({"precond_ab": precond_ab(args.N, args.ab_chi, args.ab_sweeps,
                                                   args.krylov), "device": info}))
        return
    budget = Budget(args.budget)
    modes = args.modes.split(",")
    L_ = _lib
    res = {}
    head = modes[0]
    warm = min(args.warmup, 1)
    notes = []
    if warm < args.warmup:
        notes.append("warmup capped at 1 sweep per form (asked %d) to keep the run within "
                     "the %.0f s budget" % (args.warmup, args.budget))
    # the headline form event-times every split-GEMM and W-pass launch of its
    # timed sweeps, plus the eigensolver kernels of the middle bond's
    # factorisation in the first timed sweep
    timed = (L_.K_GEMM_SPLIT, L_.K_WPASS) if head == "bf16x3" else (L_.K_GETT, L_.K_WPASS)
    fk = (L_.K_TRD_SYMV, L_.K_TRD_POST)
    shares = {head: 0.30 * args.budget}
    for m in modes[1:]:
        shares[m] = 0.14 * args.budget
    for m in modes:
        res[m] = run_mode(m, args.N, args.chi, args.steps, warm, args.krylov, budget,
                          shares[m], args.config3, timed_kinds=timed if m == head else (),
                          factorize_kinds=fk if m == head else ())
        log(m, {k: v for k, v in res[m].items() if k != "sweeps"})
        if res[m]["steps"] < args.steps:
            notes.append("%s: %d of %d timed sweeps fit its share of the budget"
                         % (m, res[m]["steps"], args.steps))
    r = res[head]
    heff = r["phases"].get("davidson") or r["phases"].get("heff") or \
        {"ms": float("nan"), "flops": 0.0, "calls": 0}
    # converged precision check; reserve time for the CPU baseline
    pc, pc_rec = None, None
    if len(modes) > 1 and args.pc_chi > 0:
        if budget.left() > 150:
            pc, pc_rec = precision_check(args.N, args.pc_chi, args.pc_sweeps, args.krylov,
                                         modes, budget, reserve=60.0)
        else:
            notes.append("precision check skipped: %.0f s left" % budget.left())
    cpu = None
    if not args.no_cpu_baseline:
        cchi = args.cpu_chi or args.chi
        if budget.left() > 90:
            t_bond, parts = cpu_baseline(cchi, krylov=args.krylov)
            units = sweep_cost_units(args.N, 2, args.chi, 5) * (args.chi / cchi) ** 3
            cpu = {"value": t_bond * units, "unit": "s/sweep", "cores": blas_threads(),
                   "kind": "port",
                   "sample": "oracle BLAS form (numpy FP64, OpenBLAS/LAPACK) on one bulk bond "
                             "update at chi=%d (w=5, d=2), timed part by part: Heff apply "
                             "%.2f s (x%d Krylov vectors), density matrix %.2f s, LAPACK "
                             "syevd at n=%d %.2f s, environment update %.2f s: %.1f s per "
                             "bond, times %.1f bond equivalents per sweep (the sweep's Heff "
                             "FLOPs in units of one full-size bond)"
                             % (cchi, parts["heff_apply"], args.krylov,
                                parts["density_matrix"], 2 * cchi, parts["eigh_scaled"],
                                parts["environment"], t_bond, units),
                   "parts_s": parts}
            log("cpu baseline", cpu)
        else:
            notes.append("CPU baseline skipped: %.0f s left" % budget.left())
    learned = None
    if head == "bf16x3" and not args.no_learned and not args.config3:
        est = 1.3 * r["s_per_sweep"] + 30.0
        if budget.left() > est + 20:
            learned = learned_form(args.N, args.chi, args.krylov)
            log("learned form", {k: v for k, v in learned.items() if k != "phases"})
        else:
            notes.append("learned form skipped: %.0f s left, ~%.0f s needed"
                         % (budget.left(), est))
    diag = None
    if learned is not None:
        if budget.left() > r["s_per_sweep"] + 25:
            diag = diag_form(args.N, args.chi, args.krylov)
            log("diag form", diag)
        else:
            notes.append("diagonal-preconditioner form skipped: %.0f s left" % budget.left())
    eig = None
    if budget.left() > 15:
        eig = eigensolver_kernels(2 * args.chi)
        log("eigensolver", eig)
    kern = r.get("kernels", {})
    # roofline of the Heff's split GEMM (every launch of the timed sweeps)
    roof_heff = None
    kname = L_.KERNEL_NAMES[L_.K_GEMM_SPLIT if head == "bf16x3" else L_.K_GETT]
    if kname in kern:
        roof_heff = roofline_from(kern[kname], "mfma", PEAK[head], "TFLOP/s",
                                  "every launch of the timed sweeps")
        roof_heff["kernel"] = kname
        tpath = os.path.join(ROOT, "profiles", "gemm_split_wide_kernel_traffic_chi2048.json")
        if kname.startswith("gemm_split") and os.path.exists(tpath):
            tr = json.load(open(tpath))
            roof_heff["traffic"] = tr["bytes_per_launch"]
            roof_heff["traffic_shape"] = tr["shape"]
            roof_heff["traffic_algorithmic"] = tr["algorithmic_bytes_per_launch"]
    # roofline of the tridiagonalisation's symmetric product (the kernel with
    # the most time per sweep), timed inside the first timed sweep at the
    # middle bond (n = 2 chi)
    roof_symv = None
    sname = L_.KERNEL_NAMES[L_.K_TRD_SYMV]
    if sname in kern:
        roof_symv = roofline_from(kern[sname], "hbm", HBM_PEAK, "GB/s",
                                  "every launch of the factorisation of bond %d in the first "
                                  "timed sweep (n=%d)" % (args.N // 2, 2 * args.chi))
        roof_symv["kernel"] = sname
        tpath = os.path.join(ROOT, "profiles", "trd_symv_kernel_traffic_n4096.json")
        if os.path.exists(tpath):
            tr = json.load(open(tpath))
            roof_symv["traffic"] = tr["bytes_per_launch"]
            roof_symv["traffic_algorithmic"] = tr["algorithmic_bytes_per_launch"]
    # which of the two holds more of the sweep: the symv share of the
    # factorisation phase (from the timed bond) against the GEMM's total
    dominant = roof_heff
    per_sweep = {}
    if roof_heff is not None:
        per_sweep[roof_heff["kernel"]] = kern[roof_heff["kernel"]]["ms"] / r["steps"]
    if roof_symv is not None and "factorize_timed_bond" in r["phases"]:
        # the symv's share of the timed bond's factorisation, times the
        # sweep's factorisation time (bulk bonds share the timed bond's shape)
        fact = r["phases"]["factorize"]["ms"] / r["steps"]
        share = kern[sname]["ms"] / max(r["phases"]["factorize_timed_bond"]["ms"], 1e-9)
        per_sweep[sname] = share * fact
        if roof_heff is None or per_sweep[sname] > per_sweep[roof_heff["kernel"]]:
            dominant = roof_symv
    line = {
        "metric": METRIC,
        "value": r["s_per_sweep"],
        "unit": "s/sweep",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": r["warmup"],
        "ms_per_step": r["s_per_sweep"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3 (fp32 acc) + f64", "native": "f64", "fp32": "f32 + f64"}[head],
        "data": "synthetic",
        "config": {"workload": ("Heisenberg S=1/2 chain N=%d two-site DMRG, %d Krylov vectors; "
                                % (args.N, args.krylov)) +
                               ("random MPS link dims 10 seed 1, maxdim ramp %s, cutoff 1e-12, "
                                "then sweeps at maxdim %d" % (ramp_for(args.chi), args.chi)
                                if args.config3 else
                                "random right-canonical MPS at link dims min(2^j, 2^(N-j), %d) "
                                "seed 1, cutoff 1e-12 (flat random spectra: bulk bonds stay at %d)"
                                % (args.chi, args.chi)) +
                               "; one step = one sweep at maxdim %d" % args.chi,
                   "N": args.N, "chi": args.chi, "mpo_bond_dim": 5, "modes": modes},
        "roofline": dominant,
        "roofline_heff_gemm": roof_heff,
        "roofline_tridiagonalisation": roof_symv,
        "kernel_ms_per_sweep": per_sweep,
        "davidson": {"bonds": heff["calls"], "avg_ms": heff["ms"] / max(heff["calls"], 1),
                     "heff_tflops": heff["flops"] / max(heff["ms"], 1e-9) / 1e9},
        "eigensolver": eig,
        "results": {m: {k: v for k, v in res[m].items() if k != "sweeps"} for m in modes},
        "budget_s": args.budget,
        "notes": notes,
    }
    if len(modes) > 1:
        e = [res[m]["energy"] for m in modes]
        line["speedup_vs_" + modes[1]] = res[modes[1]]["s_per_sweep"] / r["s_per_sweep"]
        # after one sweep from a random state the two forms follow different
        # Davidson trajectories, so this difference is not a precision figure
        line["dE_per_site_timed_sweep"] = abs(e[0] - e[1]) / args.N
        if pc is not None:
            line["dE_per_site"] = abs(pc[modes[0]] - pc[modes[1]]) / args.N
            line["precision_check"] = {
                "workload": "same chain, random MPS link dims 10 seed 2, %d sweeps, maxdim ramp "
                            "to %d, cutoff 1e-12; FP64 energies of the converged states"
                            % (args.pc_sweeps, args.pc_chi),
                "energy": pc, "sweeps": pc_rec}
        elif pc_rec is not None:
            line["precision_check"] = {"incomplete": pc_rec}
    if diag is not None:
        line["diag_form"] = diag
    if learned is not None:
        learned["energy_plain_sweep"] = r["energy"]
        learned["plain_matvecs_per_bond"] = r["sweeps"][0]["matvecs"] / (2.0 * (args.N - 1)) \
            if r.get("sweeps") else None
        if len(modes) > 1:
            learned["speedup_vs_" + modes[1]] = res[modes[1]]["s_per_sweep"] / learned["s_per_sweep"]
        line["learned_form"] = learned
    if cpu is not None:
        line["cpu_baseline"] = cpu
    line["wall_s"] = budget.used()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
