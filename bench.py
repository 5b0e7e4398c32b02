#!/usr/bin/env python3
"""Benchmark of the steady-state VMC loop on one MI355X.

Workload (the metric's configuration): dissipative XYZ model on a 4x4
periodic lattice (Jx=0.9, Jy=1.0, Jz=1.1, jumps sqrt(gamma) sigma^- with
gamma=1), NDM alpha_h = alpha_a = 2 (Np = 2176), parameters N(0, 0.01^2) from
a fixed seed, 4096 chains x 4 samples per SR step (Ns = 16384), LocalRule
Metropolis with 2N+1 = 33 passes, SR additive shift 0.01, lr 0.005.

One "step" = one SR iteration: sampling, log rho / O / L_loc / force, loss
statistics, S build, shifted Cholesky solve and the parameter update, all on
the device.  Three forms are timed:
  fp64             the plain form: FP64 network and S build, the reference's
                   schedule (chains restarted and burned in (50 calls) every
                   iteration);
  fp64_persistent  the plain precision on the chosen schedule (persistent
                   chains), so schedule and precision gains stand apart;
  chosen           FP32 network kernels, BF16 S build on MFMA with FP32
                   accumulation, FP64 solve, persistent chains (burn-in only
                   once).
`value` is the chosen form's SR iterations/s.

Prints ONE JSON line.  --gpus N > 1 is not supported (one GPU per process,
libnqhip has no collectives).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neuralquantum.jl_amd"))
sys.path.insert(0, ROOT)

from nqhip import _lib, autotune, operators, problems  # noqa: E402

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}

CFG = dict(lx=4, ly=4, alpha_h=2, alpha_a=2, chains=4096, chain_length=4, passes=33, burn=50,
           eps=0.01, lr=0.005, sigma=1e-4, param_seed=1234, sampler_seed=4321)


def hip_sync():
    try:
        lib = ctypes.CDLL("libamdhip64.so")
        lib.hipDeviceSynchronize()
    except OSError:
        pass


def build_ctx(form: str, cfg=CFG):
    _, _, _, L = problems.dissipative_xyz(cfg["lx"], cfg["ly"], Jx=0.9, Jy=1.0, Jz=1.1, gamma=1.0)
    N = cfg["lx"] * cfg["ly"]
    nh, na = cfg["alpha_h"] * N, cfg["alpha_a"] * N
    ctx = _lib.Context(0, cfg["sampler_seed"])
    ctx.set_problem(operators.flatten(L))
    ctx.set_ndm(N, nh, na)
    rng = np.random.default_rng(cfg["param_seed"])
    p = rng.standard_normal(ctx.np_) * math.sqrt(cfg["sigma"])   # init_random_pars!(net; sigma=1e-4)
    ctx.set_params(p)
    if form == "fp64":
        ctx.set_precision(_lib.PREC_FP64, _lib.PREC_FP64)
        flags = 0
    elif form == "fp64_persistent":      # the precision of the plain form, the chosen form's schedule
        ctx.set_precision(_lib.PREC_FP64, _lib.PREC_FP64)
        flags = _lib.SAMPLER_PERSISTENT
    else:
        ctx.set_precision(_lib.PREC_FP32, _lib.PREC_BF16)
        flags = _lib.SAMPLER_PERSISTENT
    ctx.set_sampler(cfg["chains"], cfg["chain_length"], cfg["passes"], cfg["burn"], flags)
    shape = autotune.load(TUNED["chosen"]) if form == "chosen" else None
    if shape:
        ctx.set_launch_shape(*shape)       # the autotuner's stored choice for this workload
    return ctx, p


def time_form(form, steps, warmup, cfg=CFG):
    ctx, _ = build_ctx(form, cfg)
    ctx.set_graph(True)          # the production path: one hipGraph replay per SR iteration
    losses = []
    for _ in range(warmup):
        losses.append(ctx.sr_iteration(cfg["eps"], cfg["lr"])[0])
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        losses.append(ctx.sr_iteration(cfg["eps"], cfg["lr"])[0])
    hip_sync()
    dt = time.perf_counter() - t0
    # a second, profiled pass of the same length for the per-phase breakdown
    # (eager launches: HIP events around every phase on the library's stream)
    ctx.set_profiling(True)
    for _ in range(steps):
        losses.append(ctx.sr_iteration(cfg["eps"], cfg["lr"])[0])
    phases = ctx.phase_times()
    ctx.set_profiling(False)
    return ctx, dt, phases, losses


# stored launch shapes (nqhip/launch_shapes.json, written by --autotune)
TUNED = {"chosen": "xyz44_chosen", "instances": "ising64_packed"}

C5 = dict(instances=64, nmin=6, nmax=16, gmin=0.5, gmax=2.0, seed=2024, chains=256, chain_length=8, burn=50,
          eps=0.01, lr=0.01, sigma=1e-4)


def build_instances(cfg=C5):
    """The batched-instances workload: dissipative Ising chains (periodic,
    V=2, gamma=1) with N ~ U{6..16} and g ~ U[0.5, 2], NDM alpha = 1,
    256 chains x 8 samples, 2N+1 passes, in the chosen form."""
    rng = np.random.default_rng(cfg["seed"])
    ns = rng.integers(cfg["nmin"], cfg["nmax"] + 1, size=cfg["instances"])
    gs = rng.uniform(cfg["gmin"], cfg["gmax"], size=cfg["instances"])
    members = []
    for i, (N, g) in enumerate(zip(ns, gs)):
        N = int(N)
        _, _, _, L = problems.dissipative_ising_chain(N, V=2.0, g=float(g), gamma=1.0, periodic=True)
        ctx = _lib.Context(0, cfg["seed"] + i)
        ctx.set_problem(operators.flatten(L))
        ctx.set_ndm(N, N, N)
        prng = np.random.default_rng(cfg["seed"] + 1000 + i)
        ctx.set_params(prng.standard_normal(ctx.np_) * math.sqrt(cfg["sigma"]))
        ctx.set_precision(_lib.PREC_FP32, _lib.PREC_BF16)
        ctx.set_sampler(cfg["chains"], cfg["chain_length"], 2 * N + 1, cfg["burn"], _lib.SAMPLER_PERSISTENT)
        shape = autotune.load(TUNED["instances"])
        if shape:
            ctx.set_launch_shape(*shape)
        members.append(ctx)
    return members, ns


def time_instances(steps, warmup, cfg=C5):
    """Aggregate SR iterations/s of the batched instances, four ways: one
    after another (each member replaying its own graph), all members queued
    on their own streams then awaited, one graph per iteration holding every
    member's own launches, and the packed form (one launch per kernel for all
    members, one graph per iteration)."""
    eps, lr = cfg["eps"], cfg["lr"]
    res, losses = {}, {}
    launches = None
    for mode in ("sequential", "streams", "batch_graph", "batch_packed"):
        members, ns = build_instances(cfg)
        batch = None
        if mode.startswith("batch"):
            batch = _lib.Batch(members)
            batch.set_mode(_lib.BATCH_GRAPH if mode == "batch_graph" else _lib.BATCH_PACKED)
        else:
            for c in members:
                c.set_graph(True)

        def one_iteration():
            if mode == "sequential":
                return np.array([c.sr_iteration(eps, lr) for c in members])
            if mode == "streams":
                for c in members:
                    c.sr_iteration_async(eps, lr)
                return np.array([c.wait() for c in members])
            return batch.sr_iteration(eps, lr)

        for _ in range(warmup):
            one_iteration()
        hip_sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = one_iteration()
        hip_sync()
        dt = time.perf_counter() - t0
        res[mode] = cfg["instances"] * steps / dt
        losses[mode] = st[:, 0].copy()
        if mode == "batch_packed":
            launches = batch.launches()
        if batch is not None:
            batch.close()
        for c in members:
            c.close()
    d = max(float(np.max(np.abs(losses[m] - losses["sequential"]) / np.abs(losses["sequential"])))
            for m in ("batch_graph", "batch_packed"))
    np_total = int(sum(5 * n + 4 * n * n for n in ns))
    return {"workload": (f"{cfg['instances']} dissipative Ising chains (periodic, V=2, gamma=1), N~U{{6..16}}, "
                         f"g~U[0.5,2] (seed {cfg['seed']}), NDM alpha=1 (sum Np = {np_total}), "
                         f"{cfg['chains']} chains x {cfg['chain_length']} samples, 2N+1 passes, persistent "
                         f"chains, FP32 network + BF16 S build + FP64 solve, shift 0.01, lr 0.01"),
            "unit": "SR iters/s summed over instances",
            "aggregate_it_s": {k: round(v, 1) for k, v in res.items()},
            "steps": steps, "warmup": warmup,
            "packed_launches_per_iteration": launches,
            "launch_shape": autotune.load(TUNED["instances"]),
            "max_rel_loss_diff_batch_vs_sequential": d}


C4 = dict(mix=0.5, retrain_every=20, train_steps=50, batch=1024, lr=1e-3, buffer=65536)


def learned_proposal(steps, cfg=CFG, c4=C4):
    """The learned-proposal workload: the chosen form on the 4x4 XYZ NDM with a
    global MADE move before half of the stored slices (retrained every 20 SR
    iterations: 50 Adam steps, batch 1024, lr 1e-3, on the last 65536
    samples), next to the same run without it; then, at the parameters
    reached, tau_int of log p along chains that make one local proposal per
    stored sample, without and with a move of the trained proposal before
    every stored sample (1024 chains x 48 slices)."""
    eps, lr = cfg["eps"], cfg["lr"]
    warm = 2 * c4["retrain_every"]           # two retrainings before the timed region
    out = {}
    for name in ("local_only", "learned"):
        ctx, _ = build_ctx("chosen", cfg)
        if name == "learned":
            ctx.made_setup(c4["mix"], c4["retrain_every"], c4["train_steps"], c4["batch"], c4["lr"], c4["buffer"])
        ctx.set_graph(True)
        for _ in range(warm):
            ctx.sr_iteration(eps, lr)
        hip_sync()
        t0 = time.perf_counter()
        losses = [ctx.sr_iteration(eps, lr)[0] for _ in range(steps)]
        hip_sync()
        dt = time.perf_counter() - t0
        out[name] = {"it_s": round(steps / dt, 2), "loss_mean_last": float(np.mean(losses[-10:]))}
        if name == "learned":
            prop, acc, fill = ctx.made_stats()
            out[name]["global_moves_last_call"] = {"proposed": prop, "accepted": acc, "buffer": fill}
            theta = ctx.made_get_params()
            p = ctx.get_params()
        # the loop's own effective sample rate: tau_int on its own chains and
        # schedule (33 local proposals per stored sample, and for the learned
        # run a global move before half of them) at the parameters reached
        t_lp, t_L = measure_tau(ctx, "chosen", cfg)
        out[name]["tau_int_own_schedule"] = {"logp": round(t_lp, 3), "abs2_Lloc": round(t_L, 3)}
        out[name]["eff_samples_per_s"] = round(cfg["chains"] * cfg["chain_length"] / (2 * max(t_lp, t_L))
                                               * steps / dt, 1)
        ctx.close()
    N = cfg["lx"] * cfg["ly"]
    taus = {}
    for name, mix in (("local_only", 0.0), ("learned", 1.0)):
        c2, _ = build_ctx("chosen", cfg)
        c2.set_params(p)
        c2.set_sampler(1024, 48, 1, 50, 0)
        c2.made_setup(mix, 0, 0, c4["batch"], c4["lr"], 0)
        c2.made_set_params(theta)
        keys, _ = c2.sample()
        lp = 2.0 * c2.logpsi(keys).real.reshape(48, 1024).T
        taus[name] = round(tau_int(lp), 3)
        if mix > 0:
            prop, acc, _ = c2.made_stats()
            taus["learned_acceptance"] = round(acc / max(prop, 1), 4)
        c2.close()
    out["tau_int_logp_one_local_proposal_per_sample"] = taus
    out["config"] = (f"MADE {2 * N}->128->128->{2 * N}, mix {c4['mix']}, retrain every {c4['retrain_every']} "
                     f"iterations ({c4['train_steps']} Adam steps, batch {c4['batch']}, lr {c4['lr']}, buffer "
                     f"{c4['buffer']}); {warm} warm-up + {steps} timed SR iterations per run")
    return out


def evaluate_loss(p, cfg=CFG, repeats=8):
    """<L^dag L> of parameters p from `repeats` independent FP64 gradient
    passes (fresh chains, 50 burn-in calls, Ns samples each): (mean,
    standard error of the mean over the passes)."""
    ctx, _ = build_ctx("fp64", cfg)
    ctx.set_params(p)
    vals = []
    for r in range(repeats):
        ctx.set_iteration(100000 + r)
        vals.append(ctx.sample_grad()[0])
    ctx.close()
    v = np.array(vals)
    return float(v.mean()), float(v.std(ddof=1) / math.sqrt(len(v)))


def equal_loss(iters, cfg=CFG):
    """Every form from the same parameters and seeds for `iters` SR
    iterations.  Per form: the loss of its final parameters evaluated with
    fresh FP64 samples (mean and standard error), the mean sampled loss of
    the last 100 iterations against the 100 before (how far from converged),
    and the final parameters and context (left open: tau_int is measured on
    each form's own chains at these parameters)."""
    res, keep = {}, {}
    for form in ("fp64", "fp64_persistent", "chosen"):
        ctx, _ = build_ctx(form, cfg)
        ctx.set_graph(True)
        losses = np.array([ctx.sr_iteration(cfg["eps"], cfg["lr"])[0] for _ in range(iters)])
        p = ctx.get_params()
        m, e = evaluate_loss(p, cfg)
        last, prev = float(losses[-100:].mean()), float(losses[-200:-100].mean())
        res[form] = {"loss_final_params": m, "loss_final_params_err": e,
                     "sampled_loss_last100": last, "drift_last100_vs_previous100": (last - prev) / prev}
        keep[form] = (ctx, p)
    ref_m, ref_e = res["fp64"]["loss_final_params"], res["fp64"]["loss_final_params_err"]
    for form in ("fp64_persistent", "chosen"):
        m, e = res[form]["loss_final_params"], res[form]["loss_final_params_err"]
        res[form]["rel_diff_vs_fp64"] = (m - ref_m) / ref_m
        res[form]["rel_diff_err"] = math.hypot(e, ref_e * m / ref_m) / ref_m
    return {"iterations": iters, "forms": res,
            "evaluation": "8 independent FP64 gradient passes of 4096 x 4 fresh samples (50 burn-in calls) at "
                          "each form's final parameters: mean and standard error"}, keep


def tau_int(series):
    """Integrated autocorrelation time (Sokal automatic window, c = 5) of
    per-chain series (chains x T), in units of stored samples; tau = 1/2 for
    uncorrelated samples."""
    x = series - series.mean(axis=1, keepdims=True)
    T = x.shape[1]
    var = np.mean(x * x)
    if var == 0:
        return 0.5
    tau = 0.5
    for t in range(1, T):
        rho = np.mean(x[:, :-t] * x[:, t:]) / var
        tau += rho
        if t >= 5 * tau:
            break
    return max(tau, 0.5)


def measure_tau(ctx, form, cfg=CFG, slices=48):
    """tau_int of log p and |L_loc|^2 along the chains of `ctx` on the form's
    own schedule, at its current parameters, in stored samples (passes
    proposals apart): the plain form restarts its chains and burns them in
    (50 calls) as it does every iteration; the persistent forms continue the
    chains their training left (the checkpoint of the chains is restored
    after the sampler is resized to `slices` slices)."""
    chains = cfg["chains"]
    persistent = form != "fp64"
    blob = ctx.save_state() if persistent else None
    ctx.set_sampler(chains, slices, cfg["passes"], cfg["burn"], _lib.SAMPLER_PERSISTENT if persistent else 0)
    if persistent:
        ctx.load_state(blob)
    keys, _ = ctx.sample()
    lp = ctx.logpsi(keys).real.reshape(slices, chains).T
    L = ctx.local_estimator(keys).reshape(slices, chains).T
    return tau_int(2 * lp), tau_int(np.abs(L) ** 2)


VALU_LANE_SLOTS_PER_S = 256 * 64 * 2.4e9   # 256 CUs x 64 lanes x 2.4 GHz


def ndm_sizes(d):
    N, nh, na = d["N"], d["nh"], d["na"]
    np_ = 2 * N + 2 * nh + na + 2 * nh * N + 2 * na * N
    R = 2 * N + 2 * nh + 2 * nh * N + 2 * na + 4 * na * N
    return np_, R


def kernel_rooflines(phases, form, d, pairs):
    """Per-kernel roofline figures from the per-phase HIP-event times.

    Each timed phase is one kernel (sampler: k_metropolis) or a short fixed
    sequence (estimator: the walk launch k_estimate and the forces launch,
    k_force_mma on the matrix cores for the FP32 network; S build: centring +
    Gram launches; solve: shift, blocked Cholesky, solves).  Algorithmic work per
    launch, per unit x units:
      estimator  force contraction (8 nh N + 4 na N) flop per (sample,
                 connection) pair x M pairs; evaluation lane-slots
                 (2 nh k + 20 nh) per changed side + 2 na k + 28 na + 20 per
                 pair (k = flipped sites on that side, k ~ 1.6 here)
      sbuild     4 Np^2 Ns flop
      solve      Np^3 / 3 + 2 Np^2 flop
      sampler    (24 nh + 32 na + 20) lane-slots per chain-step
      post       2 Ns R p bytes (O read twice)"""
    N, nh, na, Ns = d["N"], d["nh"], d["na"], d["Ns"]
    np_, R = ndm_sizes(d)
    elt = 8 if form == "fp64" else 4
    per = {k: v[0] / max(v[1], 1) for k, v in phases.items()}   # ms per occurrence
    fp = "fp64" if form == "fp64" else "fp32"
    out = {}
    t = per["estimator"] * 1e-3
    fl = (8 * nh * N + 4 * na * N) * pairs
    slots = pairs * (2 * nh * 1.6 + 20 * nh + 2 * na * 1.6 + 28 * na + 20)
    ek = "k_estimate (walk + forces)" if form == "fp64" else "k_estimate (walk) + k_force_mma (forces)"
    out["estimator"] = dict(kernel=ek, bound="mfma", achieved=fl / t / 1e12, peak=PEAK_TFLOPS[fp],
                            unit="TFLOP/s", algorithmic=f"(8 nh N + 4 na N) x {pairs} pairs = {fl:.4g} flop",
                            valu_frac=slots / t / VALU_LANE_SLOTS_PER_S,
                            hbm_gbs=Ns * (2 * N / 8 + 4 * elt + R * elt) / t / 1e9, ms=per["estimator"])
    t = per["sbuild"] * 1e-3
    fl = 4.0 * np_ * np_ * Ns
    sp = "fp64" if form == "fp64" else "bf16"
    # the S build computes only what S needs: the lower triangle of the real
    # lambda and imaginary mu blocks (Ns rows each), of the complex Pi block
    # (2 Ns rows) and of the Pi-lambda / Pi-mu blocks, and none of the zero
    # lambda-mu block: 2 Ns (Np - nl)^2 flop.  The figure above counts the
    # full real GEMM, so its fraction can exceed 1
    nl = N + nh + nh * N
    fs = 2.0 * Ns * (np_ - nl) ** 2
    out["sbuild"] = dict(kernel="k_gram" if form == "fp64" else "k_center_t + k_gram_bf16_part + k_gram_finish",
                         bound="mfma",
                         achieved=fl / t / 1e12, peak=PEAK_TFLOPS[sp], unit="TFLOP/s",
                         algorithmic=f"4 Np^2 Ns = {fl:.4g} flop", ms=per["sbuild"],
                         structural_flop=fs, structural_frac=fs / t / 1e12 / PEAK_TFLOPS[sp],
                         frac_note="frac on 4 Np^2 Ns (the full real GEMM); structural_frac on "
                                   "2 Ns (Np - nl)^2, the non-zero lower-triangle blocks actually needed")
    # the factorization, timed on its own, against Np^3/3 flop (the dense
    # Cholesky of S); the whole solve phase (shift, factor, substitutions,
    # checks) is reported beside it.  S's lambda-mu block is zero, so the
    # lambda and mu columns are factored side by side (one k_chol_step launch
    # per 32 columns of lambda for both) before the Pi block: the chain is
    # ceil(nl / 32) + ceil(npi / 32) launches instead of ceil(Np / 32)
    t = per.get("factor", per["solve"]) * 1e-3
    fl = np_ ** 3 / 3
    nl, npi = N + nh + nh * N, na + 2 * na * N
    out["solve"] = dict(kernel="k_chol_step (+ k_split_copy, k_split_schur)", bound="mfma",
                        achieved=fl / t / 1e12,
                        peak=PEAK_TFLOPS["fp64"], unit="TFLOP/s", algorithmic=f"Np^3/3 = {fl:.4g} flop",
                        ms=per.get("factor", per["solve"]), phase_ms=per["solve"],
                        launches=(nl + 31) // 32 + (npi + 31) // 32)
    t = per["post"] * 1e-3
    by = 2.0 * Ns * R * elt
    out["post"] = dict(kernel="k_colsum + k_stats + k_force", bound="hbm", achieved=by / t / 1e9,
                       peak=HBM_PEAK_GBS, unit="GB/s", algorithmic=f"2 Ns R p = {by:.4g} B", ms=per["post"])
    # sampler: (24 nh + 32 na + 20) VALU lane-slots per chain-step; chain-steps
    # per iteration: chains x (burn + L) x passes when the chains restart
    # every iteration, chains x L x passes when they persist
    t = per["sampler"] * 1e-3
    calls = (d["burn"] + d["L"]) if form == "fp64" else d["L"]
    steps = d["chains"] * calls * d["passes"]
    slots = steps * (24 * nh + 32 * na + 20)
    out["sampler"] = dict(kernel="k_metropolis", bound="valu", achieved=slots / t / 1e12,
                          peak=VALU_LANE_SLOTS_PER_S / 1e12, unit="T lane-slots/s",
                          algorithmic=f"(24 nh + 32 na + 20) x {steps} chain-steps = {slots:.4g} lane-slots",
                          ms=per["sampler"])
    for v in out.values():
        if "peak" in v:
            v["frac"] = v["achieved"] / v["peak"]
    out["sampler"]["frac_note"] = "VALU lane-slots (not in the roofline JSON's hbm/mfma choice)"
    return out, per


def count_pairs(ctx, cap=512, chunk=2048):
    """Number M of (sample, connection) pairs with a non-zero matrix element
    over the last iteration's samples (the units of the estimator's work)."""
    keys = ctx.get_samples()
    total = 0
    for i in range(0, len(keys), chunk):
        cnt, _, _ = ctx.connections(keys[i:i + chunk], cap)
        total += int(cnt.astype(np.int64).sum())
    return total


def load_traffic(form, kernel="estimator"):
    """HBM bytes per k_estimate launch, or per factorization (its 51
    k_chol_step launches), measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE scaled by a copy kernel of known size), if a measurement for
    this workload is committed."""
    name = "traffic_k_estimate_xyz44.json" if kernel == "estimator" else "traffic_k_chol_step_xyz44.json"
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            t = json.load(f)
        if kernel == "estimator":
            return t.get(form, {}).get("bytes_per_launch")
        return t.get("bytes_per_factorization")
    except (OSError, ValueError):
        return None


def copy_calibration(nbytes=256 << 20):
    """One elementwise kernel of known traffic (reads `nbytes`, writes
    `nbytes`: dst = 2 src), for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.mul(src, 2.0, out=dst)
    torch.cuda.synchronize()
    del src, dst


def cpu_threads():
    """Host threads for the multi-thread CPU baseline: the cores this process
    may run on, at most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:   # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(threads=1, full=False):
    """The oracle in reference-structured mode (full network recompute per
    proposal, full Np gradient per connection, complex S = Oc Oc^H, LAPACK
    Cholesky) for one SR iteration of the workload, on `threads` host threads
    (OpenMP over chains and samples, the BLAS of the S build and the solve
    limited to the same count).  full=False times a bounded sample (512 of
    the chains, 2048 of the samples) and scales each part linearly;
    full=True runs the whole iteration."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:   # pragma: no cover
        threadpool_limits = None
    from oracle import nqref, ops_ref
    ref = nqref.Ref()
    ref.set_threads(threads)
    cfg = CFG
    N = cfg["lx"] * cfg["ly"]
    nh, na = cfg["alpha_h"] * N, cfg["alpha_a"] * N
    np_ = ref.ndm_np(N, nh, na)
    tab = nqref.Table(ops_ref.table_from_rules(ops_ref.xyz_model(cfg["lx"], cfg["ly"])))
    p = np.random.default_rng(cfg["param_seed"]).standard_normal(np_) * math.sqrt(cfg["sigma"])
    chains, Lc, Ns = cfg["chains"], cfg["chain_length"], cfg["chains"] * cfg["chain_length"]
    ctxm = threadpool_limits(limits=threads) if threadpool_limits else None
    try:
        ch_s = chains if full else 512
        t0 = time.perf_counter()
        keys, _ = ref.metropolis(N, nh, na, p, ch_s, Lc, cfg["passes"], cfg["burn"], cfg["sampler_seed"],
                                 compat=True)
        t_samp = (time.perf_counter() - t0) * chains / ch_s
        nsamp = Ns if full else 2048
        ks = keys if full else np.random.default_rng(1).integers(0, 2 ** (2 * N), size=nsamp, dtype=np.uint64)
        t0 = time.perf_counter()
        _, O = ref.logpsi_grad(N, nh, na, p, ks)
        _, dL = ref.local_estimator(tab, nh, na, p, ks, grad=True)
        t_est = (time.perf_counter() - t0) * Ns / nsamp
        del O, dL
        X = (np.random.default_rng(2).standard_normal((np_, Ns)) +
             1j * np.random.default_rng(3).standard_normal((np_, Ns)))
        t0 = time.perf_counter()
        Sc = X @ X.conj().T
        t_s = time.perf_counter() - t0
        S = Sc.real / Ns + 0.01 * np.eye(np_)
        t0 = time.perf_counter()
        nqref.sr_solve(S, np.ones(np_), 0.0)
        t_chol = time.perf_counter() - t0
    finally:
        if ctxm is not None:
            ctxm.__exit__(None, None, None)
        ref.set_threads(1)
    total = t_samp + t_est + t_s + t_chol
    if full:
        sample = (f"oracle/nqref.c reference-structured, {threads} threads, one whole iteration: Metropolis on all "
                  f"{chains} chains (restart + {cfg['burn']} burn-in calls), estimator with full per-connection "
                  f"gradients on all {Ns} samples, complex S = Oc Oc^H (BLAS) on {Ns} samples, full Cholesky")
    else:
        sample = (f"oracle/nqref.c reference-structured, {threads} thread(s): Metropolis on {ch_s} of {chains} "
                  f"chains, estimator with full per-connection gradients on {nsamp} of {Ns} samples, "
                  f"complex S = Oc Oc^H (BLAS) on all {Ns} samples, full Cholesky; the sampler and the "
                  f"estimator scaled linearly to the full iteration")
    return {"value": 1.0 / total, "unit": "SR iters/s", "cores": threads, "kind": "port", "sample": sample,
            "seconds_per_iteration": {"sampler": t_samp, "estimator": t_est, "sbuild": t_s, "solve": t_chol}}


def run_autotune(path=autotune.PATH):
    """The autotuner on the two tuned workloads; results to `path` (by default
    the package's launch_shapes.json)."""
    ctx, _ = build_ctx("chosen")
    ctx.set_launch_shape(0, 0, 0, 0)
    for _ in range(3):                   # the restart + burn-in iteration, then persistent chains
        ctx.sr_iteration(CFG["eps"], CFG["lr"])
    r = autotune.tune_context(ctx, CFG["eps"], CFG["lr"], iters=20, warmup=3)
    autotune.save(TUNED["chosen"], r, "4x4 XYZ NDM alpha 2, 4096 x 4 samples, chosen form, graph replay", path)
    print(json.dumps({"workload": TUNED["chosen"], **{k: r[k] for k in ("shape", "ms", "default_ms")}}), flush=True)
    ctx.close()
    members, _ = build_instances()
    for c in members:
        c.set_launch_shape(0, 0, 0, 0)
    batch = _lib.Batch(members)
    for _ in range(3):
        batch.sr_iteration(C5["eps"], C5["lr"])
    r = autotune.tune_batch(batch, members, C5["eps"], C5["lr"], iters=10, warmup=2)
    autotune.save(TUNED["instances"], r, "64 Ising chains N 6..16, 256 x 8 samples, packed batch", path)
    print(json.dumps({"workload": TUNED["instances"], **{k: r[k] for k in ("shape", "ms", "default_ms")}}), flush=True)
    batch.close()
    for c in members:
        c.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tau", action="store_true")
    ap.add_argument("--no-instances", action="store_true", help="skip the batched-instances measurement")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal measurement")
    ap.add_argument("--no-equal-loss", action="store_true",
                    help="skip the equal-loss run of the forms (and the tau_int measured at its parameters)")
    ap.add_argument("--equal-loss-iters", type=int, default=600)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 256 MiB device copy (the known-size kernel the --pmc byte counters are "
                         "scaled by)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the launch shapes of the chosen form and of the packed batched instances, store them "
                         "in nqhip/launch_shapes.json and exit")
    ap.add_argument("--autotune-out", default=autotune.PATH, help="where --autotune writes its results")
    args = ap.parse_args()
    if args.autotune:
        run_autotune(args.autotune_out)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: libnqhip runs on one GPU per process and has no multi-GPU path; "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    N = CFG["lx"] * CFG["ly"]
    d = dict(N=N, nh=CFG["alpha_h"] * N, na=CFG["alpha_a"] * N, Ns=CFG["chains"] * CFG["chain_length"],
             chains=CFG["chains"], L=CFG["chain_length"], burn=CFG["burn"], passes=CFG["passes"])
    forms = ("fp64", "fp64_persistent", "chosen")
    res = {}
    for form in forms:
        ctx, dt, phases, losses = time_form(form, args.steps, args.warmup)
        pairs = count_pairs(ctx)
        roofs, per = kernel_rooflines(phases, form, d, pairs)
        res[form] = dict(it_s=args.steps / dt, ms=1e3 * dt / args.steps, per=per, roofs=roofs, pairs=pairs,
                         loss_last=float(losses[-1]))
        ctx.close()
    if args.calibrate_copy:
        copy_calibration()
    ch = res["chosen"]
    out = {
        "metric": "SR iters/s + eff. samples/s, 4x4 NDM, mixed vs FP64 GPU at equal <L^dag L>",
        "value": round(ch["it_s"], 3),
        "unit": "SR iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ch["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (network), bf16->fp32 (S build), fp64 (solve)",
        "data": "synthetic: random-init NDM, samples drawn by the sampler itself",
        "config": {"workload": "dissipative XYZ 4x4 periodic, NDM alpha_h=alpha_a=2 (Np=2176), "
                               "4096 chains x 4 samples, 33 passes, SR shift 0.01, lr 0.005",
                   "forms": {"fp64": "FP64 network + FP64 S build, restart + 50 burn-in calls per iteration "
                                     "(the reference's schedule)",
                             "fp64_persistent": "FP64 network + FP64 S build, persistent chains (the chosen "
                                                "schedule at the plain precision)",
                             "chosen": "FP32 network, BF16 S build (FP32 acc), FP64 solve, persistent chains"},
                   "chosen_launch_shape": autotune.load(TUNED["chosen"])},
    }
    for form in ("fp64", "fp64_persistent"):
        out[form + "_form"] = {"it_s": round(res[form]["it_s"], 3), "ms_per_step": round(res[form]["ms"], 4),
                               "phase_ms": {k: round(v, 4) for k, v in res[form]["per"].items()},
                               "loss_last": res[form]["loss_last"]}
    out["chosen_form"] = {"phase_ms": {k: round(v, 4) for k, v in ch["per"].items()},
                          "pairs_per_iteration": ch["pairs"], "loss_last": ch["loss_last"]}
    out["speedup_chosen_vs_fp64"] = round(ch["it_s"] / res["fp64"]["it_s"], 3)
    out["speedup_schedule_fp64_persistent_vs_fp64"] = round(res["fp64_persistent"]["it_s"] / res["fp64"]["it_s"], 3)
    out["speedup_precision_chosen_vs_fp64_persistent"] = round(ch["it_s"] / res["fp64_persistent"]["it_s"], 3)
    # the dominant kernel of the chosen form: the kernel with the most device
    # time per SR iteration.  Every phase but the solve is one main kernel;
    # the solve phase is ~110 launches of several kernels, of which the
    # factorization (k_chol_step, timed on its own) is the largest, so the
    # solve competes with its factorization time
    roofs = ch["roofs"]
    single = {k: roofs[k]["ms"] for k in roofs if "frac" in roofs[k] and k not in ("solve", "sampler")}
    single["solve"] = ch["per"].get("factor", roofs["solve"]["ms"])
    dom = max(single, key=single.get)
    r = roofs[dom]
    out["roofline"] = {"kernel": r["kernel"], "bound": r["bound"], "achieved": round(r["achieved"], 4),
                       "peak": r["peak"], "unit": r["unit"], "frac": round(r["frac"], 5),
                       "traffic": load_traffic("chosen", dom) if dom in ("estimator", "solve") else None,
                       "algorithmic": r["algorithmic"], "ms_per_launch": round(r["ms"], 4)}
    if dom == "estimator":
        out["roofline"]["valu_lane_slot_frac"] = round(r["valu_frac"], 4)
        out["roofline"]["hbm_algorithmic_gbs"] = round(r["hbm_gbs"], 2)
    out["kernels"] = {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                      for k, v in roofs.items()}
    out["kernels_fp64_form"] = {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                                for k, v in res["fp64"]["roofs"].items() if k in ("sampler", "estimator", "sbuild")}
    if not args.no_equal_loss:
        eq, trained = equal_loss(args.equal_loss_iters)
        out["equal_loss"] = eq
        if not args.no_tau:
            Ns = d["Ns"]
            taus = {form: measure_tau(ctx, form) for form, (ctx, _) in trained.items()}
            out["tau_int"] = {k: {"logp": round(v[0], 3), "abs2_Lloc": round(v[1], 3)} for k, v in taus.items()}
            out["tau_int"]["note"] = ("in stored samples (33 proposals apart), each form on its own schedule at the "
                                      "parameters its equal-loss run reached; 4096 chains x 48 slices")
            out["eff_samples_per_s"] = {
                k: round(Ns / (2 * max(taus[k])) * res[k]["it_s"], 1) for k in forms}
            out["eff_samples_per_s"]["speedup_chosen_vs_fp64"] = round(
                out["eff_samples_per_s"]["chosen"] / out["eff_samples_per_s"]["fp64"], 3)
        for ctx, _ in trained.values():
            ctx.close()
    if not args.no_learned:
        out["learned_proposal"] = learned_proposal(20)
    if not args.no_instances:
        out["batched_instances"] = time_instances(max(2, args.steps // 2), 2)
    if not args.no_cpu_baseline:
        th = cpu_threads()
        base = cpu_baseline(threads=th, full=th > 1)
        base["single_thread"] = cpu_baseline(threads=1, full=False)
        out["cpu_baseline"] = base
    print(json.dumps(out))


if __name__ == "__main__":
    main()
