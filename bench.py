#!/usr/bin/env python3
"""Benchmark of the transient-elasticity hot path on one MI355X.

Metric: Newmark steps/s + adjoint-gradient time, 128x64x64 hex8, mixed vs FP64 GPU.

A bench "step" is one Newmark time step of the mixed-precision forward sweep
(FP32 multigrid with BF16 coarse coefficients inside FP64 flexible CG, rtol
1e-8) on the 128x64x64 cantilever with dt = 0.005 and a filtered-noise density
of mean 0.12.  K steps are timed after W warmup steps.  The same run also
reports:
  * the all-FP64 forward on the same steps (speed-up of mixed over FP64);
  * the adjoint gradient (forward 200 steps + backward sweep with binomial
    checkpointing in 20 state slots) in both modes, with the relative errors of
    the mixed objective and gradient against the FP64 run;
  * the operator microbenchmark (K u, M u, A u in FP64 and FP32, 1000 applies
    each) with its roofline fraction, timed with HIP events on the library's
    stream;
  * the CPU oracle on a bounded sample of the same forward sweep.
Data is synthetic (seeded); nothing is read from disk.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Newmark steps/s + adjoint-gradient time, 128x64x64 hex8, mixed vs FP64 GPU"
PEAK_HBM = 8.0e12          # MI355X HBM3E, B/s
PEAK_FP64 = 78.6e12        # MI355X dense FP64 (vector and matrix), flop/s
PEAK_FP32 = 157.3e12       # MI355X dense FP32 matrix, flop/s
PEAK_FP16 = 2.5e15         # MI355X dense FP16 matrix, flop/s
# matrix instructions of one FP32 A apply at 128x64x64, as rocprofv3 counts them (SQ_INSTS_MFMA summed
# over the dispatch): 520,058 v_mfma_f32_16x16x32_f16 of 16,384 flops each (2,601 tiles of 8x8x4 nodes;
# the tables per tile position leave out elements outside the grid and edge-tile elements that touch
# none of the tile's nodes)
MFMA_EXEC_A32 = 520058 * 16384
# HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, rocprofv3 --pmc, FETCH scaled by 1.99922, the factor the
# counters read for a device copy of 64 MiB; WRITE by 1.0)
TRAFFIC_A32 = 20985242     # FP32 A apply: FETCH 7050.6 KiB, WRITE 6397.7 KiB
TRAFFIC_FCG64 = 84492971   # FP64 flexible-CG apply: FETCH 28375.6 KiB, WRITE 25783.6 KiB
NEL = (128, 64, 64)
DT = 0.005
NLVLS = 4      # multigrid levels of the 128x64x64 workload (coarsest 16x8x8 elements, 4,131 dofs)
# Chebyshev smoother degree of the finest level per precision mode (each mode at its fastest setting on this
# workload); the intermediate levels take the library's default schedule (one damped-Jacobi step before the
# coarse-grid correction, none after) unless COARSE says otherwise
CHEB = {0: 2, 1: 2}
COARSE = (1, 0)


def params(capi, precision, nsteps, slots=0, nlvls=None, coarse=None):
    pre, post = coarse or COARSE
    return capi.default_params(
        nelx=NEL[0], nely=NEL[1], nelz=NEL[2], xc=(0.0, 2.0, 0.0, 1.0, 0.0, 1.0), nsteps=nsteps, t1=DT * nsteps,
        pulse_steps=20, precision=precision, rtol=1e-8, maxit=200, nlvls=nlvls or NLVLS,
        cheb_degree=CHEB.get(precision, 2), checkpoint_slots=slots, rmin=0.08, filter_rtol=1e-10,
        mg_coarse_pre=pre, mg_coarse_post=post)


def density(capi):
    """Seeded U[0,1] noise -> PDE filter (rmin 0.08) -> mean 0.12 -> clamp [0.001, 1]."""
    ctx = capi.Context(params(capi, capi.TT_PREC_FP64, 1, nlvls=1))
    n = ctx.nel
    x = np.random.default_rng(2024).uniform(0.0, 1.0, n)
    xf = ctx.pde_filter(x)
    xf = xf * (0.12 / xf.mean())
    ctx.close()
    return np.clip(xf, 0.001, 1.0)


def op_microbench(capi, torch, nrep):
    """1000 applies of K u, M u, A u at 128x64x64, FP64 and FP32; error of FP32 against FP64."""
    ctx = capi.Context(params(capi, capi.TT_PREC_FP64, 200))
    rng = np.random.default_rng(7)
    x = rng.uniform(0.001, 1.0, ctx.nel)
    ctx.set_density(x)
    nd, nel = ctx.nd, ctx.nel
    u = rng.standard_normal(nd)
    u.reshape(-1, 3)[0:(NEL[0] + 1) * 0] = 0.0
    # zero the clamped dofs (x = 0 face)
    u3 = u.reshape(NEL[2] + 1, NEL[1] + 1, NEL[0] + 1, 3)
    u3[:, :, 0, :] = 0.0
    dev = torch.device("cuda:0")
    u64 = torch.from_numpy(u).to(dev)
    u32 = u64.float()
    y64 = torch.empty_like(u64)
    y32 = torch.empty_like(u32)
    out = {}
    # minimum unique HBM traffic per apply: read u, write y, read kfac and mfac (LinearElasticity.cc:2502-2619)
    for name, which in (("K", capi.TT_OP_K), ("M", capi.TT_OP_M), ("A", capi.TT_OP_A)):
        ctx.op_apply(which, u64, y64, prec=0)
        ctx.op_apply(which, u32, y32, prec=1)
        ref = y64.double()
        err = float(torch.linalg.norm(y32.double() - ref) / torch.linalg.norm(ref))
        ms64 = ctx.time_op_apply(which, u64, y64, 0, nrep)
        ms32 = ctx.time_op_apply(which, u32, y32, 1, nrep)
        # algorithmic flops: K dense 24x24 per element; M as (8x8 scalar) x I3 per element
        fl = 2 * nel * ((576 if which != capi.TT_OP_M else 0) + (192 if which != capi.TT_OP_K else 0))
        b64 = 2 * nd * 8 + 2 * nel * 8
        b32 = 2 * nd * 4 + 2 * nel * 4
        out[name] = {
            "fp64_ms": ms64, "fp32_ms": ms32, "fp64_applies_per_s": 1e3 / ms64, "fp32_applies_per_s": 1e3 / ms32,
            "fp32_rel_err_vs_fp64": err, "flops": fl, "fp64_tflops": fl / (ms64 * 1e-3) / 1e12,
            "fp32_tflops": fl / (ms32 * 1e-3) / 1e12, "fp64_GBs": b64 / (ms64 * 1e-3) / 1e9,
            "fp32_GBs": b32 / (ms32 * 1e-3) / 1e9, "nel": nel, "nd": nd,
        }
    # the FP64 flexible-CG apply (the Krylov iteration's outer residual), timed the same way
    out["fcg_fp64_ms"] = ctx.time_fcg_apply(nrep)
    ctx.close()
    # calibration for the HBM traffic counters: a plain device copy of 64 MiB (64 MiB read + 64 MiB written)
    a = torch.empty(1 << 24, dtype=torch.float32, device=dev).normal_()
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    return out


def forward_run(capi, x, precision, nsteps, coarse=None, nlvls=None):
    ctx = capi.Context(params(capi, precision, nsteps, nlvls=nlvls, coarse=coarse))
    ctx.set_density(x)
    ctx.complete_initial_state()
    ctx.reset_stats()
    t0 = time.perf_counter()
    ctx.forward()
    el = time.perf_counter() - t0
    st = ctx.stats()
    u, _, _ = ctx.get_state(nsteps)
    res = {"seconds": el, "iters_per_solve": st.krylov_iters / max(1, st.solves), "solves": st.solves,
           "max_relres": st.max_relres, "lambda_max": list(st.lambda_max)[:st.nlvls]}
    # the multigrid preconditioner of the last Newmark operator alone: GPU time per application, and the
    # host's time to enqueue one (close to the GPU time: the V-cycle is bound by kernel launches)
    g_ms, h_ms = ctx.time_vcycle(100)
    res["vcycle_ms"] = {"gpu": g_ms, "host_enqueue": h_ms}
    ctx.close()
    return res, u


def gradient_run(capi, x, precision, nsteps, slots):
    """One design iteration's physics on resident data: tt_complete_initial_state (multigrid set-up for
    the Newmark operator, coarsest factorisation, the initial-acceleration solve), then the forward
    sweep (binomial checkpoints in `slots` HBM slots) and the backward adjoint sweep with the
    sensitivities.  `seconds` is the whole of it; the three parts are reported as well."""
    ctx = capi.Context(params(capi, precision, nsteps, slots=slots))
    ctx.set_density(x)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.complete_initial_state()
    t1 = time.perf_counter()
    ctx.forward()
    t2 = time.perf_counter()
    fx, df = ctx.objective_gradient()
    t3 = time.perf_counter()
    st = ctx.stats()
    ctx.close()
    return {"seconds": t3 - t0, "setup_seconds": t1 - t0, "forward_seconds": t2 - t1, "adjoint_seconds": t3 - t2,
            "steps": nsteps, "slots": slots, "forward_steps": st.forward_steps, "adjoint_steps": st.adjoint_steps,
            "solves": st.solves, "iters_per_solve": st.krylov_iters / max(1, st.solves)}, fx, df


def reduced_run(capi, x, precision, nsteps, nb, omega0):
    """The SOMMG reduced-order path on the same workload: basis construction (nb multigrid solves with
    K - omega0^2 M), reduced forward sweep, reduced adjoint with full-mesh sensitivities."""
    p = params(capi, precision, nsteps)
    p.reduction, p.basis, p.omega0 = 1, nb, omega0
    ctx = capi.Context(p)
    ctx.set_density(x)
    t0 = time.perf_counter()
    ctx.complete_initial_state()
    t1 = time.perf_counter()
    ctx.forward()
    t2 = time.perf_counter()
    fx, df = ctx.objective_gradient()
    t3 = time.perf_counter()
    st = ctx.stats()
    ctx.close()
    return {"basis": nb, "omega0": omega0, "seconds": t3 - t0, "basis_seconds": t1 - t0, "forward_seconds": t2 - t1,
            "adjoint_seconds": t3 - t2, "basis_solves": st.solves,
            "iters_per_solve": st.krylov_iters / max(1, st.solves)}, fx, df


def step_roofline_ms(precision, iters, cheb=2):
    """Roofline time of one Newmark step of the 128x64x64 workload: the sum over its kernels of
    max(algorithmic flops / matrix peak, algorithmic bytes / HBM bandwidth), with the measured Krylov
    iterations per solve.  Kernels: the fused right-hand side (K and M contractions of two state
    combinations), the initial residual, the initial guess, the corrector; per iteration the FP64 CG
    apply (stages z, reads p_old and q_old, writes p and q), the CG vector update, and the V-cycle
    (2 cheb operator applies on the finest level, one on each intermediate level (one damped-Jacobi
    step before the coarse-grid correction, none after), with their vector traffic, restriction and
    prolongation, the dense coarsest solve on the packed inverse)."""
    ex, ey, ez = NEL

    def lvl(l):
        f = 1 << l
        nel = (ex // f) * (ey // f) * (ez // f)
        nn = (ex // f + 1) * (ey // f + 1) * (ez // f + 1)
        return nel, nn, 3 * nn

    def t(flops, nbytes, peak):
        return max(flops / peak, nbytes / PEAK_HBM)

    nel0, nn0, nd0 = lvl(0)
    c = 2 * 1152  # flops per element of the stacked [K; M] contraction
    step = t(c * nel0, 8 * (5 * nd0 + 2 * nel0) + nn0, PEAK_FP64)  # right-hand side
    step += t(c * nel0, 8 * (3 * nd0 + 2 * nel0) + nn0, PEAK_FP64)  # initial residual
    step += t(4 * nd0, 8 * 3 * nd0, PEAK_FP64)  # b.b, r.r
    step += t(4 * nd0, 8 * 4 * nd0, PEAK_FP64) + t(6 * nd0, 8 * 6 * nd0, PEAK_FP64)  # guess, corrector
    it = t(c * nel0, 8 * (5 * nd0 + 2 * nel0) + nn0, PEAK_FP64)  # CG apply
    it += t(6 * nd0, 8 * 7 * nd0, PEAK_FP64)  # CG update
    mixed = precision == 1
    for l in range(NLVLS - 1):
        nel, nn, nd = lvl(l)
        es = 4 if mixed else 8
        ces = (4 if l == 0 else 2) if mixed else 8  # coefficient storage (FP32 / BF16 / FP64)
        peak = PEAK_FP32 if mixed else PEAK_FP64
        napp = 2 * cheb if l == 0 else 1  # (pre-smoothing steps - 1) + residual + post-smoothing steps
        it += napp * t(c * nel, es * 6 * nd + ces * 2 * nel + nn, peak)
        _, _, ndc = lvl(l + 1)
        it += t(54 * nd, es * (nd + ndc), peak) + t(16 * nd, es * (2 * nd + ndc), peak)  # restrict, prolong
    _, _, nc = lvl(NLVLS - 1)
    it += t(2 * nc * nc, 8 * nc * (nc + 1) // 2, PEAK_FP64)  # coarsest solve on the packed inverse
    return 1e3 * (step + iters * it)


def gradient_roofline_s(precision, forward_steps, adjoint_steps, iters):
    """Roofline time of the adjoint gradient: forward steps (first sweep and revolve recomputation)
    plus adjoint steps, each a solve of the same operator plus the split K / M apply of the previous
    multiplier, the adjoint vector combination and the sensitivity contraction (two stacked GEMMs,
    4 x 1152 flops per element, reading 7 dof vectors)."""
    ex, ey, ez = NEL
    nel = ex * ey * ez
    nd = 3 * (ex + 1) * (ey + 1) * (ez + 1)
    extra = (max(2 * 1152 * nel / PEAK_FP64, 8 * (3 * nd + 2 * nel) / PEAK_HBM)  # split apply
             + 8 * 10 * nd / PEAK_HBM  # adjoint combination
             + max(4 * 1152 * nel / PEAK_FP64, 8 * (7 * nd + 2 * nel) / PEAK_HBM))  # sensitivities
    step = step_roofline_ms(precision, iters) * 1e-3
    return forward_steps * step + adjoint_steps * (step + extra)


def cpu_baseline(x, steps):
    """The oracle (the FP64 C restatement of the reference's Newmark sweep, OpenMP) on the first `steps`
    Newmark steps of the same forward sweep: same mesh, load, dt, density and rtol; its linear solver is
    Jacobi-preconditioned CG from the previous displacement (the reference's own solver is PETSc's
    FGMRES with a Galerkin multigrid, which needs PETSc)."""
    from oracle import oracle as O
    p = O.cantilever_params(NEL[0], NEL[1], NEL[2], nsteps=steps, dt=DT, pulse_steps=20, rtol=1e-8)
    t0 = time.perf_counter()
    O.forward(p, x)
    el = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": steps / el, "unit": "Newmark steps/s", "cores": threads, "kind": "port", "seconds": el,
            "sample": f"oracle forward, first {steps} of the 200 steps of the 128x64x64 sweep (dt 0.005, rtol 1e-8, "
                      f"the same filtered-noise density), Jacobi-preconditioned CG, {threads} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--slots", type=int, default=20)
    ap.add_argument("--grad-steps", type=int, default=200,
                    help="time steps of the adjoint-gradient leg (the 200-step, 20-slot gradient workload, "
                         "whatever --steps is)")
    ap.add_argument("--nrep", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-reduced", action="store_true", help="skip the SOMMG reduced-order measurement")
    ap.add_argument("--reduced-only", action="store_true",
                    help="only the SOMMG reduced-order gradient, mixed precision (for profiling)")
    ap.add_argument("--ops-only", action="store_true", help="only the operator microbenchmark (for profiling)")
    ap.add_argument("--nlvls", type=int, default=4, help="multigrid levels (4: coarsest grid 16x8x8 elements)")
    ap.add_argument("--cheb-mixed", type=int, default=None, help="Chebyshev degree of the mixed-precision multigrid")
    ap.add_argument("--cheb-fp64", type=int, default=None, help="Chebyshev degree of the FP64 multigrid")
    args = ap.parse_args()
    global NLVLS
    NLVLS = args.nlvls
    if args.cheb_mixed:
        CHEB[1] = args.cheb_mixed
    if args.cheb_fp64:
        CHEB[0] = args.cheb_fp64
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU: --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is described, not built).", file=sys.stderr)
        sys.exit(2)
    import torch
    from topopt_in_petsc_transient_amd import capi
    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.set_device(0)
    capi.lib()

    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    log("density field")
    x = density(capi)
    if args.reduced_only:
        rr, fxr, _ = reduced_run(capi, x, capi.TT_PREC_MIXED, args.steps, 5, 0.01)
        rr["fx"] = fxr
        print(json.dumps({"reduced_model": {"mixed": rr}}))
        return
    log("operator microbenchmark")
    ops = op_microbench(capi, torch, args.nrep)
    if args.ops_only:
        print(json.dumps({"operator": ops}))
        return
    log("warmup")
    forward_run(capi, x, capi.TT_PREC_MIXED, max(1, args.warmup))
    torch.cuda.synchronize()
    log("timed mixed forward")
    mixed, u_mixed = forward_run(capi, x, capi.TT_PREC_MIXED, args.steps)
    torch.cuda.synchronize()
    log("FP64 forward")
    fp64, u_fp64 = forward_run(capi, x, capi.TT_PREC_FP64, args.steps)
    u_err = float(np.linalg.norm(u_mixed - u_fp64) / max(np.linalg.norm(u_fp64), 1e-300))
    # the same sweeps with Chebyshev(2) smoothing on every level (the multigrid schedule as first configured
    # for this workload), for comparison with the default schedule above
    every = {}
    for name, prec in (("mixed", capi.TT_PREC_MIXED), ("fp64", capi.TT_PREC_FP64)):
        log(f"{name} forward, Chebyshev(2) on every level")
        r, _ = forward_run(capi, x, prec, args.steps, coarse=(2, 2))
        every[name] = {"steps_per_s": args.steps / r["seconds"], "iters_per_solve": r["iters_per_solve"],
                       "seconds": r["seconds"]}
    # the same sweeps with one more multigrid level (coarsest grid 8x4x4 elements, a 675-dof dense
    # inverse, instead of 16x8x8 and 4,131 dofs): the workload above is defined with 4 levels
    five = {}
    if NLVLS == 4:
        for name, prec in (("mixed", capi.TT_PREC_MIXED), ("fp64", capi.TT_PREC_FP64)):
            log(f"{name} forward, 5 levels")
            r, _ = forward_run(capi, x, prec, args.steps, nlvls=5)
            five[name] = {"steps_per_s": args.steps / r["seconds"], "iters_per_solve": r["iters_per_solve"],
                          "seconds": r["seconds"]}
    grad = None
    if not args.no_grad:
        log(f"adjoint gradient ({args.grad_steps} steps, {args.slots} slots), mixed")
        gm, fxm, dfm = gradient_run(capi, x, capi.TT_PREC_MIXED, args.grad_steps, args.slots)
        log("adjoint gradient, FP64")
        gd, fxd, dfd = gradient_run(capi, x, capi.TT_PREC_FP64, args.grad_steps, args.slots)
        grad = {"mixed": gm, "fp64": gd, "speedup": gd["seconds"] / gm["seconds"],
                "speedup_sweeps": (gd["forward_seconds"] + gd["adjoint_seconds"]) /
                                  (gm["forward_seconds"] + gm["adjoint_seconds"]),
                "fx_rel_err": abs(fxm - fxd) / abs(fxd),
                "dfdx_rel_err": float(np.linalg.norm(dfm - dfd) / np.linalg.norm(dfd)), "fx": fxd,
                "steps": args.grad_steps, "slots": args.slots}
    reduced = None
    if grad and not args.no_reduced:
        log("reduced-order model (SOMMG, basis 5), mixed and FP64")
        reduced = {}
        for name, prec in (("mixed", capi.TT_PREC_MIXED), ("fp64", capi.TT_PREC_FP64)):
            try:
                rr, fxr, dfr = reduced_run(capi, x, prec, args.grad_steps, 5, 0.01)
                rr["fx"] = fxr
                rr["fx_rel_diff_vs_full"] = abs(fxr - fxd) / abs(fxd)
                rr["dfdx_rel_diff_vs_full"] = float(np.linalg.norm(dfr - dfd) / np.linalg.norm(dfd))
                reduced[name] = rr
            except capi.TTError as e:
                reduced[name] = {"error": str(e)}
        reduced["note"] = ("quasi-static basis: omega0 below the first eigenfrequency (~0.03 rad/s here), as with the "
                           "reference's defaults; this 0.1 s pulse is far from quasi-static, so the reduced response "
                           "differs from the full one (the reduced path's parity is tested against the oracle)")
    cpu = None
    if not args.no_cpu:
        log("CPU oracle sample")
        cpu = cpu_baseline(x, args.cpu_steps)

    # roofline of the dominant kernel of the step: the FP64 flexible-CG apply of every Krylov iteration
    # (z staged, A z contracted, p = z + beta p_old and q = A z + beta q_old = A p formed at the owned
    # nodes, p.q partials), the largest single kernel of the mixed step.  Algorithmic work per launch: the
    # stacked [K; M] 48x24 element contraction, 2 x 1152 flops per element; its algorithmic bytes: z, p_old,
    # q_old read, p, q written (FP64), kfac, mfac, mask.
    fl = 2 * 1152 * ops["A"]["nel"]
    nd0, nel0 = ops["A"]["nd"], ops["A"]["nel"]
    fcg_ms = ops["fcg_fp64_ms"]
    ach = fl / (fcg_ms * 1e-3) / 1e12
    roof = {"bound": "mfma", "achieved": ach, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
            "frac": ach / (PEAK_FP64 / 1e12), "traffic": TRAFFIC_FCG64,
            "kernel": "k_apply_tiles<double,double,4,true,StageOne<double>,EpiFcg> (flexible-CG apply: z staged, p = "
                      "z + beta p_old, q = A z + beta q_old = A p, p.q partials; FP64, 128x64x64, 8x8x4-node tiles, "
                      "v_mfma_f64_16x16x4f64)",
            "kernel_ms": fcg_ms, "flops_per_launch": fl,
            "algorithmic_bytes_per_launch": 8 * (5 * nd0 + 2 * nel0) + nd0 // 3}
    # second entry: the FP32 Newmark-operator apply of the V-cycle smoothers (FP16-split matrix cores)
    a32 = ops["A"]
    ach32 = fl / (a32["fp32_ms"] * 1e-3) / 1e12
    roof["fp32_apply"] = {
        "bound": "mfma", "achieved": ach32, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
        "frac": ach32 / (PEAK_FP32 / 1e12), "traffic": TRAFFIC_A32,
        "kernel": "k_apply_tiles<float,float,4,EpiBC> (A u = (1+a2 beta) K u + (a5+a2 alpha) M u, FP32, "
                  "128x64x64, 8x8x4-node tiles, FP16-split v_mfma_f32_16x16x32_f16)",
        "kernel_ms": a32["fp32_ms"], "flops_per_launch": fl,
        # the FP32-accurate contraction runs as three FP16 products on the matrix cores: executed
        # matrix-instruction work per launch (the counted v_mfma_f32_16x16x32_f16, 16,384 flops each)
        # against the FP16 dense peak
        "matrix_pipe": {"instr": "v_mfma_f32_16x16x32_f16", "executed_flops_per_launch": MFMA_EXEC_A32,
                        "achieved": MFMA_EXEC_A32 / (a32["fp32_ms"] * 1e-3) / 1e12, "peak": PEAK_FP16 / 1e12,
                        "unit": "TFLOP/s", "frac": MFMA_EXEC_A32 / (a32["fp32_ms"] * 1e-3) / PEAK_FP16}}
    # whole-step and whole-gradient rooflines (algorithmic work x measured iterations, against the
    # measured times)
    steps_roof = {}
    for name, run, prec in (("mixed", mixed, 1), ("fp64", fp64, 0)):
        tr = step_roofline_ms(prec, run["iters_per_solve"], CHEB[prec])
        tm = 1e3 * run["seconds"] / args.steps
        steps_roof[name] = {"t_roof_ms": tr, "t_measured_ms": tm, "frac": tr / tm}
    if grad:
        for name, prec in (("mixed", 1), ("fp64", 0)):
            g = grad[name]
            tr = gradient_roofline_s(prec, g["forward_steps"], g["adjoint_steps"], g["iters_per_solve"])
            steps_roof["gradient_" + name] = {"t_roof_s": tr, "t_measured_s": g["seconds"], "frac": tr / g["seconds"]}
    roof["newmark_step"] = steps_roof
    value = args.steps / mixed["seconds"]
    rec = {
        "metric": METRIC, "value": value, "unit": "Newmark steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * mixed["seconds"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64 (mixed: FP32/BF16 multigrid, FP64 outer CG)",
        "data": "synthetic (seeded filtered-noise density, half-sine line load)",
        "config": {"workload": "128x64x64 hex8 cantilever, Newmark dt=0.005, 4-level... see nlvls",
                   "nlvls": NLVLS, "rtol": 1e-8, "cheb_degree_mixed": CHEB[1], "cheb_degree_fp64": CHEB[0],
                   "mg_coarse_pre": COARSE[0], "mg_coarse_post": COARSE[1],
                   "parallelism": "single GPU"},
        "fp64_steps_per_s": args.steps / fp64["seconds"], "mixed_vs_fp64_speedup": fp64["seconds"] / mixed["seconds"],
        "u_final_rel_err_mixed_vs_fp64": u_err, "forward_mixed": mixed, "forward_fp64": fp64,
        "forward_cheb2_every_level": every, "forward_5_levels": five,
        "adjoint_gradient": grad, "reduced_model": reduced, "operator": ops, "roofline": roof, "cpu_baseline": cpu,
    }
    rec["config"]["workload"] = (f"128x64x64 hex8 cantilever (1,635,075 dofs), Newmark beta=1/4 dt={DT}, "
                                 f"{rec['config']['nlvls']}-level MG, Chebyshev({CHEB[1]}) mixed / ({CHEB[0]}) FP64 "
                                 f"on the finest level, one damped-Jacobi step on the intermediate levels, "
                                 f"in FP64 FCG rtol 1e-8")
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
