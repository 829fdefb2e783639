"""Benchmark of the mechanics hot path on one MI355X.

Headline metric: T10 element residual + tangent evaluations per second on the
~200k-element T10 left-ventricle mesh (Holzapfel-Ogden + active fibre tension,
FP64), inputs resident in HBM.  One step = one pass of the fused residual +
analytic-tangent element kernel over all elements of a seeded deformed state.
Also reported: the same pass including deterministic assembly into the block
pattern, a few Newmark beat steps (wall time per step), the kernel's roofline
fraction, and the reference algorithm (central-difference tangent, 61 residual
evaluations per element) on the host cores as the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

# algorithmic HBM bytes per element (residual + analytic tangent): inputs 1548 B
# (conn 40 + coordinates 240 + dN/dX 960 + bases 288 + scalars 20) + outputs
# f 240 + K_e 7200 (FP64) / 3600 (FP32 values)
BYTES_PER_ELEM = {0: 1548 + 240 + 7200, 1: 1548 + 240 + 3600}
FLOPS_PER_ELEM = 42_700
# residual only: inputs 1548 B + f 240 B
RES_BYTES_PER_ELEM = 1548 + 240
# measured HBM bytes per element of the default kernels (FP64 variant 6, mixed variant 5):
# rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, calibrated on a copy kernel of known size,
# on the 200,340-element LV (profiles/element_residual_tangent_lv200k_fp64.md)
TRAFFIC_PER_ELEM = {0: 8_881, 1: 5_263}
HBM_PEAK = 8.0e12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--elements", type=int, default=200_000)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--beat-ms", type=float, default=10.0,
                    help="measured window [0, this] ms of the LV beat (Newmark beta 1/4, gamma 1/2, dt 1 ms with "
                         "CBTiming halving), in FP64 and mixed precision, with and without the learned warm start "
                         "(0: skip)")
    ap.add_argument("--beat-tol", default="1e-12,1e-12,1e-14",
                    help="Newton abstol, rtol, stol of the beat window runs")
    ap.add_argument("--beat-runs", default="fp64_bj,fp64,mixed,mixed_warm",
                    help="comma list of fp64_bj (block Jacobi), fp64, mixed, mixed_warm (vertex-patch Schwarz)")
    ap.add_argument("--krylov-method", type=int, default=None,
                    help="Krylov solver of the beat window runs (0 GMRES, 1 BiCGStab with device-side scalars, "
                         "2 BiCGStab with host-side scalars; default: the library's)")
    ap.add_argument("--precond-reuse", type=int, default=None,
                    help="1: the beat window runs build the preconditioner once per Newton solve (default: the "
                         "library's)")
    ap.add_argument("--newton-stall", type=int, default=5,
                    help="Newton iterations of < 10 %% residual reduction in a row that end a step as failed, "
                         "for every run of the beat window and of --beat-full (0: run on to Solver.MaxIterations "
                         "as the reference does)")
    ap.add_argument("--beat-train", type=int, default=1,
                    help="training windows for the warm start (Tmax +-15 %%, activation +-10 ms, seeded)")
    ap.add_argument("--beat-cpu-steps", type=int, default=1,
                    help="Newmark steps of the same beat run by the CPU reference solver for the beat's CPU "
                         "baseline (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU baseline sample: repeat 20k-element chunks of the mesh for about this long")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-element", type=int, default=0,
                    help="profiling mode: N launches of the element kernel (default variant, or "
                         "--profile-variant) and of a known-size copy kernel, nothing else")
    ap.add_argument("--profile-variant", type=int, default=0,
                    help="element kernel variant of --profile-element (0: the precision's default)")
    ap.add_argument("--profile-assembly", type=int, default=0,
                    help="profiling mode: N passes of element kernel + deterministic assembly "
                         "(cmg_residual_tangent_device) and N launches of a known-size copy kernel, nothing else")
    ap.add_argument("--profile-beat", type=int, default=0,
                    help="profiling mode: the first N Newmark steps (1 ms) of the LV beat with the default solver")
    ap.add_argument("--sweep", action="store_true",
                    help="run the batched passive-parameter sweep (Guccione sets, quasi-static LV inflation) "
                         "instead of the element-kernel benchmark")
    ap.add_argument("--sweep-sets", type=int, default=64)
    ap.add_argument("--sweep-elements", type=int, default=50_000)
    ap.add_argument("--sweep-steps", type=int, default=20)
    ap.add_argument("--sweep-check", type=int, default=2, help="sets re-solved one at a time for the check")
    ap.add_argument("--sweep-check-from", type=int, default=0, help="first set of the check")
    ap.add_argument("--warm-start", action="store_true",
                    help="train the learned Newton warm start on jittered beats and compare a beat window "
                         "with and without it")
    ap.add_argument("--ws-elements", type=int, default=20_000)
    ap.add_argument("--ws-steps", type=int, default=150, help="Newmark steps (dt = 1 ms) of each beat window")
    ap.add_argument("--ws-train", type=int, default=5, help="training beats (Tmax +-15 %%, timing +-10 ms)")
    ap.add_argument("--ws-epochs", type=int, default=40)
    ap.add_argument("--ws-bf16", action="store_true", help="bfloat16 inference")
    ap.add_argument("--beat-full", action="store_true",
                    help="run whole beats (0 to --beat-end s, CBTiming step control) in FP64 and in mixed "
                         "precision and compare their displacements at every common time")
    ap.add_argument("--beat-elements", type=int, default=20_000)
    ap.add_argument("--beat-end", type=float, default=0.8)
    ap.add_argument("--beat-precisions", default="fp64,mixed", help="comma list of fp64, mixed")
    ap.add_argument("--beat-newmark", default="0.3025,0.6",
                    help="Newmark beta,gamma of the whole-beat runs (gamma > 1/2 damps the high-frequency motion "
                         "that stalls the undamped trapezoidal rule late in the beat; 0.25,0.5 is the trapezoidal rule)")
    ap.add_argument("--beat-rayleigh", default="0,0",
                    help="Rayleigh damping alpha (1/s), beta (s); the reference's LV settings use 500,0.01")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU: sharding the mesh over GPUs (element "
                                   "partitions with an interface-row exchange) is not implemented", "n_gpus": args.gpus}))
        sys.exit(2)

    from cardiomechanics_amd import materials as MT
    from cardiomechanics_amd import mesh as M
    from cardiomechanics_amd import solver as SV
    from cardiomechanics_amd.context import MechanicsContext
    from cardiomechanics_amd.hip import DeviceArray, Event

    prec = 1 if args.precision == "mixed" else 0
    if args.sweep:
        return sweep(args, prec)
    if args.warm_start:
        return warm_start(args, prec)
    if args.beat_full:
        return beat_full(args)
    t0 = time.time()
    mesh = M.lv_mesh(args.elements)
    E = mesh.n_elements
    log(f"[bench] mesh {mesh.name}: {E} T10 elements, {mesh.n_nodes} nodes ({time.time() - t0:.1f} s)")
    mats = [MT.lv_holzapfel()]
    L = np.ptp(mesh.coords, axis=0).max()
    x = mesh.coords + M.smooth_displacement(mesh.coords, L, 0.02, 1234)
    t_eval = 0.12
    xd = None

    def time_kernel(ctx, steps, warmup, tangent=1):
        stream = ctx.L.cmg_stream(ctx.h)
        for _ in range(warmup):
            ctx._check(ctx.L.cmg_element_kernel_device(ctx.h, xd.ptr, t_eval, tangent), "element kernel")
        ctx._check(ctx.L.cmg_synchronize(ctx.h), "sync")
        e0, e1 = Event(), Event()
        w0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            ctx._check(ctx.L.cmg_element_kernel_device(ctx.h, xd.ptr, t_eval, tangent), "element kernel")
        e1.record(stream)
        e1.synchronize()
        ctx._check(ctx.L.cmg_synchronize(ctx.h), "sync")
        return e0.elapsed_ms(e1) / steps, (time.perf_counter() - w0) / steps * 1e3

    if args.profile_assembly > 0:
        ctx = MechanicsContext(mesh, mats, precision=prec)
        xd = DeviceArray(x.ravel())
        fd_ = DeviceArray(np.zeros(ctx.n_dofs))
        vals = DeviceArray(np.zeros(ctx.nnzb * 9, np.float32 if prec else np.float64))
        ncopy = 64 * 1024 * 1024
        src = DeviceArray(np.ones(ncopy))
        dst = DeviceArray(np.zeros(ncopy))
        a0 = time.perf_counter()
        for _ in range(args.profile_assembly):
            ctx._check(ctx.L.cmg_residual_tangent_device(ctx.h, xd.ptr, t_eval, 1, fd_.ptr, vals.ptr), "rt")
        ctx._check(ctx.L.cmg_synchronize(ctx.h), "sync")
        pass_ms = (time.perf_counter() - a0) / args.profile_assembly * 1e3
        for _ in range(args.profile_assembly):
            ctx._check(ctx.L.cmg_copy_device(ctx.h, src.ptr, dst.ptr, ncopy), "copy")
        ctx._check(ctx.L.cmg_synchronize(ctx.h), "sync")
        print(json.dumps({"profile": "residual + tangent + assembly + copy", "elements": E, "dofs": ctx.n_dofs,
                          "blocks": ctx.nnzb, "pass_ms_wall": pass_ms, "copy_bytes_read": 8 * ncopy,
                          "copy_bytes_written": 8 * ncopy}))
        return
    if args.profile_beat > 0:
        ctx = MechanicsContext(mesh, mats, precision=prec)
        plug = SV.sinus_pulse(-16000.0, 0.0, 0.8)
        steps = []
        for k in range(1, args.profile_beat + 1):
            tk = k * 1e-3
            plug.apply(tk, 1e-3)
            ctx.set_pressures(plug.values(ctx.n_groups))
            rc, st = ctx.newmark_step(tk, 1e-3)
            steps.append({"t_ms": k, "rc": rc, "newton": st.newton_iterations, "krylov": st.krylov_iterations,
                          "wall_s": st.wall_seconds, "tangent_s": st.tangent_seconds, "krylov_s": st.krylov_seconds})
            log(f"[bench] profile step {k}: {steps[-1]}")
        print(json.dumps({"profile": "Newmark steps", "elements": E, "dofs": ctx.n_dofs, "steps": steps}))
        ctx.close()
        return
    if args.profile_element > 0:
        ctx = MechanicsContext(mesh, mats, precision=prec, element_variant=args.profile_variant)
        xd = DeviceArray(x.ravel())
        ncopy = 64 * 1024 * 1024
        src = DeviceArray(np.ones(ncopy))
        dst = DeviceArray(np.zeros(ncopy))
        ms, _ = time_kernel(ctx, args.profile_element, 1)
        for _ in range(args.profile_element):
            ctx._check(ctx.L.cmg_copy_device(ctx.h, src.ptr, dst.ptr, ncopy), "copy")
        ctx._check(ctx.L.cmg_synchronize(ctx.h), "sync")
        print(json.dumps({"profile": "element kernel + copy", "elements": E, "element_ms": ms,
                          "copy_bytes_read": 8 * ncopy, "copy_bytes_written": 8 * ncopy,
                          "bytes_per_element_algorithmic": BYTES_PER_ELEM[prec]}))
        return

    # kernel variants of the fused residual + tangent element kernel; the fastest is the headline
    variants = {}
    for var in (2, 4, 5, 6, 7, 8, 3, 1):
        ctx = MechanicsContext(mesh, mats, precision=prec, element_variant=var)
        if xd is None:
            xd = DeviceArray(x.ravel())
            log(f"[bench] context ready: {ctx.n_dofs} dofs, {ctx.nnzb} blocks ({time.time() - t0:.1f} s)")
        ms, wall_ms = time_kernel(ctx, max(3, args.steps // 4), args.warmup)
        variants[var] = ms
        log(f"[bench] variant {var}: {ms:.3f} ms/launch")
        ctx.close()
    best = min(variants, key=variants.get)
    ctx = MechanicsContext(mesh, mats, precision=prec, element_variant=best)
    kernel_ms, wall_ms = time_kernel(ctx, args.steps, args.warmup)
    value = E / (kernel_ms * 1e-3)
    log(f"[bench] element residual+tangent kernel (variant {best}): {kernel_ms:.3f} ms/launch, {value:.3e} elem/s "
        f"(wall {wall_ms:.3f} ms/step)")

    # the same pass with deterministic assembly of the residual vector and block matrix
    fd_ = DeviceArray(np.zeros(ctx.n_dofs))
    vals = DeviceArray(np.zeros(ctx.nnzb * 9, np.float32 if prec else np.float64))
    for _ in range(2):
        ctx._check(ctx.L.cmg_residual_tangent_device(ctx.h, xd.ptr, t_eval, 1, fd_.ptr, vals.ptr), "rt")
    ctx._check(ctx.L.cmg_synchronize(ctx.h), "sync")
    n_asm = max(3, args.steps // 5)
    a0 = time.perf_counter()
    for _ in range(n_asm):
        ctx._check(ctx.L.cmg_residual_tangent_device(ctx.h, xd.ptr, t_eval, 1, fd_.ptr, vals.ptr), "rt")
    ctx._check(ctx.L.cmg_synchronize(ctx.h), "sync")
    asm_ms = (time.perf_counter() - a0) / n_asm * 1e3
    log(f"[bench] residual+tangent+assembly: {asm_ms:.3f} ms/pass")

    bytes_launch = BYTES_PER_ELEM[prec] * E
    achieved = bytes_launch / (kernel_ms * 1e-3)
    traffic = TRAFFIC_PER_ELEM[prec] * E if TRAFFIC_PER_ELEM[prec] else None
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": traffic,
                "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch of this kernel on this mesh, "
                                  "calibrated on a copy kernel (profiles/); bytes per element times E, not "
                                  "re-measured in this run",
                "bytes_per_element": BYTES_PER_ELEM[prec], "kernel": f"k_residual_tangent variant {best}",
                "fp64_flops_per_s": FLOPS_PER_ELEM * E / (kernel_ms * 1e-3)}
    # residual only (the line-search / convergence-check evaluation): 1,788 B and 3,160 flops per element
    res_ms, _ = time_kernel(ctx, args.steps, args.warmup, tangent=0)
    residual_only = {"elem_per_s": E / (res_ms * 1e-3), "ms": res_ms, "bytes_per_element": RES_BYTES_PER_ELEM,
                     "achieved_GBps": RES_BYTES_PER_ELEM * E / (res_ms * 1e-3) / 1e9,
                     "frac": RES_BYTES_PER_ELEM * E / (res_ms * 1e-3) / HBM_PEAK}
    log(f"[bench] residual-only kernel: {res_ms:.3f} ms/launch, {E / (res_ms * 1e-3):.3e} elem/s, "
        f"{residual_only['frac']:.2f} of 8 TB/s")

    beat = None
    if args.beat_ms > 0:
        beat = beat_window_runs(args, mesh, mats)

    cpu = None
    if not args.no_cpu:
        import oracle as O
        # the reference algorithm on the host cores: chunks of 20k elements of the same mesh and
        # state, cycling through the mesh until about --cpu-seconds have passed
        chunk = 20_000
        done = 0
        c0 = time.perf_counter()
        while True:
            lo = done % E
            hi = min(lo + chunk, E)
            sub = M.Mesh(coords=mesh.coords, elements=mesh.elements[lo:hi], bases=mesh.bases[lo:hi],
                         bc=mesh.bc, material=mesh.material[lo:hi], activation=mesh.activation[lo:hi])
            rc, Ke, st = O.tangent_batch(O.ElementData(sub), x, mats, t_eval, kind="fd", eps=1e-12)
            done += hi - lo
            c1 = time.perf_counter() - c0
            if c1 >= args.cpu_seconds:
                break
        cpu = {"value": done / c1, "unit": "elem/s", "cores": int(O.lib().orc_num_threads()), "kind": "port",
               "sample": f"{done} element evaluations ({c1:.1f} s) in 20k-element chunks of the same mesh and "
                         f"state; reference algorithm: residual + central-difference tangent (61 residual "
                         f"evaluations per element), FP64, OpenMP"}
        log(f"[bench] CPU baseline: {cpu['value']:.3e} elem/s on {cpu['cores']} threads ({c1:.1f} s)")

    out = {"metric": "T10 elem residual+tangent/s", "value": value, "unit": "elem/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": kernel_ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == 0 else "f64/f32",
           "data": "synthetic: generated truncated-ellipsoid LV, seeded smooth displacement (seed 1234)",
           "config": {"workload": "LV T10 residual+analytic tangent, Holzapfel-Ogden + active tension",
                      "elements": E, "nodes": mesh.n_nodes, "precision": args.precision,
                      "element_variant": best, "variant_ms": variants, "assembly_ms_per_pass": asm_ms,
                      "residual_only": residual_only},
           "roofline": roofline, "cpu_baseline": cpu}
    if beat:
        out["beat"] = beat
        if cpu is not None and beat.get("cpu_reference"):
            cpu["beat"] = beat.pop("cpu_reference")
    print(json.dumps(out))
    ctx.close()


def beat_window_runs(args, mesh, mats):
    """The 200k-element LV beat over [0, --beat-ms]: Newmark beta 1/4, gamma 1/2, dt 1 ms
    halved by CBTiming on failed steps (CBTiming.cpp:77-136), cavity pressure -16 kPa
    sin^2 over 0.8 s, active tension 60 kPa sin^2 from the apex-to-base activation, all
    runs at the same Newton tolerance.  Runs: FP64 with block Jacobi (the first GPU path),
    FP64 and mixed precision with vertex-patch Schwarz, and mixed precision with the
    learned warm start (trained on --beat-train jittered windows).  Each run's nodes are
    compared with the FP64 run's at every time both accepted.  The CPU reference solver
    (central-difference tangent, OpenMP) runs the first --beat-cpu-steps steps."""
    from cardiomechanics_amd import materials as MT
    from cardiomechanics_amd import mesh as M
    from cardiomechanics_amd import solver as SV
    from cardiomechanics_amd.context import MechanicsContext
    from cardiomechanics_amd.warmstart import WarmStartNet

    atol, rtol, stol = (float(x) for x in args.beat_tol.split(","))
    T = args.beat_ms * 1e-3
    dt0 = 1e-3
    X0 = mesh.coords.ravel()
    labels = [r.strip() for r in args.beat_runs.split(",") if r.strip()]
    conf = {"fp64_bj": (0, 1), "fp64": (0, 2), "mixed": (1, 2), "mixed_warm": (1, 2)}
    for lab in labels:
        if lab not in conf:
            raise SystemExit(f"--beat-runs: unknown run {lab}")

    def window(prec, pc, mat, act=None, weights=None, record=False, keep=True):
        msh = mesh if act is None else M.Mesh(coords=mesh.coords, elements=mesh.elements, bases=mesh.bases,
                                              bc=mesh.bc, material=mesh.material, activation=act,
                                              surfaces=mesh.surfaces, surface_id=mesh.surface_id)
        plug = SV.sinus_pulse(-16000.0, 0.0, 0.8)
        timing = SV.Timing(0.0, T, dt0)
        res = dict(newton=0, krylov=0, accepted=0, halved=0, failed_attempts=0, kept=0, proposed=0, states={},
                   samples=[])
        kw = {} if args.krylov_method is None else {"krylov_method": args.krylov_method}
        if args.precond_reuse is not None:
            kw["precond_reuse"] = args.precond_reuse
        if args.newton_stall is not None:
            kw["newton_stall"] = args.newton_stall
        with MechanicsContext(msh, [mat], precision=prec, preconditioner=pc, newton_atol=atol, newton_rtol=rtol,
                              newton_stol=stol, newmark_beta=0.25, newmark_gamma=0.5, **kw) as ctx:
            net = WarmStartNet(ctx, seed=5) if (weights is not None or record) else None
            if weights is not None:
                net.set_weights(weights)
            t, p_prev, full = 0.0, 0.0, 0
            t0 = time.perf_counter()
            while t < T - 1e-12:
                dt = timing.dt
                plug.apply(t + dt, dt)
                p = plug.values(ctx.n_groups)
                ctx.set_pressures(p)
                rate = (p[0] - p_prev) / dt
                ok_dt = abs(dt - dt0) < 1e-15
                if record and ok_dt and full >= 2:
                    net.record_before_step(t + dt, rate)
                if weights is not None and ok_dt and full >= 2:
                    net.propose(t + dt, rate, bf16=True)   # bfloat16 inference on the matrix cores
                    res["proposed"] += 1
                rc, st = ctx.newmark_step(t + dt, dt)
                res["newton"] += st.newton_iterations
                res["krylov"] += st.krylov_iterations
                if rc == 2:
                    t += dt
                    p_prev = p[0]
                    res["accepted"] += 1
                    res["halved"] += not ok_dt
                    res["kept"] += st.guess_source == 2
                    full = full + 1 if ok_dt else 0
                    if record and getattr(net, "_pending", None) is not None:
                        net.record_after_step()
                    if keep:
                        res["states"][round(t * 1e7)] = ctx.get_state()["nodes"] - X0
                else:
                    res["failed_attempts"] += 1
                    if record:
                        net._pending = None
                if not timing.compute_next(rc):
                    raise RuntimeError(f"beat window: minimum time step at t={t + dt}")
            res["wall"] = time.perf_counter() - t0
            if record:
                res["samples"] = net.samples
            if net is not None:
                net.close()
        return res

    runs = {}
    nominal = MT.lv_holzapfel()
    for lab in labels:
        prec, pc = conf[lab]
        weights = None
        train = None
        if lab == "mixed_warm":
            rng = np.random.default_rng(0)
            samples = []
            t0 = time.perf_counter()
            for b in range(args.beat_train):
                m = MT.lv_holzapfel(tension_max=60e3 * rng.uniform(0.85, 1.15))
                act = np.maximum(mesh.activation + rng.uniform(-0.01, 0.01), 0.0)
                r = window(1, 2, m, act=act, record=True, keep=False)
                samples += r["samples"]
                log(f"[beat] training window {b}: {r['accepted']} steps, {len(r['samples'])} samples, "
                    f"{r['wall']:.1f} s")
            t_rec = time.perf_counter() - t0
            with MechanicsContext(mesh, [nominal], precision=1) as tctx:
                net = WarmStartNet(tctx, seed=5)
                net.samples = samples
                t1 = time.perf_counter()
                hist = net.train(epochs=40, lr=2e-3)
                weights = net.weights()
                net.close()
            train = {"windows": args.beat_train, "samples": len(samples), "record_s": t_rec,
                     "train_s": time.perf_counter() - t1, "loss_first": hist[0], "loss_last": hist[-1]}
        r = window(prec, pc, nominal, weights=weights)
        r["train"] = train
        runs[lab] = r
        log(f"[beat] {lab}: {r['accepted']} steps ({r['halved']} shorter than 1 ms, {r['failed_attempts']} failed "
            f"attempts), newton {r['newton']}, krylov {r['krylov']}, {r['wall']:.1f} s"
            + (f", learned guess kept {r['kept']}/{r['proposed']}" if weights is not None else ""))
    out = {"workload": f"LV beat window 0-{args.beat_ms:g} ms, Newmark beta 1/4 gamma 1/2, dt 1 ms with CBTiming "
                       f"halving, -16 kPa sin^2 cavity pressure, 60 kPa sin^2 active tension apex->base",
           "elements": mesh.n_elements, "newton_tolerance": [atol, rtol, stol],
           "newton_stall": args.newton_stall, "runs": {}}
    ref = runs.get("fp64")
    umax = max(float(np.abs(u).max()) for u in ref["states"].values()) if ref else None
    for lab, r in runs.items():
        rec = {k: r[k] for k in ("wall", "accepted", "halved", "failed_attempts", "newton", "krylov")}
        rec["precision"] = "mixed" if conf[lab][0] else "fp64"
        rec["preconditioner"] = "block Jacobi" if conf[lab][1] == 1 else "vertex-patch Schwarz"
        if r["train"]:
            rec["warm_start"] = dict(r["train"], kept=r["kept"], proposed=r["proposed"],
                                     inference="bf16 (v_mfma_f32_16x16x16_bf16), FP32 training")
        if ref is not None and lab != "fp64":
            common = sorted(set(ref["states"]) & set(r["states"]))
            if common:
                e_rel = [float(np.linalg.norm(r["states"][k] - ref["states"][k]) /
                               max(np.linalg.norm(ref["states"][k]), 1e-300)) for k in common]
                e_win = [float(np.abs(r["states"][k] - ref["states"][k]).max()) / umax for k in common]
                rec["vs_fp64"] = {"common_times": len(common), "max_rel_displacement_difference": max(e_rel),
                                  "last_rel_displacement_difference": e_rel[-1],
                                  "max_abs_difference_over_window_max_displacement": max(e_win)}
            if ref["wall"] > 0:
                rec["speedup_vs_fp64"] = ref["wall"] / r["wall"]
        out["runs"][lab] = rec
    if umax is not None:
        out["max_displacement_m"] = umax
    if "fp64_bj" in runs and "mixed_warm" in runs:
        out["speedup_mixed_warm_vs_fp64_bj"] = runs["fp64_bj"]["wall"] / runs["mixed_warm"]["wall"]
    # the CPU reference solver on the first steps of the same beat
    if not args.no_cpu and args.beat_cpu_steps > 0:
        from oracle.drivers import CpuReferenceSolver
        import oracle as O
        ref_cpu = CpuReferenceSolver(mesh, [nominal])
        plug = SV.sinus_pulse(-16000.0, 0.0, 0.8)
        walls, stats = [], []
        for k in range(1, args.beat_cpu_steps + 1):
            tk = k * dt0
            plug.apply(tk, dt0)
            ref_cpu.set_pressures(plug.values(1))
            rc, st = ref_cpu.newmark_step(tk, dt0, atol=atol, rtol=rtol, stol=stol)
            walls.append(st["wall_s"])
            stats.append(st)
            log(f"[beat] CPU reference step t={tk * 1e3:.0f} ms: rc {rc}, newton {st['newton']:.0f}, krylov "
                f"{st['krylov']:.0f}, {st['wall_s']:.1f} s (tangent {st['tangent_s']:.1f} s, linear "
                f"{st['linear_s']:.1f} s)")
        ref_cpu.close()
        out["cpu_reference"] = {
            "value": float(np.mean(walls)), "unit": "s per Newmark step", "cores": int(O.lib().orc_num_threads()),
            "kind": "port", "steps": args.beat_cpu_steps,
            "sample": f"the first {args.beat_cpu_steps} Newmark step(s) of the same beat (t = 1 ms ...), reference "
                      f"algorithm: central-difference element tangents (h = Solver.Epsilon), BiCGStab + block "
                      f"Jacobi in place of MUMPS LU, same Newton tolerance, OpenMP",
            "newton_krylov": [[s_["newton"], s_["krylov"]] for s_ in stats]}
    return out


def beat_full(args):
    """Whole LV beats (Newmark, dt = 1 ms, CBTiming halving on failed steps, cavity pressure
    -16 kPa sin(pi t / 0.8 s), active tension 60 kPa sin^2 over 0.3 s from the apex-to-base
    activation) in FP64 and/or mixed precision at the reference's Solver.Precision; with
    both, the displacement difference at every time both runs accepted, relative to the
    FP64 run's largest displacement over the beat."""
    from cardiomechanics_amd import materials as MT
    from cardiomechanics_amd import mesh as M
    from cardiomechanics_amd import solver as SV
    from cardiomechanics_amd.context import MechanicsContext

    mesh = M.lv_mesh(args.beat_elements)
    mats = [MT.lv_holzapfel()]
    X0 = mesh.coords.ravel()
    runs = {}
    labels = [x.strip() for x in args.beat_precisions.split(",") if x.strip()]
    if not labels or any(x not in ("fp64", "mixed") for x in labels):
        raise SystemExit("--beat-precisions takes fp64 and/or mixed")
    alpha, beta_r = (float(x) for x in args.beat_rayleigh.split(","))
    nm_beta, nm_gamma = (float(x) for x in args.beat_newmark.split(","))
    for label, prec in (("fp64", 0), ("mixed", 1)):
        if label not in labels:
            continue
        ctx = MechanicsContext(mesh, mats, precision=prec, rayleigh_alpha=alpha, rayleigh_beta=beta_r,
                               newmark_beta=nm_beta, newmark_gamma=nm_gamma, newton_stall=args.newton_stall)
        states, stats, umax = {}, [], [0.0]

        def on(rec, st, ctx=ctx, states=states, stats=stats, label=label, umax=umax):
            u = ctx.get_state()["nodes"] - X0
            umax[0] = max(umax[0], float(np.abs(u).max()))
            if len(labels) == 2:                     # kept only for the comparison
                states[round(rec.time * 1e6)] = u
            stats.append((rec.dt, st.newton_iterations, st.krylov_iterations))
            if len(stats) % 50 == 0:
                log(f"[beat] {label} t={rec.time:.4f} s dt={rec.dt:.2e} newton={st.newton_iterations} "
                    f"krylov={st.krylov_iterations}")

        t0 = time.perf_counter()
        SV.run(ctx, "newmark", SV.Timing(0.0, args.beat_end, 1e-3), SV.sinus_pulse(-16000.0, 0.0, 0.8),
               on_step=on)
        wall = time.perf_counter() - t0
        ctx.close()
        runs[label] = (states, stats, wall, umax[0])
        log(f"[beat] {label}: {len(stats)} accepted steps ({sum(1 for s in stats if s[0] < 0.999e-3)} shorter "
            f"than 1 ms), newton {sum(s[1] for s in stats)}, krylov {sum(s[2] for s in stats)}, {wall:.1f} s")
    umax = runs[labels[0]][3]
    common, diff = [], None
    if len(runs) == 2:
        ref, mix = runs["fp64"][0], runs["mixed"][0]
        common = sorted(set(ref) & set(mix))
        diff = max(float(np.abs(mix[k] - ref[k]).max()) for k in common) / umax
    last = labels[-1]
    out = {"metric": "Newmark beat wall time", "unit": "s", "higher_is_better": False, "n_gpus": 1,
           "config": {"workload": "LV beat, Holzapfel-Ogden + active tension, Newmark dt 1 ms with halving",
                      "elements": mesh.n_elements, "dofs": 3 * mesh.n_nodes, "end_time_s": args.beat_end,
                      "rayleigh_alpha": alpha, "rayleigh_beta": beta_r, "newmark_beta": nm_beta,
                      "newmark_gamma": nm_gamma, "newton_stall": args.newton_stall},
           "value": runs[last][2],
           "dtype": "mixed (fp32 tangent and Krylov operator, fp64 residual)" if last == "mixed" else "f64",
           "wall_s": {k: v[2] for k, v in runs.items()},
           "accepted_steps": {k: len(v[1]) for k, v in runs.items()},
           "newton": {k: sum(s[1] for s in v[1]) for k, v in runs.items()},
           "krylov": {k: sum(s[2] for s in v[1]) for k, v in runs.items()},
           "common_times": len(common), "max_displacement_m": umax,
           "max_abs_difference_over_max_displacement": diff}
    print(json.dumps(out))


def sweep(args, prec):
    """Passive-parameter sweep: S Guccione sets on one LV mesh, quasi-static inflation
    0 -> 2 kPa in `sweep_steps` load steps, all sets in the same launches; a few sets are
    re-solved one at a time (same library, single-set path) for timing and checking.
    A load step that fails for any set is rolled back and redone as two half steps
    (CBTiming's halving, CBTiming.cpp:90-100), for the batch and the single sets alike;
    the final state at 2 kPa does not depend on the load path of this elastic problem."""
    from cardiomechanics_amd import materials as MT
    from cardiomechanics_amd import mesh as M
    from cardiomechanics_amd.context import MechanicsContext, ParameterBatch

    mesh = M.lv_mesh(args.sweep_elements)
    sets = MT.guccione_sweep(args.sweep_sets, seed=0)
    # the endocardium is a closed T3 cavity with a clamped rim: its assembled pressure tangent is
    # symmetric, so the batched conjugate gradient keeps it (quadratic Newton convergence)
    opts = dict(precision=prec, krylov_maxit=20000, newton_atol=1e-12, newton_rtol=1e-10, newton_stol=1e-14,
                newton_maxit=100)
    steps = args.sweep_steps
    P = -2000.0
    log(f"[sweep] {mesh.n_elements} T10 elements, {mesh.n_nodes} nodes, {len(sets)} Guccione sets, {steps} steps")

    def load_step(ctx, step, get, put, t, h, depth, counts):
        saved = get()
        ctx.set_pressures([P * t])
        rc, st = step(t, h)
        if rc == 2 or depth >= 5:
            return rc, st
        counts[0] += 1
        put(saved)
        rc, st = load_step(ctx, step, get, put, t - 0.5 * h, 0.5 * h, depth + 1, counts)
        if rc != 2:
            return rc, st
        return load_step(ctx, step, get, put, t, 0.5 * h, depth + 1, counts)

    failed_sets = []
    with MechanicsContext(mesh, [sets[0]], **opts) as ctx:
        with ParameterBatch(ctx, [[m] for m in sets]) as batch:
            t0 = time.perf_counter()
            its = []
            halved = [0]
            for k in range(1, steps + 1):
                rc, st = load_step(ctx, batch.static_step, batch.nodes, batch.set_nodes, k / steps, 1.0 / steps, 0,
                                   halved)
                its.append((max(s.newton_iterations for s in st), max(s.krylov_iterations for s in st)))
                log(f"[sweep] step {k}: rc={rc} max newton {its[-1][0]} max cg {its[-1][1]} halvings {halved[0]} "
                    f"{time.perf_counter() - t0:.1f} s")
                if rc != 2:
                    failed_sets = [i for i, s in enumerate(st) if s.converged_reason <= 0]
                    log(f"[sweep] sets that did not converge: {failed_sets}")
                    break
            t_batch = time.perf_counter() - t0
            xb = batch.nodes()
            info = batch.info()
    t_single = []
    errs = []
    check = list(range(args.sweep_check_from, min(args.sweep_check_from + args.sweep_check, len(sets))))
    for i in check:
        with MechanicsContext(mesh, [sets[i]], **opts) as one:
            t0 = time.perf_counter()
            h1 = [0]
            for k in range(1, steps + 1):
                rc, st = load_step(one, one.static_step, lambda: one.get_state()["nodes"],
                                   lambda x: one.set_state(nodes=x), k / steps, 1.0 / steps, 0, h1)
                if rc != 2:
                    break
            t_single.append(time.perf_counter() - t0)
            g = one.get_state()["nodes"]
        errs.append(float(np.linalg.norm(xb[i] - g) / np.linalg.norm(g - mesh.coords.ravel())))
        log(f"[sweep] set {i} alone: rc {rc}, {t_single[-1]:.1f} s, halvings {h1[0]}, "
            f"rel. difference to the batch {errs[-1]:.2e}")
    seq = float(np.mean(t_single)) * len(sets) if t_single else None
    out = {"metric": "passive-parameter sweep wall time", "value": t_batch, "unit": "s", "n_gpus": 1,
           "higher_is_better": False, "dtype": "f64" if prec == 0 else "f64/f32",
           "config": {"workload": "LV quasi-static inflation 0-2 kPa, Guccione parameter sweep",
                      "elements": mesh.n_elements, "sets": len(sets), "steps": steps},
           "batch_seconds": t_batch, "halvings": halved[0], "failed_sets": failed_sets,
           "sequential_seconds_extrapolated": seq, "speedup_vs_sequential": (seq / t_batch) if seq else None,
           "checked_sets": [check[0], check[-1]] if check else [], "rel_difference_vs_single": errs,
           "max_rel_difference_vs_single": max(errs) if errs else None,
           "autotuner": info, "max_newton_cg_per_step": its}
    print(json.dumps(out))


def warm_start(args, prec):
    """Learned warm start on a reduced version of the warm-start workload: training beats
    with Tmax jittered by +-15 % and activation shifted by +-10 ms (seeded), then the
    nominal beat window with and without the network at the same Newton tolerance."""
    from cardiomechanics_amd import materials as MT
    from cardiomechanics_amd import mesh as M
    from cardiomechanics_amd import solver as SV
    from cardiomechanics_amd.context import MechanicsContext
    from cardiomechanics_amd.warmstart import WarmStartNet

    mesh = M.lv_mesh(args.ws_elements)
    dt, steps = 1e-3, args.ws_steps
    tol = dict(precision=prec, newton_atol=1e-8, newton_rtol=1e-8, newton_stol=1e-8, newton_maxit=100)
    log(f"[warm start] {mesh.n_elements} T10 elements, {steps} steps of {dt * 1e3:.0f} ms, "
        f"{args.ws_train} training beats")

    def window(m, act, weights=None, record=False):
        msh = M.Mesh(coords=mesh.coords, elements=mesh.elements, bases=mesh.bases, bc=mesh.bc,
                     material=mesh.material, activation=act, surfaces=mesh.surfaces, surface_id=mesh.surface_id)
        plug = SV.sinus_pulse(-16000.0, 0.0, 0.8)
        newton = krylov = kept = halved = 0
        wall = 0.0
        nodes = []
        halved_at = []

        def advance(ctx, t_end, h, depth):
            # a failed step halves dt (CBTiming, CBTiming.cpp:90-100): two half steps
            plug.apply(t_end, h)
            ctx.set_pressures(plug.values(ctx.n_groups))
            rc, st = ctx.newmark_step(t_end, h)
            if rc == 2:
                return [st]
            if depth >= 5:
                raise RuntimeError(f"Newmark step at t={t_end} failed at dt={h}: {st.as_dict()}")
            return advance(ctx, t_end - 0.5 * h, 0.5 * h, depth + 1) + advance(ctx, t_end, 0.5 * h, depth + 1)

        with MechanicsContext(msh, [m], **tol) as ctx:
            net = WarmStartNet(ctx, seed=5)
            if weights is not None:
                net.set_weights(weights)
            p_prev = 0.0
            for k in range(1, steps + 1):
                t = k * dt
                plug.apply(t, dt)
                p = plug.values(ctx.n_groups)
                ctx.set_pressures(p)
                rate = (p[0] - p_prev) / dt
                p_prev = p[0]
                if record:
                    net.record_before_step(t, rate)
                t0 = time.perf_counter()
                if weights is not None and k > 2:
                    net.propose(t, rate, bf16=args.ws_bf16)
                sts = advance(ctx, t, dt, 0)
                wall += time.perf_counter() - t0
                if record:
                    if len(sts) == 1:
                        net.record_after_step()
                    else:   # the increment history now holds a partial step: not a training pair
                        net._pending = None
                halved += len(sts) > 1
                halved_at.append(len(sts) > 1)
                newton += sum(st.newton_iterations for st in sts)
                krylov += sum(st.krylov_iterations for st in sts)
                kept += sts[0].guess_source == 2
                nodes.append(ctx.get_state()["nodes"])
            samples = net.samples
            net.close()
        return dict(newton=newton, krylov=krylov, wall=wall, kept=kept, halved=halved, halved_at=halved_at,
                    nodes=nodes, samples=samples)

    rng = np.random.default_rng(0)
    samples = []
    t0 = time.perf_counter()
    for b in range(args.ws_train):
        m = MT.lv_holzapfel(tension_max=60e3 * rng.uniform(0.85, 1.15))
        # timing jitter: activation times shifted by up to +-10 ms, clipped at t = 0 so no
        # element starts its twitch before the beat does
        r = window(m, np.maximum(mesh.activation + rng.uniform(-0.01, 0.01), 0.0), record=True)
        samples += r["samples"]
        log(f"[warm start] training beat {b}: newton {r['newton']} krylov {r['krylov']} halved {r['halved']} "
            f"{r['wall']:.1f} s")
    t_rec = time.perf_counter() - t0
    with MechanicsContext(mesh, [MT.lv_holzapfel()], **tol) as tctx:
        net = WarmStartNet(tctx, seed=5)
        net.samples = samples
        t0 = time.perf_counter()
        hist = net.train(epochs=args.ws_epochs, lr=2e-3, log=log)
        t_train = time.perf_counter() - t0
        weights = net.weights()
        net.close()
    plain = window(MT.lv_holzapfel(), mesh.activation)
    warm = window(MT.lv_holzapfel(), mesh.activation, weights=weights)
    # same time discretisation only up to the first step either run had to split
    same = 0
    while same < steps and not plain["halved_at"][same] and not warm["halved_at"][same]:
        same += 1
    errs = [float(np.linalg.norm(a - b) / max(np.linalg.norm(a - mesh.coords.ravel()), 1e-300))
            for a, b in zip(plain["nodes"][:same], warm["nodes"][:same])] or [0.0]
    log(f"[warm start] plain: newton {plain['newton']} krylov {plain['krylov']} {plain['wall']:.2f} s; "
        f"warm: newton {warm['newton']} krylov {warm['krylov']} {warm['wall']:.2f} s, "
        f"guess kept {warm['kept']}/{steps - 2}; max rel. difference over the first {same} steps "
        f"(no step split in either run) {max(errs):.2e}")
    out = {"metric": "learned warm start: Newmark window wall time", "value": warm["wall"], "unit": "s",
           "n_gpus": 1, "higher_is_better": False, "dtype": "f64 (network f32" + ("/bf16" if args.ws_bf16 else "") + ")",
           "config": {"workload": "LV Newmark beat window, Holzapfel-Ogden + active tension", "elements": mesh.n_elements,
                      "steps": steps, "dt": dt, "training_beats": args.ws_train, "epochs": args.ws_epochs},
           "plain": {k: plain[k] for k in ("newton", "krylov", "wall", "halved")},
           "warm": {k: warm[k] for k in ("newton", "krylov", "wall", "kept", "halved")},
           "speedup": plain["wall"] / warm["wall"], "steps_compared": same, "max_rel_difference": max(errs),
           "training": {"record_seconds": t_rec, "train_seconds": t_train, "loss_first": hist[0],
                        "loss_last": hist[-1], "samples": len(samples)}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
