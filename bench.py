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
    ap.add_argument("--beat-steps", type=int, default=2, help="Newmark steps of the beat to time (0: skip)")
    ap.add_argument("--krylov", default="gmres,bicgstab",
                    help="Krylov methods to time the beat steps with (comma list of gmres, bicgstab)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU baseline sample: repeat 20k-element chunks of the mesh for about this long")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-element", type=int, default=0,
                    help="profiling mode: N launches of the element kernel (default variant) and of a "
                         "known-size copy kernel, nothing else")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU sharding is a design note only "
                                   "(see the sharding section of the design document)", "n_gpus": args.gpus}))
        sys.exit(2)

    from cardiomechanics_amd import materials as MT
    from cardiomechanics_amd import mesh as M
    from cardiomechanics_amd import solver as SV
    from cardiomechanics_amd.context import MechanicsContext
    from cardiomechanics_amd.hip import DeviceArray, Event

    prec = 1 if args.precision == "mixed" else 0
    t0 = time.time()
    mesh = M.lv_mesh(args.elements)
    E = mesh.n_elements
    log(f"[bench] mesh {mesh.name}: {E} T10 elements, {mesh.n_nodes} nodes ({time.time() - t0:.1f} s)")
    mats = [MT.lv_holzapfel()]
    L = np.ptp(mesh.coords, axis=0).max()
    x = mesh.coords + M.smooth_displacement(mesh.coords, L, 0.02, 1234)
    t_eval = 0.12
    xd = None

    def time_kernel(ctx, steps, warmup):
        stream = ctx.L.cmg_stream(ctx.h)
        for _ in range(warmup):
            ctx._check(ctx.L.cmg_element_kernel_device(ctx.h, xd.ptr, t_eval, 1), "element kernel")
        ctx._check(ctx.L.cmg_synchronize(ctx.h), "sync")
        e0, e1 = Event(), Event()
        w0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            ctx._check(ctx.L.cmg_element_kernel_device(ctx.h, xd.ptr, t_eval, 1), "element kernel")
        e1.record(stream)
        e1.synchronize()
        ctx._check(ctx.L.cmg_synchronize(ctx.h), "sync")
        return e0.elapsed_ms(e1) / steps, (time.perf_counter() - w0) / steps * 1e3

    if args.profile_element > 0:
        ctx = MechanicsContext(mesh, mats, precision=prec)
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
    for var in (2, 3, 1):
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
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": None,
                "bytes_per_element": BYTES_PER_ELEM[prec], "kernel": f"k_residual_tangent variant {best}",
                "fp64_flops_per_s": FLOPS_PER_ELEM * E / (kernel_ms * 1e-3)}

    beat = None
    if args.beat_steps > 0:
        beat = {}
        for name in args.krylov.split(","):
            method = {"gmres": 0, "bicgstab": 1}[name.strip()]
            bctx = MechanicsContext(mesh, mats, precision=prec, element_variant=best, krylov_method=method)
            plug = SV.sinus_pulse(-16000.0, 0.0, 0.8)
            dt = 1e-3
            times = []
            its = []
            for k in range(1, args.beat_steps + 1):
                tk = k * dt
                plug.apply(tk, dt)
                bctx.set_pressures(plug.values(bctx.n_groups))
                b0 = time.perf_counter()
                rc, st = bctx.newmark_step(tk, dt)
                times.append(time.perf_counter() - b0)
                its.append((st.newton_iterations, st.krylov_iterations))
                log(f"[bench] {name} newmark step t={tk * 1e3:.0f} ms rc={rc} newton={st.newton_iterations} "
                    f"krylov={st.krylov_iterations} {times[-1]:.3f} s")
            bctx.close()
            beat[name] = {"steps": args.beat_steps, "dt": dt, "s_per_step": float(np.mean(times)),
                          "newton_krylov": its, "beat_800_steps_s_extrapolated": float(np.mean(times) * 800)}

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
                      "element_variant": best, "variant_ms": variants, "assembly_ms_per_pass": asm_ms},
           "roofline": roofline, "cpu_baseline": cpu}
    if beat:
        out["beat"] = beat
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
