#!/usr/bin/env python3
"""Benchmark of the dense BEM hot path on one MI355X.

Workload (one "step"): mixed-precision (FP32 far-field / FP64 near-field and singular,
complex64 output) dense Galerkin assembly of the Helmholtz single-layer and double-layer
operators, P1 x P1 continuous, wavenumber k = 4, on the level-5 icosahedral unit sphere
(20,480 flat triangles, 10,242 vertices): two 10,242 x 10,242 complex64 matrices in HBM.
Reported value: matrix entries per second (2 * 10,242^2 per step). Outside the timed region
the same operators are assembled in FP64 to report the relative Frobenius error of the
mixed result, the FP64 throughput, the potential evaluation of a seeded complex Gaussian P1
density at 16,384 points on the sphere of radius 2, kernel-level roofline figures (HIP
events on the launch stream) and the CPU baseline (FP64 oracle on the host cores, row sample).

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "bempp-cl_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "BEM matrix entries/s at ≤1e-6 rel err vs FP64 (mixed vs FP64 GPU); fwd+bwd s"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# algorithmic FLOP per kernel evaluation (add/sub/mul 1, FMA 2, sqrt/rsqrt/rcp/sin/cos/exp 1)
FLOP_PER_EVAL = {"helmholtz_single_layer": 35, "helmholtz_double_layer": 45}
FLOP_PER_POTENTIAL_EVAL = 23


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--k", type=float, default=4.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target duration of the CPU sample")
    ap.add_argument("--no-reference", action="store_true", help="skip the FP64 comparison (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device to device (reference dispatch for HBM byte counters)")
    ap.add_argument("--no-sweep", action="store_true", help="skip the frequency-sweep measurement")
    ap.add_argument("--sweep-level", type=int, default=4)
    ap.add_argument("--sweep-reps", type=int, default=3)
    return ap.parse_args()


def frequency_sweep_bench(level, reps):
    """Helmholtz single layer, DP0 x DP0, 32 wavenumbers linspace(1, 16, 32), one persistent launch.

    Mixed sweep (complex64) timed with HIP events on the launch stream; every matrix compared
    with the FP64 sweep (which equals the FP64 single-wavenumber assembly to ~1e-16).
    """
    import numpy as np
    import torch

    import bempp_amd.api as bempp
    from bempp_amd.api.operators.boundary import helmholtz
    from bempp_amd.core.dense_assembler import hip_dense_assemble
    from bempp_amd.core.frequency_sweep import hip_kbatch_assemble
    from bempp_amd.core.singular_assembler import singular_rule

    grid = bempp.shapes.icosphere(level)
    sp = bempp.function_space(grid, "DP", 0)
    F = grid.number_of_elements
    ks = np.linspace(1.0, 16.0, 32)
    params = bempp.GLOBAL_PARAMETERS
    n_sing = singular_rule(grid, params.quadrature.singular, sp.support, sp.support).host.number_of_pairs
    n_geom = (F * F - n_sing) * 36
    out = hip_kbatch_assemble(sp, sp, ks, "single", parameters=params)  # tunes on first use
    torch.cuda.synchronize()
    ev = [{key: torch.cuda.Event(enable_timing=True) for key in ("regular_start", "regular_end", "end")}
          for _ in range(reps)]
    t0 = time.perf_counter()
    for e in ev:
        hip_kbatch_assemble(sp, sp, ks, "single", parameters=params, out=out, events=e)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    reg_ms = float(np.mean([e["regular_start"].elapsed_time(e["regular_end"]) for e in ev]))
    tot_ms = float(np.mean([e["regular_start"].elapsed_time(e["end"]) for e in ev]))
    ref = hip_kbatch_assemble(sp, sp, ks, "double", parameters=params, config=(16, 4))
    torch.cuda.synchronize()
    errs = [rel_fro(out[m], ref[m]) for m in range(len(ks))]
    del ref
    # the plain FP64 path: one single-wavenumber assembly per k
    one = torch.empty((F, F), dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k in ks:
        d = helmholtz.single_layer(sp, sp, sp, k, precision="double").descriptor
        hip_dense_assemble(sp, sp, params, d, out=one, accumulate=False)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t
    del one, out
    entries = len(ks) * F * F
    flops = n_geom * (10 + 9 * len(ks))
    return {
        "workload": f"Helmholtz single layer DP0xDP0, icosphere level {level} ({F} triangles), "
                    f"{len(ks)} wavenumbers linspace(1,16,32), mixed -> complex64",
        "entries_per_s": entries / (wall),
        "ms": 1e3 * wall,
        "regular_kernel_ms": reg_ms,
        "near_singular_ms": tot_ms - reg_ms,
        "max_rel_err_vs_fp64": float(max(errs)),
        "fp64_single_k_total_ms": 1e3 * t64,
        "speedup_vs_fp64_single_k": t64 / wall,
        "regular_algorithmic_tflops": flops / (reg_ms * 1e-3) / 1e12,
        "flop_per_eval": "10 + 9 per wavenumber",
        "regular_write_gbs": entries * 8 / (reg_ms * 1e-3) / 1e9,
        "note": "the mixed kernel expands e^{ik(r-r0)} in moments of r - r0, so the algorithmic "
                "FLOP rate counts the work of direct evaluation",
    }


def traffic_from_profiles(kernel_key):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), if recorded."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    entry = table.get(kernel_key)
    return None if entry is None else float(entry["hbm_bytes_per_launch"])


def rel_fro(a, b):
    import torch

    return float(torch.linalg.vector_norm(a.to(b.dtype) - b) / torch.linalg.vector_norm(b))


def cpu_baseline(grid, space, k, n_entries_step, target_seconds):
    """Time the FP64 oracle (restated reference algorithm) on a block of test elements."""
    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0))
    F = grid.number_of_elements

    def timed(op, rows, parts):
        t = time.perf_counter()
        O.assemble(op, grid.vertices, grid.elements, space, space, k, parts=parts, test_range=(0, rows),
                   nthreads=cores)
        return time.perf_counter() - t

    # two-point calibration separates the fixed cost (result allocation, colouring) from the
    # per-row cost, then one block sized for ~target_seconds is timed per operator
    ta = timed("helmholtz_single_layer", 64, 1)
    tb = timed("helmholtz_single_layer", 192, 1)
    per_row = max((tb - ta) / 128, 1e-6)
    fixed = max(ta - 64 * per_row, 0.0)
    n1 = int(min(F, max(192, target_seconds / 2.0 / per_row)))
    est, sampled = 0.0, 0.0
    for op in ("helmholtz_single_layer", "helmholtz_double_layer"):
        t_reg = timed(op, n1, 1)
        sampled += t_reg
        ns = max(1, n1 // 8)
        t_sing = timed(op, ns, 2)
        est += fixed + (t_reg - fixed) * F / n1 + max(t_sing - fixed, 0.0) * F / ns
    return {
        "value": n_entries_step / est,
        "unit": "matrix entries/s",
        "cores": cores,
        "kind": "port",
        "sample": f"FP64 oracle (C/OpenMP restatement of the reference Numba algorithm), Helmholtz SL+DL P1: "
                  f"regular pairs of test elements [0,{n1}) of {F} ({sampled:.1f} s) and singular pairs of "
                  f"[0,{max(1, n1 // 8)}), per-row cost extrapolated linearly to all rows; estimated full step "
                  f"{est:.0f} s",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import bempp_amd.api as bempp
    from bempp_amd.api.operators.boundary import helmholtz
    from bempp_amd.api.operators.potential import helmholtz as hpot
    from bempp_amd.core import hip_backend as hb
    from bempp_amd.core.dense_assembler import hip_dense_assemble
    from bempp_amd.core.singular_assembler import singular_rule

    torch.cuda.set_device(0)
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    k = args.k
    t = time.perf_counter()
    grid = bempp.shapes.icosphere(args.level)
    space = bempp.function_space(grid, "P", 1)
    F, N = grid.number_of_elements, space.global_dof_count
    log(f"grid level {args.level}: {F} triangles, {N} dofs ({time.perf_counter() - t:.1f} s setup)")
    params = bempp.GLOBAL_PARAMETERS
    ops = {
        "helmholtz_single_layer": (helmholtz.single_layer, "single"),
        "helmholtz_double_layer": (helmholtz.double_layer, "single"),
    }
    desc = {name: f(space, space, space, k, precision="single").descriptor for name, (f, _) in ops.items()}
    desc64 = {name: f(space, space, space, k, precision="double").descriptor for name, (f, _) in ops.items()}
    srule = singular_rule(grid, params.quadrature.singular, space.support, space.support)
    n_sing = srule.host.number_of_pairs
    n_regular_pairs = F * F - n_sing
    n_evals = n_regular_pairs * 36
    outs = {name: torch.empty((N, N), dtype=torch.complex64, device="cuda") for name in ops}

    def step(events=None):
        for name in ops:
            hip_dense_assemble(space, space, params, desc[name], out=outs[name], accumulate=False,
                               events=None if events is None else events[name])

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [{name: {key: torch.cuda.Event(enable_timing=True) for key in ("regular_start", "regular_end", "singular_end")}
           for name in ops} for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(ev[s])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = 1e3 * (t1 - t0) / args.steps
    entries_per_step = 2 * N * N
    value = entries_per_step / ((t1 - t0) / args.steps)
    kern = {}
    for name in ops:
        reg = np.mean([e[name]["regular_start"].elapsed_time(e[name]["regular_end"]) for e in ev])
        sing = np.mean([e[name]["regular_end"].elapsed_time(e[name]["singular_end"]) for e in ev])
        flops = n_evals * FLOP_PER_EVAL[name]
        kern[name] = {"regular_ms": float(reg), "singular_ms": float(sing),
                      "regular_tflops": flops / (reg * 1e-3) / 1e12}
    log(f"mixed step {ms_per_step:.2f} ms, {value / 1e9:.2f} G entries/s; kernels {kern}")

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "matrix entries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "Helmholtz single+double layer, P1xP1, icosphere level 5 (20480 triangles, 10242 vertices), "
                        "k=4.0, mixed FP32/FP64 -> complex64, dense in HBM",
            "triangles": F,
            "dofs": N,
            "wavenumber": k,
            "quadrature": "regular order 4 (6 points), singular order 4",
            "regular_pairs": n_regular_pairs,
            "singular_pairs": n_sing,
        },
    }
    name0 = "helmholtz_single_layer"
    sl = kern[name0]
    result["roofline"] = {
        "kernel": f"regular_kernel<{name0}, P1xP1, mixed>",
        "bound": "valu",
        "achieved": sl["regular_tflops"],
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": sl["regular_tflops"] / PEAK_FP32_TFLOPS,
        "traffic": traffic_from_profiles("regular_kernel_helmholtz_single_layer_p1_level5_mixed"),
        "algorithmic_flop_per_launch": n_evals * FLOP_PER_EVAL[name0],
        "flop_per_eval": FLOP_PER_EVAL[name0],
        "evals_per_launch": n_evals,
        "launch_ms": sl["regular_ms"],
    }
    result["kernels"] = kern

    if not args.no_reference:
        errs, t64 = {}, 0.0
        for name, (f, _) in ops.items():
            torch.cuda.synchronize()
            t = time.perf_counter()
            a64 = hip_dense_assemble(space, space, params, desc64[name])
            torch.cuda.synchronize()
            t64 += time.perf_counter() - t
            errs[name] = rel_fro(outs[name], a64)
            del a64
        result["rel_err_vs_fp64"] = errs
        result["fp64_entries_per_s"] = entries_per_step / t64
        result["speedup_vs_fp64"] = value / result["fp64_entries_per_s"]
        log(f"fp64 step {1e3 * t64:.1f} ms, rel err {errs}")

        # the same step with Kahan-compensated FP32 sums (parameters.hip.kahan)
        from bempp_amd.api.utils.parameters import DefaultParameters

        pk = DefaultParameters()
        pk.hip.kahan = True
        kres = {}
        for name, (f, _) in ops.items():
            dk = f(space, space, space, k, precision="single", parameters=pk).descriptor
            hip_dense_assemble(space, space, pk, dk, out=outs[name], accumulate=False)
            torch.cuda.synchronize()
            t = time.perf_counter()
            hip_dense_assemble(space, space, pk, dk, out=outs[name], accumulate=False)
            torch.cuda.synchronize()
            ms = 1e3 * (time.perf_counter() - t)
            a64 = hip_dense_assemble(space, space, params, desc64[name])
            kres[name] = {"ms": ms, "rel_err_vs_fp64": rel_fro(outs[name], a64)}
            del a64
        result["kahan_variant"] = kres
        log(f"kahan variant {kres}")

        # potential: complex Gaussian P1 density at 16384 points on the radius-2 sphere
        rng = np.random.default_rng(0)
        c = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        g = np.random.default_rng(1).standard_normal((3, 16384))
        pts = 2.0 * g / np.linalg.norm(g, axis=0)
        fun = bempp.GridFunction(space, coefficients=c)
        pot32 = hpot.single_layer(space, pts, k, precision="single")
        pot64 = hpot.single_layer(space, pts, k, precision="double")
        u64 = pot64.evaluate(fun)
        pot32.evaluate(fun)
        torch.cuda.synchronize()
        t = time.perf_counter()
        reps = 5
        for _ in range(reps):
            u32 = pot32.evaluate(fun)
        t_pot = (time.perf_counter() - t) / reps
        result["potential"] = {
            "points": pts.shape[1],
            "sources": F * 6,
            "evals_per_s": pts.shape[1] / t_pot,
            "ms": 1e3 * t_pot,
            "tflops": pts.shape[1] * F * 6 * FLOP_PER_POTENTIAL_EVAL / t_pot / 1e12,
            "rel_err_vs_fp64": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
        }
        log(f"potential {result['potential']}")

    if not args.no_sweep:
        result["frequency_sweep"] = frequency_sweep_bench(args.sweep_level, args.sweep_reps)
        log(f"frequency sweep {result['frequency_sweep']}")

    if not args.no_cpu_baseline:
        log("timing the CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline(grid, space, k, entries_per_step, args.cpu_seconds)
        log(f"cpu baseline {result['cpu_baseline']}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
