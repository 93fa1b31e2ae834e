#!/usr/bin/env python3
"""Benchmark of the dense BEM hot path on one MI355X.

Workload (one "step"): mixed-precision (FP32 far-field / FP64 near-field and singular,
complex64 output) dense Galerkin assembly of the Helmholtz single-layer and double-layer
operators, P1 x P1 continuous, wavenumber k = 4, on the level-5 icosahedral unit sphere
(20,480 flat triangles, 10,242 vertices): two 10,242 x 10,242 complex64 matrices in HBM,
both assembled in one symmetric pass (every unordered element pair once; shared point-pair
geometry and e^{ikr} for the single layer, the double layer and the reversed pair).
Reported value: matrix entries per second (2 * 10,242^2 per step). Outside the timed region:
the pass over ordered pairs, the two operators assembled separately, the FP64 assemblies
(separate = the plain path, ordered fused, symmetric) with the relative Frobenius error of
the mixed result, a Kahan-summation variant, the
potential of a seeded complex Gaussian P1 density at 16,384 points on the sphere of radius 2,
the Laplace DP0 FP64 assembly at level 3, the frequency sweep (32 wavenumbers, DP0, level 4),
the Maxwell electric field operator (RWG x SNC, level 4, symmetric and ordered traversals) and
its potential at 65,536 points, the loss/gradient workload (level 6), kernel-level roofline
figures (HIP events on the launch stream) and the CPU baseline (FP64 oracle on the host cores,
row sample).

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--no-sweep] [--no-adjoint]
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
# FP32 VALU FLOP the symmetric mixed kernel executes per unordered point pair, counted from its
# gfx950 assembly (inner loop over one test point and the 6 trial points: 205 FMA (2 FLOP),
# 126 add/sub/mul, 24 sqrt/rcp/sin/cos -> 560 FLOP per 6 point pairs)
EXEC_FLOP_PER_SYM_POINT_PAIR = 560.0 / 6.0
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
    ap.add_argument("--no-efie", action="store_true", help="skip the Maxwell EFIE measurement")
    ap.add_argument("--no-c1", action="store_true", help="skip the Laplace DP0 level-3 FP64 measurement")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the loss/gradient measurement")
    ap.add_argument("--no-solve", action="store_true", help="skip the GMRES solve on the assembled operator")
    ap.add_argument("--no-hypersingular", action="store_true", help="skip the hypersingular assembly")
    ap.add_argument("--no-maxwell-extras", action="store_true",
                    help="skip the MFIE assembly, the magnetic potential and the far fields")
    ap.add_argument("--adjoint-level", type=int, default=6)
    ap.add_argument("--adjoint-k", type=float, default=8.0)
    ap.add_argument("--fd-vertices", type=int, default=32)
    ap.add_argument("--sweep-level", type=int, default=4)
    ap.add_argument("--sweep-reps", type=int, default=3)
    return ap.parse_args()


def _ev():
    import torch

    return torch.cuda.Event(enable_timing=True)


def adjoint_bench(level, k, fd_vertices, with_cpu=False, cpu_seconds=4.0):
    """Loss L = ||V c - b||^2 + sum_p |u(p)|^2 on a perturbed sphere and its gradient.

    V: DP0 Helmholtz single layer (mixed FP32/FP64, complex64 in HBM), u: its potential at 4096
    points on the radius-2 sphere. Forward and backward (torch.autograd; backward = adjoint
    kernels + V^H g) are timed on the device after one warm-up pass. The mixed loss and
    gradients are compared with an FP64 evaluation, and the FP64 adjoint gradient with central
    finite differences (h = 1e-6) of the FP64 loss on `fd_vertices` random vertices (all three
    coordinates), each by re-assembling only the rows and columns of the elements at the
    vertex (dL from the changed entries, so that the large constant part of L cancels exactly).
    """
    import numpy as np
    import torch

    import bempp_amd.api as bempp
    from bempp_amd.api.differentiable import (SingleLayerProblem, _device_rule, single_layer_apply,
                                              single_layer_potential)
    from bempp_amd.api.grid.shapes import perturbed_icosphere
    from bempp_amd.api.space.space import function_space
    from bempp_amd.core import clear_caches
    from bempp_amd.core import hip_backend as hb
    from bempp_amd.core.dense_assembler import hip_dense_assemble
    from bempp_amd.core.singular_assembler import singular_rule

    grid = perturbed_icosphere(level)
    F, NV = grid.number_of_elements, grid.number_of_vertices
    rng5, rng6, rng7 = np.random.default_rng(5), np.random.default_rng(6), np.random.default_rng(7)
    pts = rng5.standard_normal((4096, 3))
    pts = 2.0 * pts / np.linalg.norm(pts, axis=1)[:, None]
    c0 = (rng6.standard_normal(F) + 1j * rng6.standard_normal(F)) / np.sqrt(2)
    b0 = (rng6.standard_normal(F) + 1j * rng6.standard_normal(F)) / np.sqrt(2)
    dev = hb.device()
    pts_t = torch.from_numpy(pts).to(dev)
    b = torch.from_numpy(b0).to(dev)
    v_host = np.ascontiguousarray(grid.vertices.T)
    params = bempp.GLOBAL_PARAMETERS
    prob = SingleLayerProblem(grid.elements, k, precision="single", parameters=params)

    def run():
        v = torch.from_numpy(v_host).to(dev).requires_grad_(True)
        c = torch.from_numpy(c0).to(dev).requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        y = single_layer_apply(v, c, prob)
        u = single_layer_potential(v, c, pts_t, prob)
        L = (y - b).abs().pow(2).sum() + u.abs().pow(2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return float(L.detach()), v.grad.detach(), c.grad.detach(), t1 - t0, t2 - t1

    run()  # warm-up: host set-up (near-field candidates, singular pairs), code objects
    L32, gv32, gc32, t_fwd, t_bwd = run()
    log(f"adjoint L{level}: forward {t_fwd:.3f} s, backward {t_bwd:.3f} s")

    # the forward's regular section (events on the launch stream): the upper-triangle kernel, which
    # also stores the mirror image of every entry
    space32 = prob.space(prob.grid(v_host))
    fe = {key: _ev() for key in ("regular_start", "regular_end", "singular_end")}
    A32 = hip_dense_assemble(space32, space32, params, prob.operator(space32).descriptor, events=fe)
    torch.cuda.synchronize()
    fwd_section_ms = fe["regular_start"].elapsed_time(fe["regular_end"])
    del A32
    torch.cuda.empty_cache()

    # kernel-level timing of the regular adjoint launch (events on the launch stream)
    space = prob.space(prob.grid(v_host))
    dsp = hb.device_space(space)
    rule = _device_rule(params)
    g_y = torch.from_numpy(c0).to(dev)  # any cotangent: the kernel work does not depend on it
    cvec = torch.from_numpy(c0).to(dev)
    grad = torch.zeros((NV, 3), dtype=torch.float64, device=dev)
    e0, e1 = _ev(), _ev()
    e0.record()
    hb.adjoint_single_layer(dsp, rule, k, hb.PRECISION_MIXED, 2.0, g_y, cvec, grad)
    e1.record()
    torch.cuda.synchronize()
    adj_ms = e0.elapsed_time(e1)
    n_sing = singular_rule(grid, params.quadrature.singular, space.support, space.support).host.number_of_pairs
    n_evals = (F * F - n_sing) * 36
    del grad

    # FP64 reference: loss and gradients (manual forward/backward, keeping V for the FD check)
    sp64 = function_space(grid, "DP", 0)
    from bempp_amd.api.operators.boundary import helmholtz
    from bempp_amd.api.operators.potential import helmholtz as hpot
    from bempp_amd.core.dense_potential_assembler import hip_potential_assembler

    A64 = hip_dense_assemble(sp64, sp64, params, helmholtz.single_layer(sp64, sp64, sp64, k, precision="double").descriptor)
    c = torch.from_numpy(c0).to(dev)
    r = hb.dense_matvec(A64, c) - b
    pot_desc = hpot.single_layer(sp64, pts.T, k, precision="double").descriptor
    u64, _ = hip_potential_assembler(sp64, pot_desc, pts.T, params)(c, as_tensor=True)
    u64 = u64[0]
    L64 = float((r.abs().pow(2).sum() + u64.abs().pow(2).sum()).item())
    gc64 = hb.dense_matvec(A64, (2 * r).contiguous(), trans=2)
    gv64 = torch.zeros((NV, 3), dtype=torch.float64, device=dev)
    cand = None
    srule = singular_rule(grid, params.quadrature.singular, sp64.support, sp64.support)
    dsp64 = hb.device_space(sp64)
    hb.adjoint_single_layer(dsp64, rule, k, hb.PRECISION_FP64, 2.0, (2 * r).contiguous(), c, gv64, singular=srule)
    hb.adjoint_potential_single_layer(dsp64, rule, k, c, pts_t, (2 * u64).contiguous(), grad_coefficients=gc64,
                                      grad_vertices=gv64)
    torch.cuda.synchronize()
    res = {
        "workload": f"Helmholtz single layer DP0, perturbed icosphere level {level} ({F} triangles, {NV} vertices), "
                    f"k={k}, loss ||Vc-b||^2 + sum|u|^2 (u: potential at 4096 points, radius 2), mixed FP32/FP64",
        "forward_s": t_fwd,
        "backward_s": t_bwd,
        "backward_over_forward": t_bwd / t_fwd,
        "adjoint_regular_kernel_ms": adj_ms,
        "adjoint_algorithmic_tflops": n_evals * 51 / (adj_ms * 1e-3) / 1e12,
        "adjoint_evals": n_evals,
        "forward_regular_section_ms": fwd_section_ms,
        "roofline": roofline("regular_kernel<helmholtz_single_layer, DP0, mixed, upper triangle and its mirror "
                             "image>", "valu",
                             n_evals * 19 / (fwd_section_ms * 1e-3) / 1e12, PEAK_FP32_TFLOPS,
                             "TFLOP/s", traffic_key="regular_kernel_helmholtz_sl_dp0_upper_level6_mixed",
                             evals_per_launch=n_evals, flop_per_eval=19, launch_ms=fwd_section_ms),
        "roofline_backward": roofline("adjoint_far_kernel<32, true> (unordered far pairs, mixed)", "valu",
                                      n_evals * 51 / (adj_ms * 1e-3) / 1e12, PEAK_FP32_TFLOPS, "TFLOP/s",
                                      traffic_key="adjoint_far_kernel_dp0_level6_mixed",
                                      evals_per_launch=n_evals, flop_per_eval=51, launch_ms=adj_ms),
        "loss_rel_err_vs_fp64": abs(L32 - L64) / abs(L64),
        "grad_vertices_rel_err_vs_fp64": float(torch.linalg.norm(gv32.to(dev) - gv64) / torch.linalg.norm(gv64)),
        "grad_c_rel_err_vs_fp64": float(torch.linalg.norm(gc32.to(dev) - gc64) / torch.linalg.norm(gc64)),
    }

    # central differences of the FP64 loss by partial re-assembly
    if fd_vertices > 0:
        t = time.perf_counter()
        topo = prob.grid(v_host)
        picks = rng7.choice(NV, fd_vertices, replace=False)
        r_np = r
        fd, ana_64, ana_32 = [], [], []
        el = grid.elements
        for vi in picks:
            E = np.nonzero((el == vi).any(axis=0))[0]
            Et = torch.from_numpy(E).to(dev)
            cE = c[Et]
            colsV = A64[:, Et]
            # potential of the elements at vi, unperturbed
            spE0 = function_space(grid, "DP", 0, support_elements=E)
            ev0 = hip_potential_assembler(spE0, hpot.single_layer(spE0, pts.T, k, precision="double").descriptor,
                                          pts.T, params)
            uE0 = ev0(_full_dp0(spE0, cE), as_tensor=True)[0][0]
            for d in range(3):
                dl = []
                for sgn in (1.0, -1.0):
                    vv = v_host.copy()
                    vv[vi, d] += sgn * 1e-6
                    g2 = topo.with_vertices(vv.T)
                    sp_full = function_space(g2, "DP", 0)
                    spE = function_space(g2, "DP", 0, support_elements=E)
                    dE = helmholtz.single_layer(sp_full, sp_full, spE, k, precision="double").descriptor
                    R = hip_dense_assemble(sp_full, spE, params, dE)  # rows of E
                    C = hip_dense_assemble(spE, sp_full, params, dE)  # columns of E
                    delta = (C.to(torch.complex128) - colsV) @ cE
                    delta[Et] = R @ c - b[Et] - r[Et]
                    dr = (2 * (torch.conj(r) * delta).real + delta.abs().pow(2)).sum()
                    evE = hip_potential_assembler(spE, hpot.single_layer(spE, pts.T, k, precision="double").descriptor,
                                                  pts.T, params)
                    du = evE(_full_dp0(spE, cE), as_tensor=True)[0][0] - uE0
                    dpot = (2 * (torch.conj(u64) * du).real + du.abs().pow(2)).sum()
                    dl.append(float((dr + dpot).item()))
                    clear_caches()
                fd.append((dl[0] - dl[1]) / 2e-6)
                ana_64.append(float(gv64[vi, d]))
                ana_32.append(float(gv32[vi, d]))
        fd, ana_64, ana_32 = map(np.array, (fd, ana_64, ana_32))
        res["fd_check"] = {
            "vertices": int(fd_vertices),
            "h": 1e-6,
            "fp64_adjoint_rel_err_vs_fd": float(np.linalg.norm(ana_64 - fd) / np.linalg.norm(fd)),
            "mixed_adjoint_rel_err_vs_fd": float(np.linalg.norm(ana_32 - fd) / np.linalg.norm(fd)),
            "seconds": time.perf_counter() - t,
        }
    del A64
    if with_cpu:
        res["cpu_baseline"] = cpu_rows_baseline(["helmholtz_single_layer"], grid, sp64, sp64, k, F * F, cpu_seconds,
                                                f"Helmholtz SL DP0 perturbed level {level}, forward assembly")
    return res


def _full_dp0(space, coeffs):
    """Per-element coefficients of a DP0 space (device) from its dof vector."""
    import numpy as np
    import torch

    l2g = torch.from_numpy(space.local2global[:, 0].astype(np.int64)).to(coeffs.device)
    mult = torch.from_numpy(np.ascontiguousarray(space.local_multipliers[:, 0])).to(coeffs.device)
    return (coeffs[l2g] * mult).contiguous()


def efie_bench(level, k, reps, with_cpu=False, cpu_seconds=4.0):
    """Maxwell electric field operator, RWG x SNC, and the E-field potential at 65,536 points.

    Mixed assembly (complex64) timed with HIP events; compared with the FP64 assembly. The
    potential of a seeded complex Gaussian RWG density is evaluated on the radius-10 sphere.
    """
    import numpy as np
    import torch

    import bempp_amd.api as bempp
    from bempp_amd.api.operators.boundary import maxwell
    from bempp_amd.api.operators.potential import maxwell as mpot
    from bempp_amd.core import hip_backend as hb
    from bempp_amd.core.dense_assembler import hip_dense_assemble
    from bempp_amd.core.singular_assembler import singular_rule

    grid = bempp.shapes.icosphere(level)
    rwg = bempp.function_space(grid, "RWG", 0)
    snc = bempp.function_space(grid, "SNC", 0)
    params = bempp.GLOBAL_PARAMETERS
    F, N = grid.number_of_elements, rwg.global_dof_count
    d32 = maxwell.electric_field(rwg, rwg, snc, k, precision="single").descriptor
    d64 = maxwell.electric_field(rwg, rwg, snc, k, precision="double").descriptor
    out = hip_dense_assemble(rwg, snc, params, d32)
    torch.cuda.synchronize()
    evs = [{key: _ev() for key in ("regular_start", "regular_end", "singular_end")} for _ in range(reps)]
    t0 = time.perf_counter()
    for e in evs:
        hip_dense_assemble(rwg, snc, params, d32, out=out, accumulate=False, events=e)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    reg = float(np.mean([e["regular_start"].elapsed_time(e["regular_end"]) for e in evs]))
    # the same assembly over ordered element pairs (parameters.hip.symmetric = False)
    params.hip.symmetric = False
    try:
        hip_dense_assemble(rwg, snc, params, d32, out=out, accumulate=False)
        torch.cuda.synchronize()
        oev = {key: _ev() for key in ("regular_start", "regular_end", "singular_end")}
        t = time.perf_counter()
        hip_dense_assemble(rwg, snc, params, d32, out=out, accumulate=False, events=oev)
        torch.cuda.synchronize()
        ordered_ms = 1e3 * (time.perf_counter() - t)
        ordered_reg = oev["regular_start"].elapsed_time(oev["regular_end"])
    finally:
        params.hip.symmetric = True
    hip_dense_assemble(rwg, snc, params, d32, out=out, accumulate=False)  # the symmetric result for the check
    # the regular section holds the zero fill, the symmetric kernel and the transpose-add: time the
    # two matrix passes alone on the same stream (on a scratch matrix) to isolate the kernel
    scratch = torch.empty_like(out)
    aux = [{key: _ev() for key in "abc"} for _ in range(reps)]
    for e in aux:
        e["a"].record()
        hb.dense_zero(scratch)
        e["b"].record()
        hb.dense_add_transpose(scratch)
        e["c"].record()
    torch.cuda.synchronize()
    zero_ms = float(np.mean([e["a"].elapsed_time(e["b"]) for e in aux]))
    trans_ms = float(np.mean([e["b"].elapsed_time(e["c"]) for e in aux]))
    del scratch
    kern_ms = reg - zero_ms - trans_ms
    ref = hip_dense_assemble(rwg, snc, params, d64)  # warm: blocking, singular rule, allocation
    torch.cuda.synchronize()
    t = time.perf_counter()
    hip_dense_assemble(rwg, snc, params, d64, out=ref, accumulate=False)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t
    err = rel_fro(out, ref)
    del ref, out
    n_sing = singular_rule(grid, params.quadrature.singular, rwg.support, snc.support).host.number_of_pairs
    n_evals = (F * F - n_sing) * 36
    # E-field potential: 65,536 points on the radius-10 sphere, complex Gaussian RWG density
    rng = np.random.default_rng(2)
    c = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    g = np.random.default_rng(3).standard_normal((3, 65536))
    pts = 10.0 * g / np.linalg.norm(g, axis=0)
    fun = bempp.GridFunction(rwg, coefficients=c)
    p32 = mpot.electric_field(rwg, pts, k, precision="single")
    p64 = mpot.electric_field(rwg, pts, k, precision="double")
    u64 = p64.evaluate(fun)
    p32.evaluate(fun)
    torch.cuda.synchronize()
    pe0, pe1 = _ev(), _ev()
    t = time.perf_counter()
    pe0.record()
    for _ in range(reps):
        u32 = p32.evaluate(fun)
    pe1.record()
    torch.cuda.synchronize()
    t_pot = (time.perf_counter() - t) / reps
    pot_ms = pe0.elapsed_time(pe1) / reps
    n_pot = pts.shape[1] * F * 6
    res = {
        "workload": f"Maxwell EFIE RWG x SNC, icosphere level {level} ({F} triangles, {N} edges), k={k}, "
                    "mixed -> complex64; E-field potential at 65536 points, radius 10",
        "entries_per_s": N * N / wall,
        "ms": 1e3 * wall,
        "regular_section_ms": reg,
        "zero_fill_ms": zero_ms,
        "add_transpose_ms": trans_ms,
        "regular_kernel_ms": kern_ms,
        "regular_tflops": n_evals * 73 / (kern_ms * 1e-3) / 1e12,
        "flop_per_eval": 73,
        "traversal": "symmetric (each unordered element pair of the regular part once, then the transpose; "
                     "the flop count is that of the ordered pairs)",
        "ordered_ms": ordered_ms,
        "ordered_regular_ms": ordered_reg,
        "rel_err_vs_fp64": err,
        "fp64_ms": 1e3 * t64,
        "speedup_vs_fp64": t64 / wall,
        "roofline": roofline("regular_kernel<maxwell_electric_field, RWGxSNC, mixed, symmetric>", "valu",
                             n_evals * 73 / (kern_ms * 1e-3) / 1e12, PEAK_FP32_TFLOPS, "TFLOP/s",
                             traffic_key="regular_kernel_maxwell_efie_rwg_level4_mixed", evals_per_launch=n_evals,
                             flop_per_eval=73, launch_ms=kern_ms,
                             # the symmetric kernel evaluates each unordered point pair once: the
                             # same count over the pairs it evaluates is half the algorithmic rate
                             evaluated_point_pairs_per_launch=n_evals // 2,
                             evaluated_pair_tflops=n_evals / 2 * 73 / (kern_ms * 1e-3) / 1e12),
        "potential": {
            "points": pts.shape[1],
            "evals_per_s": pts.shape[1] / t_pot,
            "ms": 1e3 * t_pot,
            "tflops": n_pot * 91 / t_pot / 1e12,
            "rel_err_vs_fp64": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
            "roofline": roofline("potential_eval_mixed_efie (+ source prep and chunk reduce)",
                                 "valu", n_pot * 91 / (pot_ms * 1e-3) / 1e12, PEAK_FP32_TFLOPS, "TFLOP/s",
                                 traffic_key="potential_eval_mixed_efie_65536", evals_per_launch=n_pot,
                                 flop_per_eval=91, launch_ms=pot_ms),
        },
    }
    if with_cpu:
        res["cpu_baseline"] = cpu_rows_baseline(["maxwell_electric_field"], grid, rwg, snc, k, N * N, cpu_seconds,
                                                f"Maxwell EFIE RWGxSNC level {level}")
        res["potential"]["cpu_baseline"] = cpu_potential_baseline(
            "maxwell_electric_field", grid, rwg, c, pts, k, pts.shape[1], cpu_seconds,
            "E-field potential of the RWG density")
    return res


def frequency_sweep_bench(level, reps, with_cpu=False, cpu_seconds=4.0):
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
    from bempp_amd.core import autotune as at

    # the tuner's record for this problem: chosen launch shape and every candidate's best time
    tuned = [v for k, v in at._load().items() if k.split("|")[3:6] == [str(F), str(F), str(len(ks))]
             and "single" in k]
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
    d = helmholtz.single_layer(sp, sp, sp, ks[0], precision="double").descriptor
    hip_dense_assemble(sp, sp, params, d, out=one, accumulate=False)  # warm
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
    res = {
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
        "autotune": tuned,
        "note": "the mixed kernel expands e^{ik(r-r0)} in moments of r - r0, so the algorithmic "
                "FLOP rate counts the work of direct evaluation; the kernel issues VALU instructions about 3/4 "
                "of its time (SQ_ACTIVE_INST_VALU) and writes the complex64 output at under half the HBM rate",
        "roofline": roofline("kbatch_far_kernel (persistent, 32 wavenumbers, mixed)", "valu",
                             flops / (reg_ms * 1e-3) / 1e12, PEAK_FP32_TFLOPS, "TFLOP/s",
                             traffic_key="kbatch_far_kernel_dp0_level4_32k", algorithmic_flop_per_launch=flops,
                             flop_per_eval="10 + 9 per wavenumber", algorithmic_bytes_per_launch=entries * 8,
                             hbm_write_gbs=entries * 8 / (reg_ms * 1e-3) / 1e9,
                             hbm_write_frac=entries * 8 / (reg_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, launch_ms=reg_ms),
    }
    if with_cpu:
        res["cpu_baseline"] = cpu_rows_baseline(["helmholtz_single_layer"], grid, sp, sp, float(ks[0]), entries,
                                                cpu_seconds, f"Helmholtz SL DP0 level {level}, one wavenumber",
                                                repeat=len(ks))
    return res


def solve_bench(space, A, k, tol=1e-5):
    """GMRES (restart 20, rtol 1e-5: the reference's defaults) on the assembled mixed single layer.

    The right-hand side is the projection of the plane wave e^{ik x} onto the P1 space; the
    matrix is the complex64 operator the assembly step left in HBM. Also times the HIP product
    alone (bem_dense_matvec, one complex64 matrix read per product) for its HBM roofline.
    """
    import numpy as np
    import torch

    import bempp_amd.api as bempp
    from bempp_amd.api.linalg.iterative_solvers import IterationCounter, device_gmres
    from bempp_amd.core import hip_backend as hb

    @bempp.complex_callable
    def plane_wave(x, n, domain_index, res):
        res[0] = np.exp(1j * k * x[0])

    b = bempp.GridFunction(space, fun=plane_wave).projections()
    counter = IterationCounter(True)
    device_gmres(A, b, tol, None, None, counter)  # warm-up (library paths, rocBLAS handles)
    torch.cuda.synchronize()
    counter = IterationCounter(True)
    t = time.perf_counter()
    x, info = device_gmres(A, b, tol, None, None, counter)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t
    xh = x.cpu().numpy()
    # true relative residual with the FP64 product on the device
    r = torch.as_tensor(b, device=A.device) - hb.dense_matvec(A, x)
    rel_res = float(torch.linalg.vector_norm(r)) / float(np.linalg.norm(b))
    # the product alone, HIP events on the current stream
    v = torch.as_tensor(xh, device=A.device)
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    hb.dense_matvec(A, v)
    e0.record()
    for _ in range(reps):
        hb.dense_matvec(A, v)
    e1.record()
    torch.cuda.synchronize()
    mv_ms = e0.elapsed_time(e1) / reps
    n = A.shape[0]
    bytes_mv = n * n * A.element_size()
    return {
        "workload": f"GMRES(20), rtol {tol}, Helmholtz single layer P1 {n} dofs (complex64, mixed), plane-wave rhs",
        "iterations": counter.count,
        "info": int(info),
        "seconds": t_solve,
        "ms_per_iteration": 1e3 * t_solve / max(counter.count, 1),
        "final_rel_residual": rel_res,
        "matvec_ms": mv_ms,
        "matvec_hbm_gb_s": bytes_mv / (mv_ms * 1e-3) / 1e9,
        "matvec_hbm_frac": bytes_mv / (mv_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
    }


def hypersingular_bench(space, k, reps=3):
    """Helmholtz hypersingular P1 x P1 on the same space: mixed vs FP64 (time and error)."""
    import torch

    from bempp_amd.api.operators.boundary import helmholtz

    def run(precision):
        op = helmholtz.hypersingular(space, space, space, k, precision=precision)
        from bempp_amd.core.dense_assembler import hip_dense_assemble

        out = hip_dense_assemble(space, space, op.assembler.parameters, op.descriptor)  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            out = hip_dense_assemble(space, space, op.assembler.parameters, op.descriptor, out=out, accumulate=False)
        torch.cuda.synchronize()
        return out, (time.perf_counter() - t) / reps

    m32, t32 = run("single")
    m64, t64 = run("double")
    n = space.global_dof_count
    return {
        "workload": f"Helmholtz hypersingular P1 x P1, {n} dofs, k={k}",
        "mixed_ms": 1e3 * t32,
        "fp64_ms": 1e3 * t64,
        "mixed_entries_per_s": n * n / t32,
        "speedup_vs_fp64": t64 / t32,
        "rel_err_vs_fp64": rel_fro(m32, m64),
    }


def maxwell_extras_bench(level, k, reps=3):
    """C4's grid and wavenumber: the MFIE (RWG x SNC, mixed and FP64), the magnetic field potential
    at 65,536 points on radius 10 and the electric / magnetic far fields at 65,536 directions of
    the same seeded RWG density."""
    import numpy as np
    import torch

    import bempp_amd.api as bempp
    from bempp_amd.api.operators.boundary.maxwell import magnetic_field
    from bempp_amd.api.operators.far_field import maxwell as ff
    from bempp_amd.api.operators.potential import maxwell as mpot
    from bempp_amd.core.dense_assembler import hip_dense_assemble

    grid = bempp.shapes.icosphere(level)
    rwg = bempp.function_space(grid, "RWG", 0)
    snc = bempp.function_space(grid, "SNC", 0)
    n = rwg.global_dof_count
    op = magnetic_field(rwg, rwg, snc, k, precision="single")
    out = hip_dense_assemble(rwg, snc, op.assembler.parameters, op.descriptor)  # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        out = hip_dense_assemble(rwg, snc, op.assembler.parameters, op.descriptor, out=out, accumulate=False)
    torch.cuda.synchronize()
    t_mfie = (time.perf_counter() - t) / reps
    rng = np.random.default_rng(2)
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    fun = bempp.GridFunction(rwg, coefficients=c)
    g = np.random.default_rng(3).standard_normal((3, 65536))
    dirs = g / np.linalg.norm(g, axis=0)
    op64 = magnetic_field(rwg, rwg, snc, k, precision="double")
    ref = hip_dense_assemble(rwg, snc, op64.assembler.parameters, op64.descriptor)  # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    ref = hip_dense_assemble(rwg, snc, op64.assembler.parameters, op64.descriptor, out=ref, accumulate=False)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t
    res = {"workload": f"icosphere level {level} ({grid.number_of_elements} triangles, {n} RWG dofs), k={k}",
           "mfie_ms": 1e3 * t_mfie, "mfie_entries_per_s": n * n / t_mfie, "mfie_fp64_ms": 1e3 * t64,
           "mfie_speedup_vs_fp64": t64 / t_mfie, "mfie_rel_err_vs_fp64": rel_fro(out, ref)}
    del ref
    import numpy as np

    for name, factory, pts in (("magnetic_potential", mpot.magnetic_field, 10.0 * dirs),
                               ("electric_far_field", ff.electric_field, dirs),
                               ("magnetic_far_field", ff.magnetic_field, dirs)):
        vals = {}
        for precision in ("single", "double"):
            opr = factory(rwg, pts, k, precision=precision)
            opr.evaluate(fun)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                vals[precision] = opr.evaluate(fun)
            dt = (time.perf_counter() - t) / reps
            if precision == "single":
                res[f"{name}_ms"] = 1e3 * dt
                res[f"{name}_evals_per_s"] = pts.shape[1] * grid.number_of_elements * 6 / dt
            else:
                res[f"{name}_fp64_ms"] = 1e3 * dt
        u32, u64 = vals["single"].astype(np.complex128), vals["double"]
        res[f"{name}_rel_err_vs_fp64"] = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
    return res


# rocprofv3 --pmc traffic per kernel: measured on the kernel sources whose hash is recorded with it
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
CSRC = os.path.join(ROOT, "bempp-cl_amd", "csrc")


def kernel_sources_sha(files):
    """sha256 (16 hex digits) of the given HIP source files, in order."""
    import hashlib

    h = hashlib.sha256()
    for name in files:
        with open(os.path.join(CSRC, name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def traffic_from_profiles(kernel_key):
    """HBM bytes per launch of a kernel from the recorded PMC runs (FETCH_SIZE, WRITE_SIZE), with
    the commit they were measured at. A record whose kernel sources changed since is stale:
    traffic is then reported as null."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, None
    with open(TRAFFIC_FILE) as f:
        entry = json.load(f).get(kernel_key)
    if entry is None:
        return None, None
    files = entry.get("kernel_sources")
    info = {"measured_at_commit": entry.get("measured_at_commit"), "profile": entry.get("source")}
    if not files or kernel_sources_sha(files) != entry.get("kernel_sources_sha"):
        info["stale"] = "kernel sources changed since the PMC run"
        return None, info
    return float(entry["hbm_bytes_per_launch"]), info


def roofline(kernel, bound, achieved, peak, unit, traffic_key=None, **extra):
    traffic, info = traffic_from_profiles(traffic_key) if traffic_key else (None, None)
    out = {"kernel": kernel, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
           "frac": achieved / peak, "traffic": traffic}
    if peak == PEAK_FP32_TFLOPS:
        # the published FP32 peak counts packed FP32 instructions (two FMAs per lane); this build
        # compiles without them, so one FP32 instruction per lane and cycle, the FP64 rate, is the ceiling
        out["peak_nonpacked_fp32"] = PEAK_FP64_TFLOPS
        out["frac_nonpacked_fp32"] = achieved / PEAK_FP64_TFLOPS
    if info:
        out["traffic_record"] = info
    out.update(extra)
    return out


def rel_fro(a, b):
    import torch

    return float(torch.linalg.vector_norm(a.to(b.dtype) - b) / torch.linalg.vector_norm(b))


def host_cores():
    return len(os.sched_getaffinity(0))


def cpu_rows_baseline(ops, grid, domain, dual, k, n_entries, target_seconds, label, repeat=1):
    """The FP64 oracle (C/OpenMP restatement of the reference's Numba algorithm) on all host cores,
    timed on a block of test elements and extrapolated linearly to all rows.

    ops: operator names assembled per step (their times add); repeat: identical assemblies per
    step (e.g. one per wavenumber of a sweep). n_entries: matrix entries of one step. A two-point
    calibration separates the fixed cost (result allocation, colouring) from the per-row cost;
    the singular pairs are timed on a smaller block."""
    from oracle import oracle as O

    cores = host_cores()
    F = grid.number_of_elements

    def timed(op, rows, parts):
        t = time.perf_counter()
        O.assemble(op, grid.vertices, grid.elements, domain, dual, k, parts=parts, test_range=(0, rows),
                   nthreads=cores, rows_subset=[])
        return time.perf_counter() - t

    lo, hi = min(32, F), min(96, F)
    ta = timed(ops[0], lo, 1)
    tb = timed(ops[0], hi, 1)
    per_row = max((tb - ta) / max(hi - lo, 1), 1e-7)
    fixed = max(ta - lo * per_row, 0.0)
    n1 = int(min(F, max(hi, target_seconds / len(ops) / per_row)))
    est, sampled = 0.0, 0.0
    for op in ops:
        t_reg = timed(op, n1, 1)
        sampled += t_reg
        ns = max(1, n1 // 8)
        t_sing = timed(op, ns, 2)
        est += fixed + (t_reg - fixed) * F / n1 + max(t_sing - fixed, 0.0) * F / ns
    est *= repeat
    return {
        "value": n_entries / est,
        "unit": "matrix entries/s",
        "cores": cores,
        "kind": "port",
        "sample": f"FP64 oracle (C/OpenMP restatement of the reference Numba algorithm), {label}: regular pairs of "
                  f"test elements [0,{n1}) of {F} ({sampled:.1f} s) and singular pairs of [0,{max(1, n1 // 8)}), "
                  f"per-row cost extrapolated linearly to all rows"
                  + (f", times {repeat} identical assemblies" if repeat > 1 else "")
                  + f"; estimated full step {est:.1f} s",
    }


def cpu_potential_baseline(op, grid, space, coeffs, points, k, n_points_step, target_seconds, label):
    """The FP64 oracle potential on all host cores for a subset of the points, scaled to the step."""
    from oracle import oracle as O

    cores = host_cores()
    m = min(points.shape[1], 64)
    t = time.perf_counter()
    O.potential(op, grid.vertices, grid.elements, space, coeffs, points[:, :m], k, nthreads=cores)
    per_point = max((time.perf_counter() - t) / m, 1e-9)
    m = int(min(points.shape[1], max(m, target_seconds / per_point)))
    t = time.perf_counter()
    O.potential(op, grid.vertices, grid.elements, space, coeffs, points[:, :m], k, nthreads=cores)
    dt = time.perf_counter() - t
    est = dt * n_points_step / m
    return {
        "value": n_points_step / est,
        "unit": "potential evaluation points/s",
        "cores": cores,
        "kind": "port",
        "sample": f"FP64 oracle potential (restated reference algorithm), {label}: {m} of {n_points_step} points "
                  f"({dt:.1f} s), scaled linearly; estimated full evaluation {est:.1f} s",
    }


def cpu_baseline(grid, space, k, n_entries_step, target_seconds):
    """C2 step: Helmholtz single + double layer P1 at level 5."""
    return cpu_rows_baseline(["helmholtz_single_layer", "helmholtz_double_layer"], grid, space, space, k,
                             n_entries_step, target_seconds, "Helmholtz SL+DL P1")


def c1_bench(reps, with_cpu, cpu_seconds):
    """C1: Laplace single layer, DP0 x DP0, level-3 icosphere (1280 triangles), FP64, regular and
    singular order 4. HIP FP64 assembly timed with HIP events; the CPU baseline is the oracle's
    full assembly (not sampled) on all host cores, which is also the parity check."""
    import numpy as np
    import torch

    import bempp_amd.api as bempp
    from bempp_amd.api.operators.boundary import laplace
    from bempp_amd.core.dense_assembler import hip_dense_assemble
    from bempp_amd.core.singular_assembler import singular_rule

    grid = bempp.shapes.icosphere(3)
    sp = bempp.function_space(grid, "DP", 0)
    params = bempp.GLOBAL_PARAMETERS
    F = grid.number_of_elements
    desc = laplace.single_layer(sp, sp, sp, precision="double").descriptor
    out = hip_dense_assemble(sp, sp, params, desc)  # host set-up, code objects
    torch.cuda.synchronize()
    evs = [{key: _ev() for key in ("regular_start", "regular_end", "singular_end")} for _ in range(reps)]
    t0 = time.perf_counter()
    for e in evs:
        hip_dense_assemble(sp, sp, params, desc, out=out, accumulate=False, events=e)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    reg = float(np.mean([e["regular_start"].elapsed_time(e["regular_end"]) for e in evs]))
    sing = float(np.mean([e["regular_end"].elapsed_time(e["singular_end"]) for e in evs]))
    n_sing = singular_rule(grid, params.quadrature.singular, sp.support, sp.support).host.number_of_pairs
    n_evals = (F * F - n_sing) * 36
    res = {
        "workload": "Laplace single layer DP0xDP0, icosphere level 3 (1280 triangles), FP64, regular and singular "
                    "order 4",
        "entries_per_s": F * F / wall,
        "ms": 1e3 * wall,
        "regular_section_ms": reg,
        "singular_ms": sing,
        "roofline": roofline("regular_kernel<laplace_single_layer, DP0, FP64, upper triangle and its mirror image>", "valu",
                             n_evals * 11 / (reg * 1e-3) / 1e12, PEAK_FP64_TFLOPS, "TFLOP/s",
                             traffic_key="regular_kernel_laplace_sl_dp0_upper_level3_fp64",
                             evals_per_launch=n_evals, flop_per_eval=11),
    }
    if with_cpu:
        from oracle import oracle as O

        cores = host_cores()
        t = time.perf_counter()
        ref = O.assemble("laplace_single_layer", grid.vertices, grid.elements, sp, sp, 0.0, nthreads=cores)
        t_cpu = time.perf_counter() - t
        got = out.cpu().numpy()
        res["rel_err_vs_oracle"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        res["cpu_baseline"] = {"value": F * F / t_cpu, "unit": "matrix entries/s", "cores": cores, "kind": "port",
                               "sample": f"full assembly (regular + singular pairs, not sampled) by the FP64 oracle, "
                                         f"the C/OpenMP restatement of the reference Numba algorithm: {t_cpu:.2f} s"}
        res["speedup_vs_cpu_baseline"] = res["entries_per_s"] / res["cpu_baseline"]["value"]
    return res


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
    from bempp_amd.core.dense_assembler import hip_dense_assemble, hip_dense_assemble_sl_dl
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
    SL, DL = "helmholtz_single_layer", "helmholtz_double_layer"
    evkeys = ("regular_start", "regular_end", "singular_end")

    def step(events=None):
        # single and double layer in one pass (shared point-pair geometry and e^{ikr})
        hip_dense_assemble_sl_dl(space, space, params, desc[SL], desc[DL], out_single=outs[SL],
                                 out_double=outs[DL], events=events)

    # cold: the first step on this grid, with the host set-up the warm steps reuse (symmetric
    # blocking, near-field candidates, singular pair tables, device copies). Code objects are
    # loaded first by a small assembly on another grid (a once-per-process cost).
    cold = {}
    small = bempp.function_space(bempp.shapes.icosphere(1), "P", 1)
    hip_dense_assemble_sl_dl(small, small, params, helmholtz.single_layer(small, small, small, k, precision="single").descriptor,
                             helmholtz.double_layer(small, small, small, k, precision="single").descriptor)
    torch.cuda.synchronize()
    from bempp_amd.core import clear_caches

    clear_caches()
    t = time.perf_counter()
    step()
    torch.cuda.synchronize()
    cold["first_step_s"] = time.perf_counter() - t
    log(f"cold first step {cold['first_step_s']:.3f} s")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [{key: torch.cuda.Event(enable_timing=True) for key in evkeys} for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(ev[s])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = 1e3 * (t1 - t0) / args.steps
    entries_per_step = 2 * N * N
    value = entries_per_step / ((t1 - t0) / args.steps)
    reg = float(np.mean([e["regular_start"].elapsed_time(e["regular_end"]) for e in ev]))
    sing = float(np.mean([e["regular_end"].elapsed_time(e["singular_end"]) for e in ev]))
    flop_pair = FLOP_PER_EVAL[SL] + FLOP_PER_EVAL[DL]
    # the regular section is zero fill (both outputs), the symmetric kernel and the transpose-add
    # of the single layer; the two helpers are timed alone on the same stream and subtracted
    aux = [{key: torch.cuda.Event(enable_timing=True) for key in ("a", "b", "c")} for _ in range(3)]
    scratch = torch.empty_like(outs[SL])
    for e in aux:
        e["a"].record()
        hb.dense_zero(scratch)
        hb.dense_zero(scratch)
        e["b"].record()
        hb.dense_add_transpose(scratch)
        e["c"].record()
    torch.cuda.synchronize()
    zero_ms = float(np.mean([e["a"].elapsed_time(e["b"]) for e in aux]))
    trans_ms = float(np.mean([e["b"].elapsed_time(e["c"]) for e in aux]))
    del scratch
    kern_ms = reg - zero_ms - trans_ms
    kern = {"symmetric_regular_section_ms": reg, "symmetric_kernel_ms": kern_ms, "zero_fill_ms": zero_ms,
            "add_transpose_ms": trans_ms, "symmetric_near_singular_ms": sing,
            "symmetric_kernel_tflops": n_evals * flop_pair / (kern_ms * 1e-3) / 1e12,
            "symmetric_kernel_executed_tflops": n_evals / 2 * EXEC_FLOP_PER_SYM_POINT_PAIR / (kern_ms * 1e-3) / 1e12}
    log(f"mixed symmetric step {ms_per_step:.2f} ms, {value / 1e9:.2f} G entries/s; kernels {kern}")

    # the same fused pass over ordered pairs (parameters.hip.symmetric = False)
    params.hip.symmetric = False
    step()
    torch.cuda.synchronize()
    oev = {key: torch.cuda.Event(enable_timing=True) for key in evkeys}
    t = time.perf_counter()
    step(oev)
    torch.cuda.synchronize()
    kern["ordered_fused_ms_per_step"] = 1e3 * (time.perf_counter() - t)
    kern["ordered_fused_regular_ms"] = oev["regular_start"].elapsed_time(oev["regular_end"])
    params.hip.symmetric = True
    log(f"ordered fused step {kern['ordered_fused_ms_per_step']:.2f} ms (regular {kern['ordered_fused_regular_ms']:.2f})")

    # the same two operators assembled separately (one weak_form() each)
    sep_ev = {name: {key: torch.cuda.Event(enable_timing=True) for key in evkeys} for name in ops}
    for name in ops:
        hip_dense_assemble(space, space, params, desc[name], out=outs[name], accumulate=False)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for name in ops:
        hip_dense_assemble(space, space, params, desc[name], out=outs[name], accumulate=False, events=sep_ev[name])
    torch.cuda.synchronize()
    sep_ms = 1e3 * (time.perf_counter() - t)
    for name in ops:
        r = sep_ev[name]["regular_start"].elapsed_time(sep_ev[name]["regular_end"])
        kern[name] = {"regular_ms": r, "singular_ms": sep_ev[name]["regular_end"].elapsed_time(sep_ev[name]["singular_end"]),
                      "regular_tflops": n_evals * FLOP_PER_EVAL[name] / (r * 1e-3) / 1e12}
    kern["separate_ms_per_step"] = sep_ms
    step()  # leave the fused result in outs for the error check
    torch.cuda.synchronize()

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
                        "k=4.0, mixed FP32/FP64 -> complex64, dense in HBM, both operators in one symmetric pass",
            "triangles": F,
            "dofs": N,
            "wavenumber": k,
            "quadrature": "regular order 4 (6 points), singular order 4",
            "regular_pairs": n_regular_pairs,
            "singular_pairs": n_sing,
        },
    }
    result["roofline"] = {
        "kernel": "regular_kernel<helmholtz_single_layer + double layer, symmetric, P1xP1, mixed>",
        "bound": "valu",
        "achieved": kern["symmetric_kernel_tflops"],
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": kern["symmetric_kernel_tflops"] / PEAK_FP32_TFLOPS,
        "traffic": traffic_from_profiles("regular_kernel_helmholtz_sl_dl_sym_p1_level5_mixed")[0],
        "traffic_record": traffic_from_profiles("regular_kernel_helmholtz_sl_dl_sym_p1_level5_mixed")[1],
        "algorithmic_flop_per_launch": n_evals * flop_pair,
        "flop_per_eval": f"{FLOP_PER_EVAL[SL]} (single layer) + {FLOP_PER_EVAL[DL]} (double layer) per ordered "
                         f"point pair; the symmetric pass evaluates each unordered point pair once",
        "evals_per_launch": n_evals,
        "executed_flop_per_launch": n_evals / 2 * EXEC_FLOP_PER_SYM_POINT_PAIR,
        "executed_tflops": kern["symmetric_kernel_executed_tflops"],
        "launch_ms": kern_ms,
    }
    result["kernels"] = kern
    result["cold"] = cold

    if not args.no_reference:
        # FP64: the plain path (one assembly per operator) and the fused FP64 pass
        # every FP64 variant is timed on its second call (the first builds blockings and allocates)
        errs, t64 = {}, 0.0
        a64 = {}
        for name in ops:
            a64[name] = hip_dense_assemble(space, space, params, desc64[name])
            torch.cuda.synchronize()
            t = time.perf_counter()
            hip_dense_assemble(space, space, params, desc64[name], out=a64[name], accumulate=False)
            torch.cuda.synchronize()
            t64 += time.perf_counter() - t
            errs[name] = rel_fro(outs[name], a64[name])
        # FP64 fused over ordered pairs, then FP64 symmetric (the same algorithm as the step)
        fp64 = {}
        for label, sym in (("ordered_fused", False), ("symmetric", True)):
            params.hip.symmetric = sym
            f1, f2 = hip_dense_assemble_sl_dl(space, space, params, desc64[SL], desc64[DL])
            torch.cuda.synchronize()
            t = time.perf_counter()
            hip_dense_assemble_sl_dl(space, space, params, desc64[SL], desc64[DL], out_single=f1, out_double=f2)
            torch.cuda.synchronize()
            fp64[label] = (time.perf_counter() - t, max(rel_fro(f1, a64[SL]), rel_fro(f2, a64[DL])))
            del f1, f2
        params.hip.symmetric = True
        del a64
        result["rel_err_vs_fp64"] = errs
        result["fp64_entries_per_s"] = entries_per_step / t64
        result["speedup_vs_fp64"] = value / result["fp64_entries_per_s"]
        for label, (tl, el) in fp64.items():
            result[f"fp64_{label}_entries_per_s"] = entries_per_step / tl
            result[f"speedup_vs_fp64_{label}"] = value / result[f"fp64_{label}_entries_per_s"]
            result[f"fp64_{label}_vs_separate_rel_err"] = el
        result["speedup_separate_mixed_vs_fp64"] = t64 / (sep_ms * 1e-3)
        result["speedup_ordered_fused_mixed_vs_fp64_ordered_fused"] = fp64["ordered_fused"][0] / (
            kern["ordered_fused_ms_per_step"] * 1e-3)
        log(f"fp64 separate {1e3 * t64:.1f} ms, ordered fused {1e3 * fp64['ordered_fused'][0]:.1f} ms, symmetric "
            f"{1e3 * fp64['symmetric'][0]:.1f} ms, rel err {errs}")

        # the same step with Kahan-compensated FP32 sums (parameters.hip.kahan)
        from bempp_amd.api.utils.parameters import DefaultParameters

        pk = DefaultParameters()
        pk.hip.kahan = True
        dk = {name: f(space, space, space, k, precision="single", parameters=pk).descriptor
              for name, (f, _) in ops.items()}
        hip_dense_assemble_sl_dl(space, space, pk, dk[SL], dk[DL], out_single=outs[SL], out_double=outs[DL])
        torch.cuda.synchronize()
        t = time.perf_counter()
        hip_dense_assemble_sl_dl(space, space, pk, dk[SL], dk[DL], out_single=outs[SL], out_double=outs[DL])
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t)
        kres = {"ms_per_step": ms}
        for name in ops:
            ref = hip_dense_assemble(space, space, params, desc64[name])
            kres[name] = {"rel_err_vs_fp64": rel_fro(outs[name], ref)}
            del ref
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
        pe0, pe1 = _ev(), _ev()
        pe0.record()
        for _ in range(reps):
            u32 = pot32.evaluate(fun)
        pe1.record()
        torch.cuda.synchronize()
        t_pot = (time.perf_counter() - t) / reps
        pot_ms = pe0.elapsed_time(pe1) / reps
        n_pot = pts.shape[1] * F * 6
        result["potential"] = {
            "points": pts.shape[1],
            "sources": F * 6,
            "evals_per_s": pts.shape[1] / t_pot,
            "ms": 1e3 * t_pot,
            "tflops": n_pot * FLOP_PER_POTENTIAL_EVAL / t_pot / 1e12,
            "rel_err_vs_fp64": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
            "roofline": roofline("potential_eval_mixed_scalar<helmholtz_single_layer> (+ source prep and chunk reduce)",
                                 "valu", n_pot * FLOP_PER_POTENTIAL_EVAL / (pot_ms * 1e-3) / 1e12, PEAK_FP32_TFLOPS,
                                 "TFLOP/s", traffic_key="potential_eval_mixed_helmholtz_sl_p1_16384",
                                 evals_per_launch=n_pot, flop_per_eval=FLOP_PER_POTENTIAL_EVAL, launch_ms=pot_ms),
        }
        if not args.no_cpu_baseline:
            result["potential"]["cpu_baseline"] = cpu_potential_baseline(
                "helmholtz_single_layer", grid, space, c, pts, k, pts.shape[1], args.cpu_seconds / 3,
                "Helmholtz SL potential of the P1 density")
        log(f"potential {result['potential']}")

    if not args.no_solve:
        result["solve"] = solve_bench(space, outs[SL], k)
        log(f"solve {result['solve']}")

    if not args.no_hypersingular:
        result["hypersingular"] = hypersingular_bench(space, k)
        log(f"hypersingular {result['hypersingular']}")

    if not args.no_maxwell_extras:
        result["maxwell_extras"] = maxwell_extras_bench(4, 2.0)
        log(f"maxwell extras {result['maxwell_extras']}")

    if not args.no_c1:
        result["c1"] = c1_bench(5, not args.no_cpu_baseline, args.cpu_seconds)
        log(f"c1 {result['c1']}")

    with_cpu = not args.no_cpu_baseline
    if not args.no_efie:
        result["efie"] = efie_bench(4, 2.0, 3, with_cpu, args.cpu_seconds / 3)
        log(f"efie {result['efie']}")

    if not args.no_sweep:
        result["frequency_sweep"] = frequency_sweep_bench(args.sweep_level, args.sweep_reps, with_cpu,
                                                          args.cpu_seconds / 3)
        log(f"frequency sweep {result['frequency_sweep']}")

    if not args.no_adjoint:
        del outs
        torch.cuda.empty_cache()
        result["adjoint"] = adjoint_bench(args.adjoint_level, args.adjoint_k, args.fd_vertices, with_cpu,
                                          args.cpu_seconds / 3)
        log(f"adjoint {result['adjoint']}")
        torch.cuda.empty_cache()

    if not args.no_cpu_baseline:
        log("timing the CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline(grid, space, k, entries_per_step, args.cpu_seconds)
        log(f"cpu baseline {result['cpu_baseline']}")
    result["configs"] = config_summary(result)
    print(json.dumps(result), flush=True)


def config_summary(result):
    """Per benchmark configuration: throughput, error, roofline fraction and CPU baseline."""

    def pick(d, *keys):
        return {k: d.get(k) for k in keys if k in d}

    out = {}
    if "c1" in result:
        c = result["c1"]
        out["C1"] = dict(pick(c, "entries_per_s", "rel_err_vs_oracle", "roofline", "cpu_baseline"))
    out["C2"] = {"entries_per_s": result["value"], "rel_err_vs_fp64": result.get("rel_err_vs_fp64"),
                 "roofline": result["roofline"], "cpu_baseline": result.get("cpu_baseline")}
    if "potential" in result:
        out["C2_potential"] = pick(result["potential"], "evals_per_s", "rel_err_vs_fp64", "roofline", "cpu_baseline")
    if "frequency_sweep" in result:
        out["C3"] = pick(result["frequency_sweep"], "entries_per_s", "max_rel_err_vs_fp64", "roofline", "cpu_baseline")
    if "efie" in result:
        out["C4"] = pick(result["efie"], "entries_per_s", "rel_err_vs_fp64", "roofline", "cpu_baseline")
        out["C4_potential"] = pick(result["efie"]["potential"], "evals_per_s", "rel_err_vs_fp64", "roofline",
                                   "cpu_baseline")
    if "adjoint" in result:
        a = result["adjoint"]
        out["C5"] = dict(pick(a, "forward_s", "backward_s", "backward_over_forward", "grad_vertices_rel_err_vs_fp64",
                              "roofline", "roofline_backward", "cpu_baseline"))
    return out


if __name__ == "__main__":
    main()
