"""Benchmark: assembly + solve time to FP64 relative residual 1e-8 on one GPU.

Workload (default): 3D heterogeneous linear elasticity on a 64^3 hexahedral
mesh of the unit cube, H1 p=2 (6,440,067 equations, 610,745,670 stored upper
CSR entries), Young's modulus exp(GRF) (correlation length 0.1, sigma 1),
nu = 0.3, clamped base, seeded smooth body load.  Synthetic seeded data.

A step = one numeric assembly (element integration + scatter into the
reference's upper CSR, rhs) + one solve to ||f - K u|| / ||f|| <= 1e-8 in
FP64, with the mesh, pattern and all inputs already resident in HBM.
The measured path is the mixed-precision FCG (FP32 inner, FP64 refinement);
the plain FP64 Jacobi-PCG GPU path is timed once beside it (same assembly,
its own solve) for the speed-up, and the CPU oracle is timed on a bounded
sample for the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C5", choices=["C1", "C2", "C3", "C4p1", "C4p2", "C5"],
                    help="workload (C5 = the headline 64^3 elasticity case)")
    ap.add_argument("--n", type=int, default=0, help="override elements per direction")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-baseline", action="store_true", help="skip the FP64 Jacobi-PCG run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle sample")
    ap.add_argument("--precond", default="multigrid", choices=["multigrid", "block_jacobi", "learned"])
    ap.add_argument("--mg-degree", type=int, default=0, help="multigrid Chebyshev degree (0 = 2)")
    ap.add_argument("--mg-range", type=float, default=0.0, help="multigrid smoothing interval ratio (0 = 30)")
    ap.add_argument("--no-bj", action="store_true", help="skip the block-Jacobi mixed solve beside multigrid")
    ap.add_argument("--refine", type=int, default=0, help="0 restart FCG per refinement, 1 reliable update")
    ap.add_argument("--inner-tol", type=float, default=1e-4)
    ap.add_argument("--no-graph", action="store_true", help="plain launches instead of hipGraph replay")
    ap.add_argument("--inner-op", default="auto", choices=["auto", "bsr", "matfree"],
                    help="FP32 operator of the inner iterations")
    ap.add_argument("--kernels-only", action="store_true",
                    help="assemble once, then time each operator kernel (5 launches) and a 1 GiB "
                         "device copy (HBM counter calibration); no solve")
    return ap.parse_args()


WORKLOADS = {
    "C5": "3D heterogeneous elasticity {n}^3 hex H1 p=2, E=exp(GRF l=0.1 s=1), nu=0.3, clamped "
          "base, body load",
    "C2": "3D elasticity {n}^3 hex H1 p=2, E lognormal iid s=1, nu=0.3, clamped x=0, unit "
          "traction x=1",
    "C1": "2D Poisson {n}x{n} quads H1 p=2, 8 Gaussian bumps, homogeneous Dirichlet",
    "C3": "3D hp Darcy on 6x{n}^3 Kuhn tetrahedra, p_e uniform in 1..5 by seed, kappa lognormal "
          "s=1 per element, f=1, homogeneous Dirichlet",
    "C4p1": "2D Darcy {n}x{n} quads H1 p=1, kappa=exp(GRF l=0.05 s=1.5), f=1",
    "C4p2": "2D Darcy {n}x{n} quads H1 p=2, kappa=exp(GRF l=0.05 s=1.5), f=1",
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(pb, iters_target, seed, tol):
    """Times the CPU oracle (a scalar C port of the reference path, 1 thread).
    Small problems (<= 300k equations) run whole: assembly + Jacobi-PCG to the
    same tolerance.  Larger ones run a bounded sample: full assembly + 10
    Jacobi-PCG iterations of a 20^3 instance of the same elasticity workload,
    scaled per element and per stored entry to the target size and to the
    GPU FP64 PCG iteration count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from neopz_amd import problems as P
    if pb.n_equations() <= 300000:
        o = O.Oracle(pb)
        t0 = time.perf_counter()
        ia, ja = o.pattern()
        a, rhs = o.assemble()
        t_asm = time.perf_counter() - t0
        t0 = time.perf_counter()
        r = O.pcg(ia, ja, a, rhs, maxit=1000000, tol=tol, wasted_spmv=True)
        t_sol = time.perf_counter() - t0
        return {"value": t_asm + t_sol, "unit": "s", "cores": 1, "kind": "port",
                "sample": f"oracle/pz_oracle.c, whole problem: pattern+assembly {t_asm:.2f}s, "
                          f"Jacobi-PCG {r['iters']} iterations {t_sol:.2f}s (reference semantics incl. "
                          f"the discarded residual SpMV)",
                "assembly_s": t_asm, "iterations": r["iters"]}
    ns = 20
    spb = P.elasticity3d(n=ns, p=2, seed=seed, field_kind="grf", corr_len=0.1,
                         load="body" if pb.bc_type[4] >= 0 else "traction")
    o = O.Oracle(spb)
    ia, ja = o.pattern()
    t0 = time.perf_counter()
    a, rhs = o.assemble()
    t_asm = time.perf_counter() - t0
    k = 10
    t0 = time.perf_counter()
    O.pcg(ia, ja, a, rhs, maxit=k, tol=0.0, wasted_spmv=True)
    t_it = (time.perf_counter() - t0) / k
    nel_s, nnz_s = ns ** 3, len(ja)
    nel_t = pb.nvol
    nnz_t = O.nnz_closed_form(3, pb.p, pb.nstate, pb.nx, pb.ny, pb.nz)
    asm_t = t_asm / nel_s * nel_t
    it_t = t_it / nnz_s * nnz_t
    total = asm_t + it_t * iters_target
    return {"value": total, "unit": "s", "cores": 1, "kind": "port",
            "sample": f"oracle/pz_oracle.c on a {ns}^3 instance ({nel_s} hexes, {nnz_s} upper "
                      f"entries): assembly {t_asm:.2f}s, Jacobi-PCG {t_it * 1e3:.1f} ms/iteration "
                      f"(reference semantics incl. the discarded residual SpMV); scaled by elements "
                      f"and stored entries to {pb.nx}^3 and x {iters_target} iterations",
            "assembly_s": asm_t, "s_per_iteration": it_t}


def traffic_of(kernel):
    """HBM bytes per launch of a kernel on the default workload, from the
    counter summary profiles/hbm_traffic_c5.txt (FETCH_SIZE and WRITE_SIZE in
    KiB per dispatch, each scaled by what the counter reads for a 1 GiB
    device copy); None when the summary has no line for the kernel."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_c5.txt")
    if not os.path.exists(path):
        return None
    scale = {"fetch": 1.0, "write": 1.0}
    found = None
    for line in open(path):
        f = dict(kv.split("=", 1) for kv in line.split() if "=" in kv)
        if "fetch_scale" in f:
            scale = {"fetch": float(f["fetch_scale"]), "write": float(f["write_scale"])}
        if f.get("kernel", "") == kernel and "fetch_KiB" in f:
            found = f
    if found is None:
        return None
    return 1024.0 * (float(found["fetch_KiB"]) * scale["fetch"] + float(found["write_KiB"]) * scale["write"])


def kernels_only(ctx, L):
    """Per-kernel timings for counter runs: FP32 block CSR SpMV, FP64 CSR SpMV,
    matrix-free operator, element assembly; then three 1 GiB device-to-device
    copies whose counters give the bytes-per-count scale of a plain copy."""
    import ctypes as C
    L.check(ctx.lib.pzg_assemble(ctx.h))
    avg, nbytes = C.c_double(), C.c_double()
    res = {}
    for which, name in ((1, "spmv_bsr32"), (0, "spmv_bsr64"), (3, "matfree"), (4, "mf_element"),
                        (2, "assembly")):
        L.check(ctx.lib.pzg_time_kernel(ctx.h, which, 5, C.byref(avg), C.byref(nbytes)))
        unit = "alg_flops" if which == 4 else "alg_bytes"
        res[name] = {"ms": avg.value, unit: nbytes.value}
        log(f"{name}: {avg.value:.4f} ms, {unit} {nbytes.value:.4e}")
    hip = C.CDLL("libamdhip64.so")
    nb = 1 << 30
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nb)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nb)):
        raise RuntimeError("hipMalloc failed")
    for _ in range(3):
        if hip.hipMemcpy(dst, src, C.c_size_t(nb), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"copy calibration: 3 x {nb} bytes device to device")
    print(json.dumps({"kernels": res, "copy_bytes": nb}))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1"}))
        sys.exit(2)
    from neopz_amd import _lib as L
    from neopz_amd import mesh as M
    from neopz_amd import problems as P
    from neopz_amd.analysis import Context, TPZSSpStructMatrixHIP, TPZStepSolver

    over = {"n": args.n} if args.n else {}
    pb = P.config(args.config, seed=args.seed, **over)
    n = pb.nx
    t0 = time.perf_counter()
    md = M.build(pb)
    log(f"mesh built: neq={md['neq']} ({time.perf_counter() - t0:.1f}s)")
    ctx = Context(0)
    sm = TPZSSpStructMatrixHIP(ctx, md)
    t0 = time.perf_counter()
    K = sm.Create()
    log(f"pattern: {ctx.sizes()} ({time.perf_counter() - t0:.1f}s)")

    if args.kernels_only:
        kernels_only(ctx, L)
        return

    def solver(kind, precond=None):
        st = TPZStepSolver(K)
        if kind == "pcg64":
            st.SetCG(200000, TPZStepSolver.jacobi(), args.tol, 0)
            st.check_every = 100
        else:
            pre = {"learned": L.PRECOND_LEARNED, "block_jacobi": L.PRECOND_BLOCK_JACOBI,
                   "multigrid": L.PRECOND_MULTIGRID}[precond or args.precond]
            op = {"auto": L.INNER_AUTO, "bsr": L.INNER_BSR, "matfree": L.INNER_MATFREE}[args.inner_op]
            st.SetMixedFCG(200000, args.tol, precond=pre, inner_tol=args.inner_tol,
                           refine=args.refine, inner_op=op, mg_degree=args.mg_degree,
                           mg_range=args.mg_range)
            # multigrid iterations are few and costly: the solver's short batches
            st.check_every = 0 if pre == L.PRECOND_MULTIGRID else 100
        st.use_graph = not args.no_graph
        return st

    def step(st):
        L.check(ctx.lib.pzg_assemble(ctx.h))
        st.Solve(None, want_solution=False)
        return st.info

    st = solver("mixed")
    for w in range(args.warmup):
        info = step(st)
        log(f"warmup {w}: {info}")
    L.check(ctx.lib.pzg_sync(ctx.h))
    times, infos = [], []
    t_all = time.perf_counter()
    for s in range(args.steps):
        t0 = time.perf_counter()
        info = step(st)
        L.check(ctx.lib.pzg_sync(ctx.h))
        times.append(time.perf_counter() - t0)
        infos.append(info)
        log(f"step {s}: {times[-1]:.3f}s {info}")
    t_all = time.perf_counter() - t_all
    ms_per_step = 1e3 * t_all / args.steps
    value = t_all / args.steps
    # solutions leave the device outside the timed regions
    u_mixed = np.zeros(md["neq"])
    L.check(ctx.lib.pzg_get_solution(ctx.h, L.ptr(u_mixed, L.dp)))
    # the block-Jacobi mixed solve on the same system, for comparison
    bj = None
    if args.precond == "multigrid" and not args.no_bj:
        stj = solver("mixed", "block_jacobi")
        t0 = time.perf_counter()
        ij = step(stj)
        L.check(ctx.lib.pzg_sync(ctx.h))
        bj = {"seconds": time.perf_counter() - t0, "iterations": ij["iterations"],
              "true_residual": ij["true_residual"]}
        log(f"mixed block Jacobi: {bj}")
    # plain FP64 Jacobi-PCG GPU path (denominator of the speed-up)
    base = None
    if not args.no_baseline:
        stb = solver("pcg64")
        t0 = time.perf_counter()
        ib = step(stb)
        L.check(ctx.lib.pzg_sync(ctx.h))
        tb = time.perf_counter() - t0
        u64 = np.zeros(md["neq"])
        L.check(ctx.lib.pzg_get_solution(ctx.h, L.ptr(u64, L.dp)))
        # the FP64 reference at full size: the FP64 PCG solution (parity-tested
        # against the CPU oracle at small sizes), both at residual 1e-8
        rel = float(np.linalg.norm(u_mixed - u64) / np.linalg.norm(u64))
        base = {"seconds": tb, "iterations": ib["iterations"], "true_residual": ib["true_residual"],
                "achieved_recursive": ib["achieved"], "mixed_solution_rel_diff": rel}
        log(f"pcg64: {tb:.3f}s {ib}, |u_mixed - u_pcg64|/|u_pcg64| = {rel:.2e}")
    # roofline of the dominant kernel (FP32 block CSR SpMV of the inner solve)
    import ctypes as C
    avg = C.c_double()
    nbytes = C.c_double()
    L.check(ctx.lib.pzg_time_kernel(ctx.h, 1, 20, C.byref(avg), C.byref(nbytes)))
    spmv_ms = avg.value
    achieved = nbytes.value / (spmv_ms * 1e-3) / 1e9
    L.check(ctx.lib.pzg_time_kernel(ctx.h, 0, 5, C.byref(avg), C.byref(nbytes)))
    spmv64_ms = avg.value
    spmv64_gbs = nbytes.value / (spmv64_ms * 1e-3) / 1e9
    L.check(ctx.lib.pzg_time_kernel(ctx.h, 2, 3, C.byref(avg), C.byref(nbytes)))
    asm_ms = avg.value
    # the matrix-free operator exists for single-rule quad/hex meshes only
    mf_ok = ctx.lib.pzg_time_kernel(ctx.h, 3, 20, C.byref(avg), C.byref(nbytes)) == L.PZG_OK
    mf_ms = mf_gbs = el_ms = el_flops = None
    if mf_ok:
        mf_ms, mf_gbs = avg.value, nbytes.value / (avg.value * 1e-3) / 1e9
        L.check(ctx.lib.pzg_time_kernel(ctx.h, 4, 20, C.byref(avg), C.byref(nbytes)))
        el_ms, el_flops = avg.value, nbytes.value
    sizes = ctx.sizes()
    # roofline of the dominant kernel of the measured path: the matrix-free
    # element kernel (FP32 contractions on the matrix cores) when the inner
    # operator is matrix-free, else the FP32 block CSR SpMV
    mf_used = args.inner_op != "bsr" and mf_ok
    # the counter summary covers the default workload only
    counted = args.config == "C5" and not args.n and args.seed == 0
    if mf_used:
        kname = (f"k_mf_hex27_mfma<{sizes['bs']}>" if pb.dim == 3 and pb.p == 2
                 else "k_mf_vol") + " (matrix-free FP32 element operator)"
        tf = el_flops / (el_ms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "kernel": kname, "achieved": tf, "peak": 157.3,
                    "unit": "TFLOP/s", "frac": tf / 157.3, "launch_ms": el_ms,
                    "flops_per_launch": el_flops,
                    "traffic": traffic_of(f"k_mf_hex27_mfma<{sizes['bs']}>") if counted else None}
    else:
        roofline = {"bound": "hbm", "kernel": f"k_spmv_bsr<float,{sizes['bs']}> (FP32 block CSR SpMV)",
                    "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                    "frac": achieved / 8000.0, "launch_ms": spmv_ms,
                    "traffic": traffic_of(f"k_spmv_bsr<float,{sizes['bs']}>") if counted else None}
    cpu = None
    if not args.no_cpu:
        its = base["iterations"] if base else infos[-1]["iterations"]
        cpu = cpu_baseline(pb, its, args.seed, args.tol)
    out = {
        "metric": "assembly+solve time to rel. residual 1e-8, learned-FCG vs Jacobi-PCG, 1 GPU",
        "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 inner / f64 refinement+assembly", "data": "synthetic",
        "config": {"workload": WORKLOADS[args.config].format(n=n) +
                               f"; mixed FCG ({args.precond}, {args.inner_op} inner operator) + "
                               f"FP64 refinement",
                   "name": args.config,
                   "neq": sizes["neq"], "nnz_upper": sizes["nnz_upper"],
                   "nnz_full": sizes["nnz_full"], "elements": sizes["nvol"], "seed": args.seed},
        "solve": infos[-1],
        "jacobi_pcg64_gpu": base,
        "mixed_block_jacobi": bj,
        "speedup_vs_pcg64": (base["seconds"] / value) if base else None,
        "assembly_ms": asm_ms,
        "elements_per_s": sizes["nvol"] / (asm_ms * 1e-3),
        "spmv": {"fp64_block_csr": {"ms": spmv64_ms, "GBps": spmv64_gbs,
                                    "frac_of_8TBps": spmv64_gbs / 8000.0},
                 "fp32_block_csr": {"ms": spmv_ms, "GBps": achieved,
                                    "frac_of_8TBps": achieved / 8000.0}},
        "matfree_operator": {"ms": mf_ms, "GBps_compulsory": mf_gbs} if mf_ok else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    tuning = L.hp_tuning(ctx)
    if tuning:
        out["hp_assembly_tuning"] = tuning
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
