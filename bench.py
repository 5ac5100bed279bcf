"""Benchmark of the Elastic solve path on one MI355X.

Workload: hex8 unit cube 100^3 (1,030,301 nodes, 3,060,300 free DOF), interior
nodes jittered by U[-0.15h, 0.15h], E = 1, nu = 0.3, z=0 clamped, traction
(0,0,-1e-3) on z=1; RCM-ordered CSRC; Jacobi PCG to a relative residual
||r||_M / ||b||_M < 1e-10.  The system is assembled on the GPU.

One step = one solve of that system with b and the matrix resident in HBM:
FP32 CSRC SpMV/preconditioner inside FP64 iterative refinement (the measured
method), timed between device synchronisations.  The plain FP64 GPU PCG is
timed beside it.  The CPU baseline is the reference's own OpenMP pcg_omp
(compiled from its sources into oracle/_ref) solving the same system on the
host cores; its solution is the FP64 oracle that both GPU results are
compared with.  Where oracle/_ref is absent the C restatement of pcg_omp
takes its place.

Prints one JSON line (see the keys below).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TOL = 1e-10
TIME_STRIDE = 4  # every 4th PCG product of the timed solves is event-timed


def spmv_bytes(neq: int, nad: int, sv: int, sx: int = 8) -> int:
    """Algorithmic bytes of one symmetric CSRC product: ia, ja, al and ad with
    sv-byte values, x and y with sx-byte entries."""
    return 4 * (neq + 1) + 4 * nad + sv * nad + sv * neq + sx * neq + sx * neq


def blocked_bytes(neq: int, nad: int, nb: int, ntiles: int, next_nodes: int, sv: int,
                  sx: int) -> int:
    """Bytes one product on the node-blocked layout must move: per off-diagonal
    node block one 4-byte word and nb*nb sv-byte values; per node the
    nb(nb+1)/2 diagonal-block values; x staged for the own rows and for every
    (tile, external node) pair, y written for the own rows and one external
    partial per (tile, external node) pair (sx bytes per entry, nb entries per
    node); per tile its share of x.(Ax) (8 bytes), its 36 bytes of descriptors,
    its rowptr/segptr (2 bytes per row and per column, plus one each) and its
    external node list (4 bytes per node)."""
    nn = neq // nb
    nblk = (nad - nn * nb * (nb - 1) // 2) // (nb * nb)
    return (nblk * (4 + sv * nb * nb) + nn * sv * nb * (nb + 1) // 2
            + sx * neq + sx * nb * next_nodes          # x: own rows, external nodes
            + sx * neq + sx * nb * next_nodes          # y and the external partials
            + 8 * ntiles + 36 * ntiles + 2 * (nn + ntiles) + 2 * (nn + next_nodes + ntiles)
            + 4 * next_nodes)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100, help="cells per cube edge")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="length of the one-core sample of the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default layout (1024-block tiles, 256 lanes)")
    ap.add_argument("--serving", type=int, default=512,
                    help="requests of the plane-strain serving stream (0: skip)")
    ap.add_argument("--serving-fixed", action="store_true",
                    help="also serve the stream with fixed FP64 and fixed mixed-precision "
                         "instances (the default stream picks per request)")
    ap.add_argument("--serving-compare", type=int, default=48,
                    help="requests in the solve-only batched vs sequential comparison")
    ap.add_argument("--plastic-compare", type=int, default=32,
                    help="J2 hex8 60^3 load paths whose tangent systems are solved "
                         "sequentially and batched (0: skip)")
    ap.add_argument("--no-tet", action="store_true",
                    help="skip the tet4 80^3 section (the ill-conditioned workload)")
    ap.add_argument("--plastic-last", type=int, default=0,
                    help="stop the J2 end-to-end run after this increment (0: all 10; "
                         "fewer for a profile)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solve path is not sharded across GPUs; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from elastic_amd import mesh as M
    from elastic_amd.solver import FP64, MIXED
    from elastic_amd.system import System

    t0 = time.time()
    mesh = M.hex8_cube(args.n, jitter=0.15, seed=M.BASE_SEED + 2)
    sysm = System(mesh)
    ctx = sysm.ctx
    neq, nad = sysm.neq, sysm.nad
    t_setup = time.time() - t0
    print("setup: neq=%d nad=%d in %.1f s" % (neq, nad, t_setup), file=sys.stderr, flush=True)
    ctx.upload_rhs(sysm.b)

    def sync():
        torch.cuda.synchronize()

    def phase(msg):
        print("%s (%.1f s)" % (msg, time.time() - t0), file=sys.stderr, flush=True)

    # reference answer on the GPU: FP64 PCG tightened to 1e-12
    st_ref = ctx.solve_resident(precision=FP64, tol=1e-12, maxit=20000)
    x_ref = ctx.download_x()

    phase("reference solve done")
    # layout autotuner (untimed set-up): tile size and lanes per tile of the
    # FP32-value operator timed inside short mixed PCG runs, then the lanes of
    # the FP64 operator on the chosen tiles
    tuned = None
    if not args.no_autotune:
        tile32, ms32 = ctx.autotune(precision=MIXED, reps=32)
        _, ms64 = ctx.autotune(precision=FP64, reps=32, tiles=False)
        tuned = {"tile_blocks": tile32, "lanes32": ctx.get_option("lanes32"),
                 "lanes64": ctx.get_option("lanes64"), "fp32_product_ms": ms32,
                 "fp64_product_ms": ms64}
        phase("autotuned: %s" % tuned)
    # plain FP64 GPU PCG (the drop-in), timed.  The product of every 4th
    # iteration is bracketed by HIP events on the solve's stream, so the
    # kernel's roofline comes from the timed solves themselves
    ctx.set_option("time_products", TIME_STRIDE)
    lat64 = []
    for i in range(args.warmup + args.steps):
        if i == args.warmup:
            ctx.product_timing(reset=True)
        sync()
        a = time.perf_counter()
        st64 = ctx.solve_resident(precision=FP64, tol=TOL, maxit=20000)
        sync()
        if i >= args.warmup:
            lat64.append(time.perf_counter() - a)
    live64_ms, live64_n = ctx.product_timing(reset=True)
    x64 = ctx.download_x()

    phase("fp64 solves timed")
    # measured method: FP32 operator inside the FP64 reliable-update PCG.  The
    # FP32 operator sums in FP64 or in FP32: both are timed once here
    # (outside the timed region) and the faster one that converges is kept
    accum = {}
    for a32 in (0, 1):
        ctx.set_option("accum32", a32)
        sync()
        a = time.perf_counter()
        st_t = ctx.solve_resident(precision=MIXED, tol=TOL, maxit=20000, allow_maxit=True)
        sync()
        if st_t["status"] == 0:
            accum[a32] = time.perf_counter() - a
    acc32 = min(accum, key=accum.get)
    ctx.set_option("accum32", acc32)
    for _ in range(args.warmup):
        ctx.solve_resident(precision=MIXED, tol=TOL, maxit=20000)
    ctx.product_timing(reset=True)
    lat = []
    sync()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        stm = ctx.solve_resident(precision=MIXED, tol=TOL, maxit=20000)
        sync()
        lat.append(time.perf_counter() - a)
    sync()
    t_total = time.perf_counter() - t_start
    live32_ms, live32_n = ctx.product_timing(reset=True)
    ctx.set_option("time_products", 0)
    xm = ctx.download_x()
    phase("mixed solves timed")
    err_m = float(np.linalg.norm(xm - x_ref) / np.linalg.norm(x_ref))
    err_64 = float(np.linalg.norm(x64 - x_ref) / np.linalg.norm(x_ref))

    # SpMV kernel roofline: the FP32 operator (the dominant kernel), its
    # average launch in the timed solves; the standalone back-to-back timing
    # (x not freshly written by the previous kernel) is reported beside it
    tot32, ker32 = ctx.time_spmv(MIXED, reps=50)
    tot64, ker64 = ctx.time_spmv(FP64, reps=50)
    # with FP32 sums the operator stages x from an FP32 copy and writes its
    # rows in FP32: x and y count 4 bytes per entry then, 8 otherwise
    b32, b64 = spmv_bytes(neq, nad, 4, 4 if acc32 else 8), spmv_bytes(neq, nad, 8)
    info = ctx.layout_info()
    nb = info["nb"]
    if nb > 1:  # the bytes the node-blocked kernel must move (the roofline's)
        m32 = blocked_bytes(neq, nad, nb, info["ntiles"], info["ext_partials"], 4,
                            4 if acc32 else 8)
        m64 = blocked_bytes(neq, nad, nb, info["ntiles"], info["ext_partials"], 8, 8)
    else:
        m32, m64 = b32, b64
    gbs32 = m32 / (live32_ms * 1e-3) / 1e9
    gbs64 = m64 / (live64_ms * 1e-3) / 1e9
    t_asm = sysm.time_assembly(reps=3)
    copy = copy_probe(torch)
    lanes32 = ctx.get_option("lanes32")  # the context is closed before the serving runs

    phase("product, assembly and copy timed")
    cpu, x_or = None, None
    if not args.no_cpu:
        cpu, x_or = cpu_baseline(sysm, st64["iters"], args.cpu_seconds)
    oracle_err = None
    if x_or is not None:
        nx = float(np.linalg.norm(x_or))
        oracle_err = {"mixed": float(np.linalg.norm(xm - x_or)) / nx,
                      "fp64": float(np.linalg.norm(x64 - x_or)) / nx,
                      "oracle": cpu["kind"] + " pcg_omp FP64 solve to %g" % TOL}

    phase("cpu baseline done")
    ctx.close()
    del sysm, ctx
    tet = None
    if not args.no_tet:
        tet = tet_section()
        phase("tet4 80^3 done")
    serving = None
    if args.serving > 0 or args.serving_compare > 0:
        serving = serving_bench(args.serving, args.serving_compare, args.serving_fixed)
    plastic = None
    if args.plastic_compare > 0:
        phase("serving done")
        plastic = plastic_compare(args.plastic_compare, last=args.plastic_last or None)

    value = neq * args.steps / t_total
    out = {
        "metric": "solved DOF/s + p50/p99 solve latency at 1e-8 rel err; CSRC SpMV GB/s vs HBM",
        "value": value,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp32 matrix values, fp32 sums in the operator, fp32 p/r/Jacobi scale in "
                  "the iteration; fp64 x, dot products and true-residual replacement"
                  if acc32 else
                  "fp32 matrix values, fp64 sums in the operator, fp64 vectors, dot products "
                  "and true-residual replacement"),
        "data": "synthetic (generated mesh, GPU-assembled system)",
        "config": {"workload": "hex8 100^3 jittered 0.15h, E=1 nu=0.3, Jacobi PCG to 1e-10 "
                               "(%d DOF, nad %d)" % (neq, nad),
                   "method": "Jacobi PCG, FP32-value CSRC SpMV, FP64 residual replacement"},
        "latency_ms_p50": 1e3 * float(np.percentile(lat, 50)),
        "latency_ms_p99": 1e3 * float(np.percentile(lat, 99)),
        "rel_err_vs_fp64": err_m,
        "rel_err_vs_oracle": oracle_err,
        "iters_mixed": stm["iters"], "refinements": stm["outer"],
        "fp64": {"dof_per_s": neq / float(np.mean(lat64)), "iters": st64["iters"],
                 "latency_ms_p50": 1e3 * float(np.percentile(lat64, 50)),
                 "rel_err_vs_tight": err_64},
        "speedup_mixed_vs_fp64": float(np.mean(lat64)) / float(np.mean(lat)),
        "operator_sums": {"chosen": "fp32" if acc32 else "fp64",
                          "trial_ms": {("fp32" if k else "fp64"): 1e3 * v for k, v in accum.items()}},
        "spmv": {"fp32_kernel_ms": live32_ms, "fp32_launches_timed": live32_n, "fp32_GBs": gbs32,
                 "fp64_kernel_ms": live64_ms, "fp64_launches_timed": live64_n, "fp64_GBs": gbs64,
                 "bytes_fp32": m32, "bytes_fp64": m64,
                 "csrc_bytes_fp32": b32, "csrc_bytes_fp64": b64,
                 "csrc_equivalent_GBs": {"fp32": b32 / (live32_ms * 1e-3) / 1e9,
                                         "fp64": b64 / (live64_ms * 1e-3) / 1e9},
                 "layout": {"nb": nb, "ntiles": info["ntiles"],
                            "external_nodes": info["ext_partials"]},
                 "standalone": {"fp32_kernel_ms": ker32, "fp32_with_sum_and_merge_ms": tot32,
                                "fp64_kernel_ms": ker64, "fp64_with_sum_and_merge_ms": tot64}},
        "autotune": tuned,
        "assembly_ms": t_asm,
        "setup_s": t_setup,
        "roofline": {"bound": "hbm", "achieved": gbs32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs32 / HBM_PEAK_GBS, "bytes_per_launch": m32,
                     "traffic": measured_traffic(neq, nad, m32),
                     "traffic_source": ("profiles/spmv_traffic_hex8_100.json: rocprofv3 FETCH_SIZE "
                                        "and WRITE_SIZE of this kernel on a layout with these "
                                        "model bytes, scaled by the copy probe's readings"),
                     "kernel": ("k_spmv_blocks<float,%s,%d,PER,%d lanes> with tile shares of p.q"
                                % ("float" if acc32 else "double", nb,
                                   lanes32))
                               if nb > 1 else "k_spmv_tiles<float,double>"},
        "copy_probe": copy,
        "cpu_baseline": cpu,
        "serving": serving,
        "batched_3d": plastic,
        "tet4_80": tet,
        "speedups_vs_plain_fp64": speedup_summary(float(np.mean(lat64)) / float(np.mean(lat)),
                                                  serving, plastic, tet),
    }
    print(json.dumps(out), flush=True)


def speedup_summary(single: float, serving, plastic, tet=None):
    """The ratios against plain FP64 GPU PCG that the mixed and batched forms
    reach, gathered from the sections above (None where a section was
    skipped)."""
    out = {"hex8_100_single_system_mixed": single}
    so = (serving or {}).get("solve_only")
    if so:
        out["plane_strain_48_batched_policy"] = so["batched_policy"]["speedup_vs_sequential_fp64"]
        out["plane_strain_48_batched_mixed"] = so["batched_mixed"]["speedup_vs_sequential_fp64"]
    if tet:
        out["tet4_80_best_mixed_vs_fp64_jacobi"] = tet["best_mixed_vs_fp64_jacobi"]
    if plastic:
        out["j2_hex8_60_32_systems_batched_mixed_fp32_sums"] = (
            plastic["batched_mixed_fp32_sums"]["speedup_vs_sequential_fp64"])
    return out


def tet_section():
    """The tet4 workload (80^3 cells split into 6 tetrahedra, jitter 0.2h,
    E = 10^U[-1,1] per element, nu = 0.45; 1,574,640 equations): PCG to 1e-10
    in FP64 and with the FP32 operator inside the FP64 reliable-update loop,
    with Jacobi and with the 3x3 node-block Jacobi (block_precond), each
    solve timed once between device synchronisations; errors against an FP64
    solve to 1e-12."""
    import torch

    from elastic_amd import mesh as M
    from elastic_amd.solver import BDIAG3, DIAG, FP64, MIXED
    from elastic_amd.system import System

    s = System(M.config_mesh(3))
    c = s.ctx
    c.upload_rhs(s.b)
    c.solve_resident(precision=FP64, tol=1e-12, maxit=200000)
    xref = c.download_x()
    nref = float(np.linalg.norm(xref))
    out = {"workload": "tet4 80^3 x 6 jittered 0.2h, E 10^U[-1,1], nu 0.45 (%d DOF), PCG to %g"
                       % (s.neq, TOL)}
    for pname, pc in (("jacobi", DIAG), ("block_jacobi_3x3", BDIAG3)):
        for name, prec in (("fp64", FP64), ("mixed", MIXED)):
            c.solve_resident(precond=pc, precision=prec, tol=TOL, maxit=200000)  # warm-up
            torch.cuda.synchronize()
            a = time.perf_counter()
            st = c.solve_resident(precond=pc, precision=prec, tol=TOL, maxit=200000)
            torch.cuda.synchronize()
            t = time.perf_counter() - a
            err = float(np.linalg.norm(c.download_x() - xref)) / nref
            out["%s_%s" % (name, pname)] = {"ms": 1e3 * t, "iters": st["iters"],
                                            "dof_per_s": s.neq / t, "rel_err_vs_tight": err}
    best = min(out["mixed_jacobi"]["ms"], out["mixed_block_jacobi_3x3"]["ms"])
    out["best_mixed_vs_fp64_jacobi"] = out["fp64_jacobi"]["ms"] / best
    c.close()
    return out


def serving_bench(nreq: int, ncompare: int, fixed: bool = False):
    """The plane-strain request stream (half quad4, half tria3, side U{32..512},
    Poisson arrivals at 200 requests/s) served by the batch with continuous
    admission, precision and preconditioner chosen per request
    (serving.request_policy; with `fixed` also all-FP64 and all-mixed
    streams): throughput and p50/p99 latency from arrival, from admission
    and from submission to x on the host (host set-up and GPU assembly
    included).  Then the solve-only comparison on the first `ncompare`
    requests with every system resident: plain FP64 PCG solved one after the
    other (the drop-in path) against the batch with FP64, mixed and policy
    instances."""
    from elastic_amd.serving import Batch, MeshRequest, Server, request_policy, request_stream
    from elastic_amd.solver import FP64, MIXED
    from elastic_amd.system import System

    workers = min(16, os.cpu_count() or 1)
    out = {"workload": ("%d plane-strain requests (even quad4, odd tria3, side U{32..512}), "
                        "Poisson arrivals 200/s, to %g" % (nreq, TOL))}
    streams = [("stream_batched_policy", None)] + ([("stream_batched_fp64", FP64),
                                                     ("stream_batched_mixed", MIXED)]
                                                    if fixed else [])
    for name, prec in streams:
        if nreq <= 0:
            break
        srv = Server(max_slots=64, iters_per_step=16, prep_workers=workers, precision=prec,
                     tol=TOL, hbm_budget=64e9, progress=sys.stderr)
        a = time.perf_counter()
        res = srv.run(request_stream(nreq, rate=200.0), realtime=True)
        print("%s: %d requests in %.1f s" % (name, res["requests"], time.perf_counter() - a),
              file=sys.stderr, flush=True)
        out[name] = {k: v for k, v in res.items() if k != "done"}
        out[name]["fp64_operator_handovers"] = int(sum(r.stats.get("op64_from", -1) >= 0
                                                       for r in res["done"]))

    if ncompare <= 0:
        return out
    print("solve-only comparison: preparing %d systems" % ncompare, file=sys.stderr, flush=True)
    systems = []
    for k in range(ncompare):
        s_ = System(MeshRequest(k)())
        s_.ctx.upload_rhs(s_.b)
        systems.append(s_)
    dof = float(sum(s_.neq for s_ in systems))
    torch_sync()
    a = time.perf_counter()
    it64 = 0
    for s_ in systems:
        st = s_.ctx.solve_resident(precision=FP64, tol=TOL, maxit=50000)
        it64 += st["iters"]
    t_seq = time.perf_counter() - a
    print("sequential fp64: %.1f s" % t_seq, file=sys.stderr, flush=True)
    x64 = [s_.ctx.download_x() for s_ in systems]
    compare = {"requests": ncompare, "dof": dof,
               "sequential_fp64": {"s": t_seq, "dof_per_s": dof / t_seq, "iters": it64}}
    policy = [request_policy(s_.ctx) for s_ in systems]
    for name, prec in (("batched_fp64", FP64), ("batched_mixed", MIXED),
                       ("batched_policy", None)):
        for s_, (pp, pc) in zip(systems, policy):
            s_.ctx.set_option("batch_precond", pc if prec is None else 2)
        a = time.perf_counter()
        done, itb = {}, 0
        with Batch(0, max_slots=64, iters_per_step=16) as b:
            queue = list(range(len(systems)))
            while len(done) < len(systems):
                while queue and b.submit(systems[queue[0]].ctx,
                                         policy[queue[0]][0] if prec is None else prec, TOL,
                                         50000, queue[0]):
                    queue.pop(0)
                b.step()
                for tag, st in b.poll():
                    done[tag] = st
                    itb += st["iters"]
        t_bat = time.perf_counter() - a
        err = max(float(np.linalg.norm(s_.ctx.download_x() - x64[i]) / np.linalg.norm(x64[i]))
                  for i, s_ in enumerate(systems))
        print("%s: %.1f s" % (name, t_bat), file=sys.stderr, flush=True)
        compare[name] = {"s": t_bat, "dof_per_s": dof / t_bat, "iters": itb,
                         "speedup_vs_sequential_fp64": t_seq / t_bat,
                         "max_rel_err_vs_fp64": err}
    for s_ in systems:
        s_.ctx.close()
    out["solve_only"] = compare
    return out


def plastic_compare(npaths: int, n: int = 60, increments: int = 10, last=None):
    """The J2 workload (hex8 60^3 between grips, E = 1, nu = 0.3, H = 0.05,
    yield stresses U[5e-4, 2e-3], load in 10 increments), run end to end: the
    paths' batched mixed-precision Newton iterations (GPU re-assembly of K_t
    and R, batched solves to 1e-11) through all increments, timed.  After
    increment 1 the run pauses (outside the timed spans) for a solve-only
    comparison: every path assembles its tangent system of increment 2's
    first Newton iteration, and those systems are solved to 1e-10 by plain
    FP64 PCG one after the other, and together in a batch in FP64 and in
    mixed precision.  The run then resumes from the committed increment 1
    (the comparison leaves u and the committed state untouched)."""
    from elastic_amd.plasticity import compare_solves, load_paths, run_paths
    from elastic_amd.serving import Batch
    from elastic_amd.solver import MIXED

    a = time.perf_counter()
    paths, _ = load_paths(n, npaths)
    t_setup = time.perf_counter() - a
    neq = paths[0].sys.neq
    counts1, counts2 = {}, {}
    slots = max(1, min(npaths, 255))
    with Batch(0, max_slots=slots, iters_per_step=16) as b:
        a = time.perf_counter()
        run_paths(paths, increments=increments, precision=MIXED, batch=b, last=1,
                  counts=counts1, progress=sys.stderr)
        t_inc1 = time.perf_counter() - a
    # plastic Gauss points of the committed state (committing the converged
    # state again changes nothing and returns the count)
    plastic = [int(p.commit()[1]) for p in paths]
    for p in paths:
        p.increment(2.0 / increments)
        p.assemble(lhs=True)
    print("plastic compare: %d systems of %d equations" % (npaths, neq),
          file=sys.stderr, flush=True)
    out = compare_solves(paths, tol=TOL)
    out["workload"] = ("J2 hex8 %d^3 tangent systems of increment 2 (first Newton iteration), "
                       "%d load paths, %d equations each" % (n, npaths, neq))
    out["setup_s"] = t_setup
    out["plastic_points_after_increment_1"] = {"min": min(plastic), "max": max(plastic),
                                               "of": 8 * paths[0].mesh.numel}
    with Batch(0, max_slots=slots, iters_per_step=16) as b:
        a = time.perf_counter()
        run_paths(paths, increments=increments, precision=MIXED, batch=b, first=2, last=last,
                  counts=counts2, progress=sys.stderr)
        t_rest = time.perf_counter() - a
    wall = t_inc1 + t_rest
    solves = counts1["linear_solves"] + counts2["linear_solves"]
    plastic_end = [int(p.commit()[1]) for p in paths]
    out["end_to_end"] = {
        "workload": "J2 hex8 %d^3 pulled 5e-3 between grips, %d load paths, sy U[5e-4, 2e-3], "
                    "H 0.05, %d increments%s, Newton to 1e-8, linear solves to 1e-11"
                    % (n, npaths, increments,
                       "" if not last or last >= increments else " (stopped after %d)" % last),
        "method": "batched mixed-precision PCG (FP32 operator, FP64 residual replacement), "
                  "GPU re-assembly of K_t and R every Newton iteration",
        "wall_s": wall, "linear_solves": solves,
        "line_search_halvings": counts1["halvings"] + counts2["halvings"],
        "solved_dof_per_s": neq * solves / wall,
        "plastic_points_at_end": {"min": min(plastic_end), "max": max(plastic_end),
                                  "of": 8 * paths[0].mesh.numel}}
    for p in paths:
        p.close()
    return out


def torch_sync():
    import torch

    torch.cuda.synchronize()


def measured_traffic(neq: int, nad: int, model_bytes: int):
    """HBM bytes per launch of the FP32 operator (FETCH_SIZE and WRITE_SIZE
    counter runs scaled by the copy-probe calibration), recorded for this
    workload in profiles/spmv_traffic_hex8_100.json together with the model
    bytes of the layout they were measured on.  None for another workload or
    when the live layout's model bytes differ from the recorded ones (the
    counters would describe another kernel)."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic_hex8_100.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    if ("neq %d" % neq) not in d["workload"] or ("nad %d" % nad) not in d["workload"]:
        return None
    if d["fp32"].get("model_bytes") != model_bytes:
        return None
    return d["fp32"]["read_bytes"] + d["fp32"]["write_bytes"]


def copy_probe(torch):
    """A plain device copy of known size (y = -x over 2^27 doubles: 1 GiB read,
    1 GiB written) timed with events on torch's stream; it is also the
    calibration point of the FETCH_SIZE / WRITE_SIZE counter runs."""
    n = 1 << 27
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.neg(x, out=y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record()
    for _ in range(reps):
        torch.neg(x, out=y)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 16 * n
    del x, y
    return {"bytes": nbytes, "ms": ms, "GBs": nbytes / (ms * 1e-3) / 1e9}


def host_cpu():
    """Host processor name and the number of CPUs the OS reports."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def reference_solve(sysm, ad, al, nth: int, tol: float, maxit: int):
    """The reference's own solv -> pcg_omp (nth > 1) or pcg (nth = 1),
    compiled from its sources into oracle/_ref/ref_harness, on this system
    (written to a scratch directory): (solution, solv seconds, iterations),
    or None when the harness is absent or fails."""
    import re
    import shutil
    import subprocess
    import tempfile

    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(exe):
        return None
    d = tempfile.mkdtemp(prefix="elastic_ref_")
    try:
        sysd, outd = os.path.join(d, "sys"), os.path.join(d, "out")
        os.makedirs(sysd)
        os.makedirs(outd)
        with open(os.path.join(sysd, "sys.txt"), "w") as fh:
            fh.write("%d %d\n" % (sysm.neq, sysm.nad))
        sysm.ia.astype(np.int32).tofile(os.path.join(sysd, "ia.i4"))
        sysm.ja.astype(np.int32).tofile(os.path.join(sysd, "ja.i4"))
        ad.tofile(os.path.join(sysd, "ad.f8"))
        al.tofile(os.path.join(sysd, "al.f8"))
        sysm.b.tofile(os.path.join(sysd, "b.f8"))
        # pool words: the system, 7 vectors, nth per-thread y buffers, slack
        words = sysm.nad * 3 + sysm.neq * (2 * (7 + nth) + 8) + (1 << 24)
        env = dict(os.environ, OMP_NUM_THREADS=str(nth))
        r = subprocess.run([exe, sysd, outd, repr(tol), str(maxit), str(nth), str(128 | 1),
                            str(words)], capture_output=True, text=True, env=env,
                           timeout=900)
        if r.returncode != 0 or not os.path.exists(os.path.join(outd, "x.f8")):
            return None
        meta = open(os.path.join(outd, "meta.txt")).read().split()
        m = re.search(r"it\s+(\d+)", open(os.path.join(outd, "solver.log")).read())
        x = np.fromfile(os.path.join(outd, "x.f8"), dtype=np.float64)
        return x, float(meta[3]), int(m.group(1)) if m else -1
    except (OSError, subprocess.SubprocessError, ValueError):
        return None
    finally:
        shutil.rmtree(d, ignore_errors=True)


def cpu_baseline(sysm, iters_fp64: int, seconds: float):
    """The reference's CPU path on the same system, on the host cores: its
    own OpenMP pcg_omp (oracle/_ref) solving to 1e-10 with the threads the
    host gives this GPU (OMP_NUM_THREADS, else every CPU), and its serial pcg
    restated (the oracle's or_pcg, bit-identical to the reference's) timed on
    one core over a bounded number of iterations.  Returns (baseline, the
    reference's solution x or None).  Without oracle/_ref the C restatement
    of pcg_omp stands in ("port")."""
    from oracle import binding as O

    nth = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    model, ncpu = host_cpu()
    ad, al = sysm.values()
    b = sysm.b
    x_ref = None
    ref = reference_solve(sysm, ad, al, nth, TOL, 20000)
    if ref is not None:
        x_ref, t, k = ref
        kind = "reference"
        sample = ("the whole FP64 solve to %g by the reference's own pcg_omp (Iters_omp.f, "
                  "built from its sources), %d iterations, %d threads, %.1f s" % (TOL, k, nth, t))
    else:
        m = 1.0 / ad
        a = time.perf_counter()
        x_ref, st = O.pcg(sysm.ia, sysm.ja, ad, al, b, tol=TOL, maxit=20000, m=m, nth=nth)
        t = time.perf_counter() - a
        k = st["iters"]
        kind = "port"
        sample = ("the whole FP64 solve to %g by the C restatement of pcg_omp, %d iterations, "
                  "%d threads, %.1f s" % (TOL, k, nth, t))
    # one core: the serial pcg (Iters.f:252-524) restated bit for bit, a
    # sample of about `seconds` of iterations extrapolated to the FP64
    # iteration count
    a = time.perf_counter()
    O.pcg(sysm.ia, sysm.ja, ad, al, b, tol=1e-30, maxit=2, m=1.0 / ad, nth=0)
    ks = int(min(200, max(5, seconds / ((time.perf_counter() - a) / 2.0))))
    a = time.perf_counter()
    O.pcg(sysm.ia, sysm.ja, ad, al, b, tol=1e-30, maxit=ks, m=1.0 / ad, nth=0)
    t1 = (time.perf_counter() - a) / ks
    base = {"value": sysm.neq / t, "unit": "DOF/s", "cores": nth, "kind": kind,
            "sample": sample, "s_per_iter": t / max(k, 1), "iters": k,
            "host": {"cpu": model, "nproc": ncpu, "threads_used": nth},
            "serial_1core": {"value": sysm.neq / (t1 * iters_fp64), "unit": "DOF/s",
                             "s_per_iter": t1, "kind": "port (or_pcg, bit-identical to pcg)",
                             "sample": "%d iterations on one core, extrapolated to the %d "
                                       "iterations of the FP64 solve" % (ks, iters_fp64)}}
    return base, x_ref


if __name__ == "__main__":
    main()
