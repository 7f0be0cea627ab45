"""Benchmark of the MI355X element K+R assembly (one JSON line on stdout).

Workload: a batch of design variants of 3D linear elasticity (E = 1, nu = 0.3), each a 64^3
trilinear hex8 lattice of the unit cube with a 4 x 4 x 4 array of spherical holes (radius
0.1 * U[0.9, 1.1] per sphere and variant, seeded), cut by the XTK restatement; the material is the
region outside the spheres.  A step is one replay of the HIP graph that assembles K and R of every
element of every variant (uncut and cut) into each variant's CSR values and R, with u, the set
data and the CSR patterns resident in HBM:

  * measured path (`value`): mixed precision + learned cut-cell quadrature (network prior, FP64
    moment fit, 27-point FP32 MFMA integrands, subdivision fallback in FP32 with FP64 sub-cell
    Jacobians, FP64 global values);
  * beside it: the plain FP64 path (XTK subdivision, TET_11 on every sub-tet) on the same sets,
    captured the same way; `speedup_vs_fp64_gpu` is the ratio of the two.

The learned path's K and R are compared with the FP64 path's for every variant (the FP64 path
matches the CPU restatement to 1e-12 in the test suite).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--variants 32] [--n 64] [--no-cpu]
       python bench.py --workload hex27 [--n 32] [--refine 1]   (the quadratic HEX27 configuration,
                                                                    optionally with one THB level; its own line)
       python bench.py --workload c1|c2 [--n N]                 (the 2D circle / 3D sphere-diffusion sets)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import moris_amd as M  # noqa: E402
from moris_amd import learned as L  # noqa: E402
from moris_amd import train_lq  # noqa: E402
from moris_amd import workloads as W  # noqa: E402

METRIC = "elements/s for K+R assembly, mixed+learned vs FP64 plain, ‖ΔK‖/‖K‖≤1e-6"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
SEED = W.BASE_SEED


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# HBM bytes per launch of k_stencil<3,3,64,false,float> on variant 0 of the 64^3 workload, measured
# with two rocprofv3 --pmc runs (FETCH_SIZE scaled x2.0, WRITE_SIZE x1.0 by the copy-kernel
# calibration): 97,555 KiB x 2 + 445,915 KiB.
STENCIL_TRAFFIC_BYTES = (97555 * 2 + 445915) * 1024
# The same for k_generic_hex27<3,double,false> on the 32^3 HEX27 gyroid set (gather mode):
# 109,644 KiB fetched, 1,470,336 KiB written (the K_e pool).
HEX27_32_TRAFFIC_BYTES = (109644 * 2 + 1470336) * 1024
# The same for k_generic_hex27<3,double,true,false> on the 48^3-base THB gyroid set (scatter mode,
# FP64 atomic adds), summed over the 42 colour launches of one call: 4,423,474 KiB fetched,
# 70,966,197 KiB written (the atomics count as write requests only).
THB48_SCATTER_TRAFFIC_BYTES = (4423474 * 2 + 70966197) * 1024
# The same for the FP64 k_stencil<3,1,ROWS,DIRECT,double> of the 128^3 diffusion set, per stencil
# variant (the autotuner's choice): (FETCH_SIZE, WRITE_SIZE) KiB per launch.
C2_STENCIL_TRAFFIC_KIB = {0: (100537, 428126), 1: (94277, 433217), 2: (68008, 426869),
                          3: (102682, 471827), 4: (95728, 464089), 5: (68942, 459620)}


def stencil_bytes(spec, S, n_elements, n_cut, ke_bytes, nf=3, nb=8, dim=3):
    """Algorithmic HBM bytes of one k_stencil launch: K values and R written once, u read once,
    per-element material (8 B per parameter), the stored cut-element K_e (FP32 on the mixed path,
    FP64 on the plain path), lattice maps and row offsets."""
    na = nf * nb
    n_nodes = (spec.n[0] + 1) * (spec.n[1] + 1) * ((spec.n[2] + 1) if dim == 3 else 1)
    n_cells = spec.n[0] * spec.n[1] * (spec.n[2] if dim == 3 else 1)
    n_mat = 1 if nf == 1 else 2
    return (8 * S.nnz + 8 * S.n_rows + 8 * S.n_rows + 8 * n_mat * n_elements + ke_bytes * na * na * n_cut
            + 4 * n_cells + 4 * n_nodes + 4 * (S.n_rows + 1))


def elasticity_flops(n_ip, n_elem, dim=3, nb=8):
    """Native-shape flops of the reference algorithm: per point D.B, B^T.DB and the residual
    products; per element the T projection (dense T of size n_p x n_a)."""
    ns = 6 if dim == 3 else 3
    npd = dim * nb
    f_ip = 2 * ns * ns * npd + 2 * npd * npd * ns + 2 * (ns * npd + ns * ns + npd * ns)
    f_t = 2 * npd * npd * npd + 2 * npd * npd * npd + 2 * npd * npd
    return n_ip * f_ip + n_elem * f_t, f_ip, f_t


def fit_flops_of(n_points, n_cut, n_moments=23, n_points_rule=27):
    """Algorithmic flops of the moment fit: per sub-cell point the IG -> IP map (4 vertices x 3
    coordinates) and the accumulation of the moments; per element the three FP64 products with
    V / P (residual, correction, final residual)."""
    return n_points * (2 * 12 + 2 * n_moments) + n_cut * 3 * 2 * n_moments * n_points_rule


def host_cpu_info():
    """The host the CPU baseline runs on: CPU model, os.cpu_count(), the CPUs this process may run
    on, OMP_NUM_THREADS, and the worker count used (the CPUs granted to this job: the affinity set,
    capped by OMP_NUM_THREADS where the job's share is set that way)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    except ValueError:
        omp = 0
    cores = min(aff, omp) if omp > 0 else aff
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": aff, "omp_num_threads": omp or None,
            "cores": max(1, cores)}


def cpu_baseline(P, S, u, budget_s=20.0, probe_n=4000, rule="TET_11", budget_1core_s=8.0):
    """Oracle (FP64 C++ restatement of the reference algorithm, its -O3 -march=x86-64-v3 build) on
    the host cores granted to this job, on a strided sample of one variant's elements sized to the
    time budget; then the same on one core on a smaller sample."""
    from oracle import py as O
    O.use_bench_build()
    host = host_cpu_info()
    cores = host["cores"]
    rp, ci = S.pattern()
    ne = P.n_elements
    elems = np.arange(ne, dtype=np.int64)
    probe = elems[:: max(1, ne // probe_n)]
    t0 = time.perf_counter()
    O.assemble_list(P.desc, rp, ci, u, probe, n_threads=cores)
    per_elem = (time.perf_counter() - t0) / len(probe)

    def timed(n_threads, budget):
        stride = max(1, int(np.ceil(ne * per_elem * (cores / n_threads) / (budget / 2))))
        sample = elems[::stride]
        runs = []
        t_end = time.perf_counter() + budget
        while not runs or (time.perf_counter() < t_end and len(runs) < 3):
            t1 = time.perf_counter()
            O.assemble_list(P.desc, rp, ci, u, sample, n_threads=n_threads)
            runs.append(time.perf_counter() - t1)
        cut_frac = float(np.mean(np.diff(P.arrays()["ig_ptr"])[sample] > 0))
        return len(sample) / min(runs), ("one element in %d of variant 0 (%d of %d elements, cut fraction %.3f), "
                                         "FP64 subdivision (%s), best of %d runs, %d worker thread%s%s" % (
                                             stride, len(sample), ne, cut_frac, rule, len(runs), n_threads,
                                             "s" if n_threads > 1 else "",
                                             " with private CSR arrays + row-wise reduction" if n_threads > 1 else ""))

    value, sample = timed(cores, budget_s)
    value1, sample1 = timed(1, budget_1core_s)
    out = {"value": value, "unit": "elements/s", "cores": cores, "kind": "port", "sample": sample,
           "value_1core": value1, "sample_1core": sample1,
           "build": "oracle/liboracle_bench.so: g++ -O3 -march=x86-64-v3 (built off the GPU host)"}
    out.update({k: v for k, v in host.items() if k != "cores"})
    return out


def run_hex27(args):
    """HEX27 IP cells with p = 2 B-splines (the quadratic configuration): 3D elasticity on an n^3
    lattice of the unit cube cut by a gyroid of period 4 cells, HEX_3x3x3 on whole cells and TET_35
    on the XTK sub-tets, generic element path.  One JSON line: FP64 and mixed-precision assembly
    rates, the moment kernel's MFMA rate, mixed-vs-FP64 error, and the oracle on a sample."""
    n = args.n if args.n_given else (16 if args.refine else 32)
    spec, keep = M.make_spec(3, (n, n, n), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_GYROID,
                             gyroid_scale=2 * np.pi * n / 4.0, bspline_order=2, lagrange_order=2,
                             refine_levels=args.refine)
    t0 = time.time()
    P = M.Problem(spec, keep)
    log("hex27 problem %.1fs: %d^3, elements %d, cut %d, sub-tets %d" % (time.time() - t0, n, P.n_elements, P.n_cut,
                                                                         P.n_ig))
    S = M.Set(P.desc, keepalive=P)
    t_host = time.time() - t0
    # sets whose element matrices exceed 2^28 entries assemble in scatter mode (no K_e pool);
    # MGPU_GENERIC_SCATTER=1 / =0 forces it on / off
    ke_entries = int(np.sum(np.diff(S.topology_offsets()).astype(np.int64) ** 2))
    force = os.environ.get("MGPU_GENERIC_SCATTER", "")[:1]
    mode = "scatter" if force == "1" or (ke_entries > 2 ** 28 and force != "0") else "gather"
    log("hex27 setup %.1fs: %d^3, elements %d, cut %d, sub-tets %d, rows %d, nnz %d, %s mode" % (
        time.time() - t0, n, P.n_elements, P.n_cut, P.n_ig, S.n_rows, S.nnz, mode))
    u = np.random.default_rng(SEED + 27).uniform(-1.0, 1.0, S.n_rows)
    ub, vb, rb = M.DeviceBuffer.from_array(u), M.DeviceBuffer(8 * S.nnz), M.DeviceBuffer(8 * S.n_rows)
    ig = np.diff(P.arrays()["ig_ptr"])
    npts = np.where(ig > 0, ig * 35, 27)
    # executed MFMA work: per 64-point chunk the 21 upper 16 x 16 tiles of M x 16 k-steps of 4 points;
    # per element and dof-type block h >= f (6 blocks, the other 3 mirrored) the 2 n_t + n_t^2 tiles
    # of K_p T and T^T (K_p T) x 7 k-steps (n_t = ceil(adofs per type / 16))
    n_t = (np.diff(S.topology_offsets()) // 3 + 15) // 16
    proj_flops = float(np.sum(6 * (2 * n_t + n_t * n_t)) * 7 * 2 * 16 * 16 * 4)
    mfma_flops = float(np.sum(np.ceil(npts / 64.0)) * 21 * 16 * 2 * 16 * 16 * 4) + proj_flops
    # the moment-fitted rule (cut elements): only the weights on MFMA, 2 tiles x 16 k-steps per
    # 64-point chunk of the TET_35 points (M by sum factorisation on the VALU)
    mfma_flops_fit = float(np.sum(np.where(ig > 0, np.ceil(npts / 64.0) * 2 * 16,
                                           np.ceil(npts / 64.0) * 21 * 16)) * 2 * 16 * 16 * 4) + proj_flops
    native, _, _ = elasticity_flops(int(npts.sum()), P.n_elements, nb=27)
    out = {"metric": "elements/s for K+R assembly, HEX27 p=2 elasticity (gyroid cut), generic path",
           "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "higher_is_better": True,
           "data": "synthetic", "config": {"workload": "%d^3 HEX27 lattice, p=2 B-splines%s, gyroid period 4 cells, "
                                                       "HEX_3x3x3 / TET_35" % (n, ", one THB level on the intersected "
                                                                                  "cells" if args.refine else ""),
                                           "elements": P.n_elements, "cut_elements": P.n_cut,
                                           "ig_subcells": P.n_ig, "points": int(npts.sum()), "nnz": S.nnz,
                                           "assembly": mode, "element_matrix_entries": ke_entries}}
    res = {}
    legs = ((M.PREC_FP64, M.QUAD_SUBDIVISION, "fp64"), (M.PREC_MIXED, M.QUAD_SUBDIVISION, "mixed"),
            (M.PREC_FP64, M.QUAD_MOMENT_FIT, "fp64_moment_fit"), (M.PREC_MIXED, M.QUAD_MOMENT_FIT, "mixed_moment_fit"))
    for prec, quad, name in legs:
        S.options(prec, quad)
        S.autotune(ub, vb, rb, iters=2)   # element order of the gather mode
        per, st = S.time_assemble(ub, vb, rb, max(3, args.steps))
        res[name] = (vb.to_array(np.float64, S.nnz), rb.to_array(np.float64, S.n_rows))
        flops = mfma_flops_fit if quad == M.QUAD_MOMENT_FIT else mfma_flops
        out[name] = {"value": P.n_elements / (per * 1e-3), "ms_per_call": per,
                     "ms_element_kernel": st[2], "ms_gather": st[3], "ms_rhs": st[4],
                     "element_kernel_mfma_TFLOP/s": flops / (st[2] * 1e-3) / 1e12,
                     "element_kernel_native_TFLOP/s": native / (st[2] * 1e-3) / 1e12}
        log("hex27 %s: %.3f ms per call, stages %s" % (name, per, ["%.4f" % x for x in st]))
    for name in ("mixed", "fp64_moment_fit", "mixed_moment_fit"):
        out[name]["rel_err_K_vs_fp64"] = float(np.linalg.norm(res[name][0] - res["fp64"][0]) /
                                               np.linalg.norm(res["fp64"][0]))
    out["value"] = out["mixed_moment_fit"]["value"]
    out["value_path"] = "mixed_moment_fit"
    out["dtype"] = "f32+f64"
    out["max_rel_err_K_mixed_vs_fp64"] = out["mixed"]["rel_err_K_vs_fp64"]
    out["mfma_flops_per_call"] = mfma_flops
    out["mfma_flops_per_call_moment_fit"] = mfma_flops_fit
    out["native_flops_per_call"] = native
    out["roofline"] = {"bound": "mfma", "kernel": "k_generic_hex27<3,double>",
                       "achieved": out["fp64"]["element_kernel_mfma_TFLOP/s"], "peak": FP64_PEAK_TFLOPS,
                       "unit": "TFLOP/s", "frac": out["fp64"]["element_kernel_mfma_TFLOP/s"] / FP64_PEAK_TFLOPS,
                       "traffic": None}
    if n == 32 and not args.refine and mode == "gather":
        out["roofline"]["traffic"] = HEX27_32_TRAFFIC_BYTES
        out["roofline"]["traffic_source"] = "rocprofv3 --pmc FETCH_SIZE (x2.0) + WRITE_SIZE (x1.0) per launch"
    elif n == 48 and args.refine == 1 and mode == "scatter":
        out["roofline"]["traffic"] = THB48_SCATTER_TRAFFIC_BYTES
        out["roofline"]["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE (x2.0) + WRITE_SIZE (x1.0), summed over "
                                             "the colour launches of one call; the atomic adds count as write "
                                             "requests only, so this is a lower bound")
    if not args.refine:
        # the set-up of the same set on the GPU (cut, numbering, T rows at the 27 nodes, pattern)
        from moris_amd import gpu_setup as G
        Gs = G.GpuSetup(spec, keep)
        Gs.close()
        t1 = time.time()
        Gs = G.GpuSetup(spec, keep)
        wall = time.time() - t1
        # the generic set is made on the host from the set-up's descriptor (its gather lists are
        # built there): the rest of a re-meshing loop's cost
        t1 = time.time()
        S_g = M.Set(Gs.desc, keepalive=Gs)
        t_set = time.time() - t1
        out["setup"] = {"gpu_device_ms": Gs.ms, "gpu_wall_ms": 1e3 * wall, "host_problem_and_set_s": t_host,
                        "host_set_from_gpu_setup_s": t_set,
                        "same_counts_as_host": [Gs.n_elements, Gs.n_ig, Gs.nnz] == [P.n_elements, P.n_ig, S.nnz]}
        log("hex27 GPU set-up: %.2f ms device, %.2f ms wall (host problem + set %.2f s); set from it %.2f s" % (
            Gs.ms, 1e3 * wall, t_host, t_set))
        S_g.close()
        Gs.close()
    if not args.no_cpu:
        log("hex27 cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(P, S, u, budget_s=15.0, probe_n=200, rule="TET_35")
    print(json.dumps(out), flush=True)


def run_hex27_blocks(args):
    """C3's whole domain on one GPU: the n^3-base lattice of the unit cube (one THB level on the
    intersected cells, HEX27 cells, p = 2 rows, gyroid of period 4 cells) split into blocks^3
    sub-domains of (n / blocks)^3 base cells, each assembled as its own set and freed before the
    next: the blocks a multi-GPU run puts one per GPU, here in sequence on one GPU.  Rows of adofs on
    block faces hold each block's partial sums (a global matrix adds them).  The time per call is
    the sum over the blocks.  One JSON line."""
    n, nbk = args.n, args.blocks
    if n % nbk:
        raise SystemExit("--n must be a multiple of --blocks")
    m = n // nbk
    hcell = 1.0 / n
    legs = ((M.PREC_FP64, M.QUAD_SUBDIVISION, "fp64"), (M.PREC_MIXED, M.QUAD_SUBDIVISION, "mixed"),
            (M.PREC_FP64, M.QUAD_MOMENT_FIT, "fp64_moment_fit"), (M.PREC_MIXED, M.QUAD_MOMENT_FIT, "mixed_moment_fit"))
    totals = {name: 0.0 for _, _, name in legs}
    elements = cut = subtets = nnz = 0
    t_setup = 0.0
    for bk in range(nbk ** 3):
        ijk = (bk % nbk, (bk // nbk) % nbk, bk // (nbk * nbk))
        lo = tuple(c * m * hcell for c in ijk)
        hi = tuple(l + m * hcell for l in lo)
        spec, keep = M.make_spec(3, (m, m, m), lo=lo, hi=hi, iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_GYROID,
                                 gyroid_scale=2 * np.pi * n / 4.0, bspline_order=2, lagrange_order=2,
                                 refine_levels=args.refine)
        t0 = time.time()
        P = M.Problem(spec, keep)
        S = M.Set(P.desc, keepalive=P)
        t_setup += time.time() - t0
        u = np.random.default_rng(SEED + 27 + bk).uniform(-1.0, 1.0, S.n_rows)
        ub, vb, rb = M.DeviceBuffer.from_array(u), M.DeviceBuffer(8 * S.nnz), M.DeviceBuffer(8 * S.n_rows)
        line = []
        for prec, quad, name in legs:
            S.options(prec, quad)
            S.autotune(ub, vb, rb, iters=1)
            per, _ = S.time_assemble(ub, vb, rb, max(2, args.steps // 5))
            totals[name] += per
            line.append("%s %.1f ms" % (name, per))
        elements += P.n_elements
        cut += P.n_cut
        subtets += P.n_ig
        nnz += S.nnz
        log("block %d/%d %s: elements %d, cut %d, nnz %d, set-up %.1fs; %s" % (
            bk + 1, nbk ** 3, ijk, P.n_elements, P.n_cut, S.nnz, time.time() - t0, ", ".join(line)))
        del ub, vb, rb, S, P
    out = {"metric": "elements/s for K+R assembly, HEX27 p=2 elasticity (gyroid cut)%s, whole domain as blocks" % (
               ", one THB level" if args.refine else ""),
           "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "higher_is_better": True, "data": "synthetic",
           "config": {"workload": "%d^3 HEX27 base lattice%s, p=2 B-splines, gyroid period 4 cells, as %d^3 blocks "
                                  "of %d^3 base cells assembled in sequence" % (
                                      n, ", one THB level on the intersected cells" if args.refine else "", nbk, m),
                      "elements": elements, "cut_elements": cut, "ig_subcells": subtets, "nnz_sum_over_blocks": nnz,
                      "host_setup_s": t_setup},
           "value": elements / (totals["mixed_moment_fit"] * 1e-3), "value_path": "mixed_moment_fit",
           "dtype": "f32+f64"}
    for _, _, name in legs:
        out[name] = {"value": elements / (totals[name] * 1e-3), "ms_per_call": totals[name]}
    print(json.dumps(out), flush=True)


def run_lattice(args):
    """One lattice set (stencil path) at a parity configuration: c1 = 64 x 64 QUAD4 plane-stress
    elasticity around a circular hole, c2 = 128^3 HEX8 diffusion around a spherical hole with a
    log-normal conductivity.  FP64 and mixed-precision subdivision; one JSON line."""
    if args.workload == "c1":
        n = args.n if args.n_given else 64
        spec, keep = W.circle_2d_spec(n)
        dim, nf, nb, rule = 2, 2, 4, "TRI_7"
    else:
        n = args.n if args.n_given else 128
        spec, keep = W.sphere_diffusion_spec(n)
        dim, nf, nb, rule = 3, 1, 8, "TET_11"
    t0 = time.time()
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    log("%s setup %.1fs: elements %d, cut %d, sub-cells %d, rows %d, nnz %d" % (
        args.workload, time.time() - t0, P.n_elements, P.n_cut, P.n_ig, S.n_rows, S.nnz))
    u = np.random.default_rng(SEED + 29).uniform(-1.0, 1.0, S.n_rows)
    ub, vb, rb = M.DeviceBuffer.from_array(u), M.DeviceBuffer(8 * S.nnz), M.DeviceBuffer(8 * S.n_rows)
    out = {"metric": "elements/s for K+R assembly, %s" % ("2D QUAD4 plane-stress elasticity (circle cut)"
                                                         if dim == 2 else "3D HEX8 diffusion (sphere cut)"),
           "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "higher_is_better": True, "data": "synthetic",
           "config": {"workload": "%s, %s lattice of the unit %s" % (args.workload, "x".join([str(n)] * dim),
                                                                    "square" if dim == 2 else "cube"),
                      "elements": P.n_elements, "cut_elements": P.n_cut, "ig_subcells": P.n_ig, "nnz": S.nnz}}
    res = {}
    for prec, name in ((M.PREC_FP64, "fp64"), (M.PREC_MIXED, "mixed")):
        S.options(prec, M.QUAD_SUBDIVISION)
        S.autotune(ub, vb, rb, iters=3)
        per, st = S.time_assemble(ub, vb, rb, max(5, args.steps))
        res[name] = vb.to_array(np.float64, S.nnz)
        st_bytes = stencil_bytes(spec, S, P.n_elements, P.n_cut, ke_bytes=8 if prec == M.PREC_FP64 else 4, nf=nf, nb=nb,
                                 dim=dim)
        out[name] = {"value": P.n_elements / (per * 1e-3), "ms_per_call": per, "stage_ms": st,
                     "k_stencil_GB/s": st_bytes / (st[4] * 1e-3) / 1e9, "k_stencil_bytes": st_bytes,
                     "tuning": list(S.tuning())}
        log("%s %s: %.4f ms per call, stages %s" % (args.workload, name, per, ["%.4f" % x for x in st]))
    out["value"] = out["fp64"]["value"]
    out["dtype"] = "f64"
    out["max_rel_err_K_mixed_vs_fp64"] = float(np.linalg.norm(res["mixed"] - res["fp64"]) / np.linalg.norm(res["fp64"]))
    gbs = out["fp64"]["k_stencil_GB/s"]
    traffic = None
    if args.workload == "c2" and n == 128:
        f_kib, w_kib = C2_STENCIL_TRAFFIC_KIB[out["fp64"]["tuning"][0]]
        traffic = (f_kib * 2 + w_kib) * 1024
    out["roofline"] = {"bound": "hbm", "kernel": "k_stencil (FP64 path)", "achieved": gbs, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": traffic}
    if traffic is not None:
        out["roofline"]["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE (x2.0) + WRITE_SIZE (x1.0) per launch of "
                                             "the tuned stencil variant")
    if not args.no_cpu:
        log("%s cpu baseline ..." % args.workload)
        out["cpu_baseline"] = cpu_baseline(P, S, u, budget_s=15.0, rule=rule)
    print(json.dumps(out), flush=True)


def timed_replays(B, steps, warmup):
    for _ in range(warmup):
        B.launch()
    B.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        B.launch()
    B.synchronize()
    return (time.perf_counter() - t0) / steps


def gpu_setup_leg(problems, sets, args, host_problem_s, host_set_s):
    """Per-variant set-up (level set, XTK cut, numbering, T rows, topology, CSR pattern) on the GPU
    beside the host set-up of the same variants; variant 0's arrays and pattern compared."""
    from moris_amd import gpu_setup as G
    dev_ms, wall_s, e2e_s = [], 0.0, 0.0
    for k, (P, _) in enumerate(problems):
        M.synchronize()
        t1 = time.perf_counter()
        Gs = G.GpuSetup(P.spec, P._keep)
        wall_s += time.perf_counter() - t1
        dev_ms.append(Gs.ms)
        # the whole per-design pipeline on the device: set-up, set from its arrays, one assembly
        M.synchronize()
        t1 = time.perf_counter()
        G2 = G.GpuSetup(P.spec, P._keep)
        D = G2.make_set()
        ub = M.DeviceBuffer.from_array(np.zeros(D.n_rows))
        vb, rb = M.DeviceBuffer(8 * D.nnz), M.DeviceBuffer(8 * D.n_rows)
        D.assemble_device(ub, vb, rb)
        M.synchronize()
        e2e_s += time.perf_counter() - t1
        del D, G2, ub, vb, rb
        if k == 0:
            a, b = P.arrays(), Gs.arrays()
            same = all(np.array_equal(a[key].view(np.uint8), b[key].view(np.uint8)) for key in a)
            rp, ci = sets[0].pattern()
            grp, gci = Gs.pattern()
            same = same and np.array_equal(rp, grp) and np.array_equal(ci, gci)
        Gs.close()
    nv = len(problems)
    log("set-up per variant: host %.1f ms (problem %.1f + set %.1f), GPU %.2f ms device, %.1f ms wall; set-up + set "
        "from its device arrays + first FP64 assembly %.1f ms; variant 0 identical: %s" % (
            1e3 * (host_problem_s + host_set_s) / nv, 1e3 * host_problem_s / nv, 1e3 * host_set_s / nv,
            float(np.median(dev_ms)), 1e3 * wall_s / nv, 1e3 * e2e_s / nv, same))
    return {"gpu_device_ms_per_variant_median": float(np.median(dev_ms)),
            "gpu_wall_ms_per_variant": 1e3 * wall_s / nv,
            "gpu_setup_set_first_assembly_ms_per_variant": 1e3 * e2e_s / nv,
            "host_ms_per_variant": 1e3 * (host_problem_s + host_set_s) / nv,
            "host_problem_ms_per_variant": 1e3 * host_problem_s / nv,
            "host_set_create_ms_per_variant": 1e3 * host_set_s / nv,
            "variant0_identical_to_host": bool(same),
            "note": "GPU: mgpu_setup_build (device time by HIP events; wall includes allocation). The "
                    "per-design pipeline: mgpu_setup_build + mgpu_set_create_from_setup + one FP64 assembly, "
                    "wall clock. Host: mgpu_problem_build on the host's threads + mgpu_set_create (pattern and "
                    "uploads). Not part of the timed assembly."}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--variants", type=int, default=32)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--branches", type=int, default=1,
                    help="parallel graph branches (0: autotuned); with 1 every kernel runs alone, so its "
                         "duration under a profiler matches the per-stage event timings")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--refine", type=int, default=0, choices=(0, 1),
                    help="hex27 workload: one THB level on the intersected cells")
    ap.add_argument("--blocks", type=int, default=1,
                    help="hex27 workload: split the n^3 domain into blocks^3 sets assembled in sequence")
    ap.add_argument("--workload", choices=("variants", "hex27", "c1", "c2"), default="variants",
                    help="variants: the headline batch (default); hex27: the quadratic HEX27 configuration; "
                         "c1: 64x64 QUAD4 plane-stress elasticity; c2: 128^3 HEX8 diffusion")
    args = ap.parse_args()
    args.n_given = "--n" in sys.argv
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if M.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "hex27":
        if args.blocks > 1:
            run_hex27_blocks(args)
        else:
            run_hex27(args)
        return
    if args.workload in ("c1", "c2"):
        run_lattice(args)
        return

    # network: the committed checkpoint, else trained here with the library's kernels
    net = train_lq.load()
    net_note = "checkpoint moris_amd/data/lq_mlp.f32"
    if net is None:
        log("no network checkpoint; training on 200000 cells ...")
        net, info = train_lq.train(samples=200_000, epochs=8, log=log)
        net_note = "trained in this run on %d cells, %d epochs" % (info["samples"], info["epochs"])

    t0 = time.time()
    problems, sets, bufs = [], [], []
    host_problem_s, host_set_s = 0.0, 0.0
    for k in range(args.variants):
        spec, keep = W.elasticity_spec(args.n, W.sphere_array(k))
        t1 = time.perf_counter()
        P = M.Problem(spec, keep)
        t2 = time.perf_counter()
        S = M.Set(P.desc, keepalive=P)
        host_problem_s += t2 - t1
        host_set_s += time.perf_counter() - t2
        L.attach(S, net)
        u = np.random.default_rng(SEED + 100 + k).uniform(-1.0, 1.0, S.n_rows)
        problems.append((P, u))
        sets.append(S)
        bufs.append((M.DeviceBuffer.from_array(u), M.DeviceBuffer(8 * S.nnz), M.DeviceBuffer(8 * S.n_rows)))
    n_el = sum(P.n_elements for P, _ in problems)
    n_cut = sum(P.n_cut for P, _ in problems)
    n_ig = sum(P.n_ig for P, _ in problems)
    nnz = sum(S.nnz for S in sets)
    log("setup %.1fs: %d variants, elements %d cut %d sub-cells %d nnz %d" % (time.time() - t0, args.variants, n_el,
                                                                              n_cut, n_ig, nnz))
    setup = gpu_setup_leg(problems, sets, args, host_problem_s, host_set_s)

    def run_mode(prec, quad, label):
        for S, (ub, vb, rb) in zip(sets, bufs):
            S.options(prec, quad)
            S.autotune(ub, vb, rb, iters=3)
        # per-stage durations with HIP events on the set's stream (variant 0, eager calls)
        per_call, stages = sets[0].time_assemble(*bufs[0], max(5, args.steps))
        B = M.Batch(sets, [b[0] for b in bufs], [b[1] for b in bufs], [b[2] for b in bufs], n_branches=args.branches)
        sec = timed_replays(B, args.steps, args.warmup)
        ev_ms, n_br = B.time(max(3, args.steps // 2))
        B.close()
        log("%s: %.3f ms per replay (%.3f by events, %d branches), stages of variant 0 %s" % (
            label, 1e3 * sec, ev_ms, n_br, ["%.4f" % t for t in stages]))
        return sec, ev_ms, n_br, per_call, stages

    # plain FP64 path
    sec64, ev64, nb64, call64, st64 = run_mode(M.PREC_FP64, M.QUAD_SUBDIVISION, "fp64 subdivision")
    ref = [(vb.to_array(np.float64, S.nnz), rb.to_array(np.float64, S.n_rows)) for S, (_, vb, rb) in zip(sets, bufs)]
    tune64 = sets[0].tuning()
    # mixed + learned
    sec, evl, nbl, call_l, st_l = run_mode(M.PREC_MIXED, M.QUAD_LEARNED, "mixed + learned")
    tune_l = sets[0].tuning()
    err_k, err_r, accepted = [], [], 0
    for S, (_, vb, rb), (v0, r0) in zip(sets, bufs, ref):
        v = vb.to_array(np.float64, S.nnz)
        r = rb.to_array(np.float64, S.n_rows)
        err_k.append(float(np.linalg.norm(v - v0) / np.linalg.norm(v0)))
        err_r.append(float(np.linalg.norm(r - r0) / np.linalg.norm(r0)))
        accepted += int(L.result(S)[2].sum())
    fl0 = L.result(sets[0])[2].copy()
    # the same path with the network replaced by the volume-scaled Gauss prior (a constant network:
    # HEX_3x3x3 weights, scaled to each element's volume moment by the fit), to show what the trained
    # network changes: acceptance, time, error
    gauss = L.gauss_prior()
    for S in sets:
        L.attach(S, gauss)
    sec_g, _, _, _, st_g = run_mode(M.PREC_MIXED, M.QUAD_LEARNED, "mixed + Gauss prior")
    err_kg, acc_g = [], 0
    for S, (_, vb, rb), (v0, r0) in zip(sets, bufs, ref):
        err_kg.append(float(np.linalg.norm(vb.to_array(np.float64, S.nnz) - v0) / np.linalg.norm(v0)))
        acc_g += int(L.result(S)[2].sum())
    for S in sets:
        L.attach(S, net)
    # the method without a network at all: the 27 weights from each element's degree-2 Lagrange
    # moments (MGPU_QUAD_MOMENT_FIT, exact for the trilinear stiffness, no fit and no fallback)
    sec_m, _, _, _, st_m = run_mode(M.PREC_MIXED, M.QUAD_MOMENT_FIT, "mixed + moment fit")
    err_km = [float(np.linalg.norm(vb.to_array(np.float64, S.nnz) - v0) / np.linalg.norm(v0))
              for S, (_, vb, rb), (v0, r0) in zip(sets, bufs, ref)]
    del ref

    value = n_el / sec
    value64 = n_el / sec64
    # dominant kernel of the measured path (variant 0, per-stage events: [0] inputs + network,
    # [1] k_lq_moments + k_lq_project, [2] cut-element integrands, [3] k_prep, [4] k_stencil) and its roofline
    P0, _ = problems[0]
    S0 = sets[0]
    st_bytes = stencil_bytes(P0.spec, S0, P0.n_elements, P0.n_cut, ke_bytes=4)
    n_learned0 = int(fl0.sum())
    ig = np.diff(P0.arrays()["ig_ptr"])
    fallback_ip = int(ig[ig > 0][fl0 == 0].sum()) * 11
    cut_flops_l, f_ip, f_t = elasticity_flops(27 * n_learned0 + fallback_ip, P0.n_cut)
    cut_flops_64, _, _ = elasticity_flops(P0.n_ig * 11, P0.n_cut)
    fit_flops = fit_flops_of(P0.n_ig * 11, P0.n_cut)
    # executed MFMA work: the cut-element kernel issues 3 tiles (16 x 16 x 4) per step of 4 points
    # (ceil(points / 4) steps per element); the learned element kernel 3 tiles x 7 steps per element
    steps = np.ceil(ig[ig > 0] * 11 / 4.0)
    mfma_64 = float(steps.sum()) * 3 * 2048
    mfma_l = float(n_learned0) * 3 * 7 * 2048 + float(steps[fl0 == 0].sum()) * 3 * 2048
    kernels = {
        "k_stencil": {"ms": st_l[4], "bound": "hbm", "algorithmic_bytes": st_bytes,
                      "GB/s": st_bytes / (st_l[4] * 1e-3) / 1e9},
        "k_lq_moments+k_lq_project": {"ms": st_l[1], "bound": "fp64 (VALU moments, MFMA fit)",
                                      "algorithmic_flops": fit_flops,
                                      "TFLOP/s": fit_flops / (st_l[1] * 1e-3) / 1e12 if st_l[1] > 0 else None},
        "cut_integrands_learned": {"ms": st_l[2], "kernels": "k_lq_element<3,float> + k_cut_elasticity_mfma<float> "
                                                             "(fallback)",
                                   "executed_mfma_flops": mfma_l,
                                   "TFLOP/s": mfma_l / (st_l[2] * 1e-3) / 1e12 if st_l[2] > 0 else None,
                                   "peak_TFLOP/s": FP32_PEAK_TFLOPS,
                                   "points": 27 * n_learned0 + fallback_ip,
                                   "reference_shape_flops": cut_flops_l},
        "k_mlp_forward": {"ms": st_l[0]},
        "k_prep": {"ms": st_l[3]},
        "cut_integrands_fp64": {"ms": st64[2], "kernels": "k_cut_elasticity_mfma<double>",
                                "executed_mfma_flops": mfma_64,
                                "TFLOP/s": mfma_64 / (st64[2] * 1e-3) / 1e12 if st64[2] > 0 else None,
                                "peak_TFLOP/s": FP64_PEAK_TFLOPS,
                                "frac": mfma_64 / (st64[2] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS if st64[2] > 0 else None,
                                "points": P0.n_ig * 11, "reference_shape_flops": cut_flops_64,
                                "reference_shape_flops_per_point": f_ip, "reference_shape_flops_T_per_element": f_t},
        "note": "TFLOP/s figures are executed work (MFMA instructions issued) per kernel time; "
                "reference_shape_flops count the reference algorithm's dense per-point products and T projection "
                "and are not a utilisation figure",
    }
    dom = max(("k_stencil", st_l[4]), ("k_lq_moments+k_lq_project", st_l[1]), ("cut_integrands_learned", st_l[2]), key=lambda kv: kv[1])[0]
    if dom == "k_stencil":
        gbs = kernels["k_stencil"]["GB/s"]
        roof = {"bound": "hbm", "kernel": dom, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs / HBM_PEAK_GBS, "traffic": STENCIL_TRAFFIC_BYTES if args.n == 64 else None,
                "traffic_source": "rocprofv3 --pmc FETCH_SIZE (x2.0) + WRITE_SIZE (x1.0) per launch, variant 0"}
    elif dom == "k_lq_moments+k_lq_project":
        tf = kernels[dom]["TFLOP/s"]
        roof = {"bound": "mfma", "kernel": dom, "achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tf / FP64_PEAK_TFLOPS, "traffic": None}
    else:
        tf = kernels[dom]["TFLOP/s"]
        roof = {"bound": "mfma", "kernel": dom, "achieved": tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tf / FP32_PEAK_TFLOPS, "traffic": None}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "%d design variants of 3D linear elasticity (E=1, nu=0.3), %d^3 hex8 unit cube, "
                               "4x4x4 sphere-array holes r=0.1*U[0.9,1.1] (XTK cut), mixed precision + learned "
                               "cut-cell quadrature, one HIP graph per step" % (args.variants, args.n),
                   "elements": n_el, "cut_elements": n_cut, "ig_subcells": n_ig, "nnz": nnz,
                   "network": net_note, "fit_tolerance": 1e-6, "cancellation_bound": 4.0},
        "fp64_plain": {"value": value64, "ms_per_step": 1e3 * sec64, "graph_branches": nb64,
                       "tuning_variant0": list(tune64), "stage_ms_variant0": st64},
        "speedup_vs_fp64_gpu": value / value64,
        "learned_accepted": accepted / max(n_cut, 1),
        "moment_fit_leg": {"value": n_el / sec_m, "ms_per_step": 1e3 * sec_m, "max_rel_err_K_vs_fp64": max(err_km),
                           "stage_ms_variant0": st_m,
                           "note": "no network: the 27 HEX_3x3x3 weights from each cut element's degree-2 Lagrange "
                                   "moments over its TET_11 points (MGPU_QUAD_MOMENT_FIT), every element accepted; "
                                   "not the metric's path"},
        "gauss_prior_leg": {"value": n_el / sec_g, "ms_per_step": 1e3 * sec_g, "learned_accepted": acc_g / max(n_cut, 1),
                            "max_rel_err_K_vs_fp64": max(err_kg), "stage_ms_variant0": st_g,
                            "note": "network replaced by the volume-scaled HEX_3x3x3 Gauss prior, same fit and checks"},
        "max_rel_err_K_vs_fp64": max(err_k),
        "max_rel_err_R_vs_fp64": max(err_r),
        "graph_branches": nbl,
        "tuning_variant0": list(tune_l),
        "stage_ms_variant0": st_l,
        "event_ms_per_replay": evl,
        "roofline": roof,
        "kernels": kernels,
        "setup": setup,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        P, u = problems[0]
        out["cpu_baseline"] = cpu_baseline(P, sets[0], u)
        out["gpu_vs_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
