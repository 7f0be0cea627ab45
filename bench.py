"""Benchmark of the MI355X element K+R assembly (one JSON line on stdout).

A step is one mgpu_set_assemble call: K and R of every element of the workload (uncut and
XTK-cut), written into the CSR values / R arrays, with u, the set data and the CSR pattern already
resident in HBM.  Workload: 3D steady heat diffusion, 128^3 trilinear hex8 on the unit cube, sphere
hole r = 0.3 at the centre cut by the XTK restatement (HEX_2x2x2 uncut, TET_11 on sub-tets), per-element
log-normal conductivity (mu 0, sigma 0.5, seeded), plain FP64 path.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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

METRIC = "elements/s for K+R assembly, mixed+learned vs FP64 plain, ‖ΔK‖/‖K‖≤1e-6"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
SEED = 20261202


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(n=128):
    return M.make_spec(3, (n, n, n), iwg=M.IWG_SPATIALDIFF_BULK, spheres=[[0.5, 0.5, 0.5, 0.3]], k_mu=0.0,
                       k_sigma=0.5, seed=SEED + 1)


def algorithmic_figures(P, S, dim=3, nf=1):
    """Algorithmic bytes of the stencil (scatter) kernel and flops of the cut-element kernel."""
    nb = 8 if dim == 3 else 4
    na = nf * nb
    nm = 2 if nf > 1 else 1
    n_nodes = (P.spec.n[0] + 1) * (P.spec.n[1] + 1) * (P.spec.n[2] + 1 if dim == 3 else 1)
    stencil_bytes = (8 * S.nnz + 8 * S.n_rows            # K values and R written once
                     + 8 * S.n_rows                       # u read
                     + 8 * nm * P.n_elements              # material
                     + 8 * na * na * P.n_cut              # cut-element K_e read back
                     + 4 * P.n_cells + 4 * n_nodes        # lattice cell->element, node->adof maps
                     + 4 * (S.n_rows + 1))                # row offsets
    n_ip_cut = P.n_ig * (11 if dim == 3 else 7)
    if nf == 1:
        f_ip = 2 * nb * nb * dim + 4 * nb * dim
        npd = nb
    else:
        ns = 6 if dim == 3 else 3
        npd = nf * nb
        f_ip = 2 * ns * ns * npd + 2 * npd * npd * ns + 2 * (ns * npd + ns * ns + npd * ns)
    f_t = 2 * npd * npd * na + 2 * npd * na * na + 2 * npd * na
    cut_flops = n_ip_cut * f_ip + P.n_cut * f_t
    return stencil_bytes, cut_flops, n_ip_cut, f_ip, f_t


def cpu_baseline(P, S, u, budget_s=20.0):
    """Oracle (FP64 C++ restatement of the reference algorithm) on the host cores, whole workload."""
    from oracle import py as O
    cores = min(os.cpu_count() or 1, 16)
    rp, ci = S.pattern()
    elems = np.arange(P.n_elements, dtype=np.int64)
    # bounded sample: time a strided subset first, grow to the whole set if the budget allows
    sample = elems[:: max(1, P.n_elements // 200000)]
    t0 = time.perf_counter()
    O.assemble_list(P.desc, rp, ci, u, sample, n_threads=cores)
    dt = time.perf_counter() - t0
    est_full = dt * P.n_elements / len(sample)
    if est_full * 2 < budget_s:
        sample = elems
    runs = []
    t_end = time.perf_counter() + budget_s
    while not runs or (time.perf_counter() < t_end and len(runs) < 3):
        t0 = time.perf_counter()
        O.assemble_list(P.desc, rp, ci, u, sample, n_threads=cores)
        runs.append(time.perf_counter() - t0)
    best = min(runs)
    cut_frac = float(np.mean(np.diff(P.arrays()["ig_ptr"])[sample] > 0))
    return {"value": len(sample) / best, "unit": "elements/s", "cores": cores, "kind": "port",
            "sample": "%d of %d elements (%s), cut fraction %.4f, best of %d runs, %d worker threads with "
                      "private CSR arrays + row-wise reduction" % (
                          len(sample), P.n_elements, "whole workload" if len(sample) == P.n_elements else "strided",
                          cut_frac, len(runs), cores)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if M.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    t0 = time.time()
    spec, keep = workload(args.n)
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    log("setup %.1fs: elements %d cut %d sub-cells %d rows %d nnz %d" % (time.time() - t0, P.n_elements, P.n_cut,
                                                                          P.n_ig, S.n_rows, S.nnz))
    u = np.random.default_rng(SEED + 9).uniform(-1.0, 1.0, S.n_rows)
    ub = M.DeviceBuffer.from_array(u)
    vb = M.DeviceBuffer(8 * S.nnz)
    rb = M.DeviceBuffer(8 * S.n_rows)

    for _ in range(args.warmup):
        S.assemble_device(ub, vb, rb)
    M.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        S.assemble_device(ub, vb, rb)
    M.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = P.n_elements / (wall / args.steps)

    # per-kernel durations with HIP events on the set's stream
    per_call, km = S.time_assemble(ub, vb, rb, max(10, args.steps // 2))
    stencil_bytes, cut_flops, n_ip_cut, f_ip, f_t = algorithmic_figures(P, S)
    t_cut, t_st = km[0], km[1]
    st_gbs = stencil_bytes / (t_st * 1e-3) / 1e9 if t_st > 0 else 0.0
    cut_tf = cut_flops / (t_cut * 1e-3) / 1e12 if t_cut > 0 else 0.0
    kernels = {
        "k_stencil": {"ms": t_st, "bound": "hbm", "algorithmic_bytes": stencil_bytes, "GB/s": st_gbs,
                      "frac": st_gbs / HBM_PEAK_GBS},
        "k_cut_moments": {"ms": t_cut, "bound": "fp64", "algorithmic_flops": cut_flops, "TFLOP/s": cut_tf,
                          "frac": cut_tf / FP64_PEAK_TFLOPS, "integration_points": n_ip_cut,
                          "flops_per_ip": f_ip, "flops_T_per_element": f_t},
    }
    if t_st >= t_cut:
        roof = {"bound": "hbm", "kernel": "k_stencil", "achieved": st_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": st_gbs / HBM_PEAK_GBS, "traffic": None}
    else:
        roof = {"bound": "mfma", "kernel": "k_cut_moments", "achieved": cut_tf, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": cut_tf / FP64_PEAK_TFLOPS, "traffic": None}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D steady heat diffusion, %d^3 hex8 unit cube, sphere hole r=0.3 (XTK cut, TET_11), "
                               "log-normal k (sigma 0.5), plain FP64 GPU path" % args.n,
                   "elements": P.n_elements, "cut_elements": P.n_cut, "ig_subcells": P.n_ig, "rows": S.n_rows,
                   "nnz": S.nnz, "precision": "fp64", "quadrature": "subdivision"},
        "roofline": roof,
        "kernels": kernels,
        "event_ms_per_call": per_call,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(P, S, u)
        out["gpu_vs_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
