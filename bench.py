"""Benchmark: grid points/s for rho, grad rho, lap rho (+ tau_PD) on one MI355X.

Workload (default): the 60-atom organic-like system -- C30 N3 O3 H24,
cc-pVDZ-like pure shells, M = 624 basis functions -- on its cubic grid
(0.2 bohr spacing, 5 bohr padding, ~2.8M points), synthetic data from a fixed
seed (no wavefunction files or network).  A step is one pass of the hot path
over the whole grid: AO values + first derivatives + Laplacians, the fused
P-contraction and the FP64 epilogue producing rho, grad rho, lap rho and
tau_PD.  Points, outputs and P are resident in HBM before timing starts.

Prints one JSON line (see --help for options).  `roofline` is the dominant
kernel (the contraction): algorithmic flops 2 M^2 n_col per point (n_col = 4
GEMM columns: phi and its three first derivatives) divided by its average
launch time measured with HIP events on the library's stream; `cpu_baseline`
is the FP64 host oracle (OpenMP) on a bounded seeded subsample of the same
grid.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16": 2.5e15}

# HBM bytes per point of a contraction call (FETCH_SIZE + WRITE_SIZE,
# calibrated against a 1 GiB device copy: FETCH x 2.0, WRITE x 1.0), measured
# with rocprofv3 --pmc on this workload; see
# profiles/contract_mixed_{factor,density}_organic60.md and
# profiles/contract_f64_organic60.md for the commands and raw counts.  The
# traffic of a call is this figure times its points (the AO reads dominate
# and scale with the points).  Default (screened factor route, one call over
# all 2 666 298 points): 7.68e6 + 0.1875e6 KiB calibrated = 15.92 GB per call.
# The mixed routes other than the default were measured per call over 8 GiB
# chunks (mean 533 260 points) and are given per point the same way; the FP64
# density route over its two calls of a step (2.0 M and 0.67 M points):
# FETCH 6.798e7 + 2.298e7 KiB, WRITE 1.254e5 + 0.413e5 KiB.
MEASURED_TRAFFIC_PER_POINT = {("organic60", "mixed", "factor"): (2 * 7.68e6 + 0.1875e6) * 1024 / 2666298,
                              ("organic60", "mixed", "factor-unscreened"): (8.44e9 + 0.038e9) / 533260,
                              ("organic60", "mixed", "density"): (24.1e9 + 0.19e9) / 533260,
                              ("organic60", "fp64", "density"): ((6.798e7 + 2.298e7) * 2048 +
                                                                  (1.254e5 + 0.4128e5) * 1024) / 2666298}


def _traffic(workload, precision, key, points):
    per_point = MEASURED_TRAFFIC_PER_POINT.get((workload, precision, key))
    return None if per_point is None else per_point * points


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="organic60")
    ap.add_argument("--precision", default="auto", choices=["auto", "fp64", "mixed"])
    ap.add_argument("--route", default="auto", choices=["auto", "density", "factor"],
                    help="mixed contraction route (contract with P, or with its pivoted Cholesky factor)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU-baseline work (0 disables)")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64-vs-mixed comparison run")
    ap.add_argument("--workspace-mb", type=int, default=0, help="per-chunk workspace budget (0: library default)")
    ap.add_argument("--screening", type=float, default=None, metavar="EPS",
                    help="spatial screening threshold of the factor route (0 turns it off; "
                         "default: the library's, 1e-15)")
    ap.add_argument("--no-extra", action="store_true", help="skip the fwd+bwd and ESP secondary measurements")
    ap.add_argument("--part", default=None, choices=["esp", "fwd_bwd", "autograd", "large300", "host"],
                    help="run only this secondary measurement and print its JSON (for profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run a device-to-device copy of MB MiB (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


def cpu_baseline(w, target_s):
    from oracle import oracle

    flat = w.basis.flat()
    rng = np.random.default_rng(12345)
    n_cal = 256
    sample = w.points[rng.choice(w.points.shape[0], size=n_cal, replace=False)]
    t0 = time.perf_counter()
    oracle.density(flat, w.P, sample, rho=True, grad=True, lap=True, tau=True)
    t_cal = time.perf_counter() - t0
    n = int(min(w.points.shape[0], max(n_cal, n_cal * target_s / max(t_cal, 1e-6))))
    sample = w.points[rng.choice(w.points.shape[0], size=n, replace=False)]
    t0 = time.perf_counter()
    oracle.density(flat, w.P, sample, rho=True, grad=True, lap=True, tau=True)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "points/s", "cores": oracle.max_threads(), "kind": "port",
            "cpu_model": _cpu_model(),
            "sample": f"{n} seeded random points of the {w.name} grid, rho+grad+lap+tau_PD in FP64 "
                      f"({dt:.1f} s, OpenMP, {oracle.max_threads()} threads)"}


def _cpu_model():
    """The host CPU's model name (what lscpu reports), from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def fwd_bwd(torch, capi, synthetic, precision, reps=3):
    """Differentiation workload: 40 atoms, M = 416, 5e5 points; rho forward
    plus adjoint (dL/dP, dL/dR) per step through the C ABI, inputs resident in
    HBM.  Also the adjoint kernels' time per call (in-library HIP events)."""
    w = synthetic.diff40()
    order = synthetic.morton_order(w.points)
    pts = torch.from_numpy(np.ascontiguousarray(w.points[order])).cuda()
    wt = torch.from_numpy(np.ascontiguousarray(w.extra["weights"][order])).cuda()
    n = pts.shape[0]
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    rho = torch.empty(n, dtype=torch.float64, device="cuda")
    dP = torch.empty((w.basis.nbasis,) * 2, dtype=torch.float64, device="cuda")
    dR = torch.empty((w.basis.natom, 3), dtype=torch.float64, device="cuda")

    def step():
        h.eval(pts.data_ptr(), n, {"rho": rho.data_ptr()}, precision)
        h.backward(pts.data_ptr(), n, wt.data_ptr(), dP.data_ptr(), dR.data_ptr(), precision)

    step()
    torch.cuda.synchronize()
    h.enable_profiling(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    rate = n * reps / (time.perf_counter() - t0)
    st = h.kernel_stats()
    h.enable_profiling(False)
    return rate, st["backward_ms"] / max(st["backward_calls"], 1)


def fwd_bwd_autograd(torch, synthetic, precision, reps=3, move=False):
    """The differentiation workload through torch.autograd: L = sum_g w_g
    (rho_g(P, R) - rho_t,g)^2 with rho_t from the perturbed P_target (FP64,
    before timing), L.backward() for dL/dP and dL/dR per step.  Mixed: the
    forward holds its error to 1e-6 of |rho - rho_t| (rho_target), so the
    gradients meet 1e-6.  move=True changes P and R every step (an optimiser's
    loop: the handle refactorises P and refreshes the screening tables)."""
    from chemtoolscuda_amd.autograd import DensityModel

    w = synthetic.diff40()
    order = synthetic.morton_order(w.points)
    pts = torch.from_numpy(np.ascontiguousarray(w.points[order])).cuda()
    wts = torch.from_numpy(np.ascontiguousarray(w.extra["weights"][order])).cuda()
    n = pts.shape[0]
    ref = DensityModel(w.basis, w.charges, precision="fp64")
    model = DensityModel(w.basis, w.charges, precision=precision)
    R0 = torch.from_numpy(w.basis.atom_xyz.copy()).cuda()
    with torch.no_grad():
        rho_t = ref.density(torch.from_numpy(w.extra["P_target"]).cuda(), R0, pts)
    P = torch.from_numpy(w.P.copy()).cuda().requires_grad_(True)
    R = R0.clone().requires_grad_(True)
    tgt = rho_t if precision == "mixed" else None

    def step(k):
        if move:
            with torch.no_grad():
                P.add_(1e-6 * (k + 1))
                R.add_(1e-4)
        P.grad = None
        R.grad = None
        loss = torch.sum(wts * (model.density(P, R, pts, rho_target=tgt) - rho_t) ** 2)
        loss.backward()

    step(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(reps):
        step(k + 1)
    torch.cuda.synchronize()
    rate = n * reps / (time.perf_counter() - t0)
    rep = model.last_report or {}
    return rate, rep.get("n_reevaluated", 0) / n


def esp_ncoef(L):
    """Polynomial coefficients of a primitive pair of total angular momentum L
    (N = 0..L, total degree max(0, 2N - L) .. N)."""
    return sum((t + 1) * (t + 2) // 2 for N in range(L + 1) for t in range(max(0, 2 * N - L), N + 1))


def esp_pair_flops(L):
    """Arithmetic of one (point, primitive pair) evaluation of the ESP kernel on
    its table branch (FMA = 2): geometry and T 9, Boys table index and offset
    4, Taylor sum 3 K = 24 (K = 8), for L > 0 e^{-dT} 18 + 2 and the downward
    recursion 3 L, monomials 3 L, polynomial 4 per coefficient, sum over N
    2 (L + 1).  The far branch (T >= 50 or 60) needs fewer."""
    boys = 4 + 24 + (20 + 3 * L if L > 0 else 0)
    return 9 + boys + 3 * L + 4 * esp_ncoef(L) + 2 * (L + 1)


def esp_rate(torch, capi, synthetic, reps=3):
    """ESP workload: 30 atoms, M = 312, 2e5 points on the 1.4 x vdW surface;
    points/s of the FP64 and mixed kernels (mixed: pair geometry, Boys function
    and pair sums in FP64, each pair's Hermite polynomial in FP32), the mixed
    kernel's largest error relative to |V_nuc| + |V_el| (the ESP crosses zero
    where they cancel) and the FP64 kernel's VALU roofline."""
    w = synthetic.esp30()
    pts = torch.from_numpy(np.ascontiguousarray(w.points)).cuda()
    n = pts.shape[0]
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    res, outs, kms = {}, {}, {}
    for prec in ("fp64", "mixed"):
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        h.esp(pts.data_ptr(), n, out.data_ptr(), prec)
        torch.cuda.synchronize()
        h.enable_profiling(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            h.esp(pts.data_ptr(), n, out.data_ptr(), prec)
        torch.cuda.synchronize()
        res[f"{prec}_points_per_s"] = n * reps / (time.perf_counter() - t0)
        st = h.kernel_stats()
        h.enable_profiling(False)
        kms[prec] = st["esp_ms"] / max(st["esp_calls"], 1)
        outs[prec] = out
    xyz = torch.from_numpy(w.basis.atom_xyz).cuda()
    Z = torch.from_numpy(np.asarray(w.charges, dtype=np.float64)).cuda()
    nuc = (Z[None, :] / torch.cdist(pts, xyz)).sum(dim=1)
    scale = nuc.abs() + (nuc - outs["fp64"]).abs()
    res["mixed_max_rel_err"] = float(((outs["mixed"] - outs["fp64"]).abs() / scale).max())
    res["mixed_over_fp64"] = res["mixed_points_per_s"] / res["fp64_points_per_s"]
    res["precision"] = {"fp64": "all FP64",
                        "mixed": "Boys function F_0..F_L, pair geometry and pair sums FP64; Hermite polynomial FP32"}
    pairs = h.esp_pairs_by_L()
    flops_pt = float(sum(int(pairs[L]) * esp_pair_flops(L) for L in range(9)))
    res["pairs_by_L"] = [int(c) for c in pairs]
    ach = flops_pt * n / (kms["fp64"] * 1e-3)
    res["roofline"] = {"bound": "valu", "achieved": ach / 1e12, "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s",
                       "frac": ach / PEAK["fp64"], "traffic": None, "kernel": "esp_kernel<double, 4, 512>",
                       "avg_launch_ms": kms["fp64"], "flops_per_point": flops_pt,
                       "mixed_avg_launch_ms": kms["mixed"]}
    res["workload"] = f"esp30: 30 atoms, M={w.basis.nbasis}, {n} points on the 1.4 x vdW surface"
    return res


def host_api_rate(w, pts_host, quantities, reps=2):
    """The drop-in API as the reference's users call it: Molecule.compute on
    numpy arrays, host points in and host outputs out (copies included), in
    both precisions."""
    import chemtoolscuda_amd as ct

    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    res = {}
    for prec in ("mixed", "fp64"):
        mol.compute(pts_host, quantities, precision=prec)
        t0 = time.perf_counter()
        for _ in range(reps):
            mol.compute(pts_host, quantities, precision=prec)
        res[f"{prec}_points_per_s"] = pts_host.shape[0] * reps / (time.perf_counter() - t0)
    res["api"] = "Molecule.compute(numpy (N,3), ['rho','grad','lap','tau_pd']) -> numpy, points and outputs in host memory"
    return res


def _err_scales(torch, ref):
    """Per-point error scales: rho, tau pointwise; grad against max(|grad|, sqrt(8 rho tau))
    (the Cauchy-Schwarz scale of d_k phi.P phi for a PSD P: grad vanishes at critical
    points); lap against |lap - 4 tau| + 4 tau, the size of its two summands; the Hessian
    H = A + T (A_kl = 2 d_kl phi.P phi, T_kl = 2 d_k phi.P d_l phi, |T_kl| <= tr T = 4 tau)
    against max|H_kl| + 8 tau >= max |A_kl| + |T_kl|; RDG against
    s (grad scale / |grad| + 4/3); Shannon against rho (|ln rho| + 1)."""
    rho = ref["rho"][:, 0].abs()
    out = {"rho": rho}
    if "tau_pd" in ref:
        tau = ref["tau_pd"][:, 0].abs()
        out["tau_pd"] = tau
        gn = torch.linalg.vector_norm(ref["grad"], dim=1)
        gs = torch.maximum(gn, torch.sqrt(8 * rho * tau))
        out["grad"] = gs
        lap_s = (ref["lap"][:, 0] - 4 * tau).abs() + 4 * tau
        out["lap"] = lap_s
        if "hess" in ref:
            out["hess"] = ref["hess"].abs().amax(dim=1) + 8 * tau
        if "rdg" in ref:
            out["rdg"] = ref["rdg"][:, 0].abs() * (gs / torch.clamp(gn, min=1e-300) + 4.0 / 3.0)
    if "shannon" in ref:
        out["shannon"] = rho * (torch.log(torch.clamp(rho, min=1e-300)).abs() + 1.0)
    return out


def _max_errors(torch, outs, ref, names):
    scales = _err_scales(torch, ref)
    mask = ref["rho"][:, 0] >= 1e-8
    res = {}
    for q in names:
        d = (outs[q] - ref[q]).abs().amax(dim=1) if q == "hess" else torch.linalg.vector_norm(outs[q] - ref[q], dim=1)
        err = d / torch.clamp(scales[q], min=1e-300)
        res[q] = float(err[mask].max())
    return res


def large_rate(torch, capi, synthetic, n_points=1_000_000):
    """300-atom workload (M = 5400 Cartesian, f on H, f+g on heavy atoms):
    Hessian, RDG and Shannon (with rho, grad, lap, tau_PD) in the mixed and
    FP64 paths, points/s of each and the mixed path's per-quantity errors."""
    w = synthetic.large300(n_points)
    order = synthetic.morton_order(w.points)
    pts = torch.from_numpy(np.ascontiguousarray(w.points[order])).cuda()
    n = pts.shape[0]
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    qs = ["rho", "grad", "hess", "lap", "tau_pd", "rdg", "shannon"]
    res, outs = {}, {}
    for prec in ("mixed", "fp64"):
        o = {q: torch.empty((n, capi.WIDTH.get(q, 1)), dtype=torch.float64, device="cuda") for q in qs}
        ptrs = {q: t.data_ptr() for q, t in o.items()}
        if prec == "mixed":  # (FP64: ~6 s per call at 10^6 points; its one call is the timed one)
            h.eval(pts.data_ptr(), n, ptrs, prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep = h.eval(pts.data_ptr(), n, ptrs, prec)
        torch.cuda.synchronize()
        res[f"{prec}_points_per_s"] = n / (time.perf_counter() - t0)
        if prec == "mixed":
            res["n_reevaluated"] = rep["n_reevaluated"]
        outs[prec] = o
    res["mixed_over_fp64"] = res["mixed_points_per_s"] / res["fp64_points_per_s"]
    res["max_rel_err"] = _max_errors(torch, outs["mixed"], outs["fp64"], ["rho", "hess", "rdg", "shannon"])
    res["workload"] = (f"large300: 300 atoms, M={w.basis.nbasis}, {n} points (per-atom Gamma(3,1) radii), "
                       "rho+grad+Hessian+lap+tau_PD+RDG+Shannon")
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(points shard across GPUs with no exchange in the forward pass; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    from chemtoolscuda_amd import capi, synthetic

    if args.copy_calibration:
        a = torch.empty(args.copy_calibration << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel()}))
        return
    if args.part:
        if args.part == "esp":
            out = esp_rate(torch, capi, synthetic)
        elif args.part == "fwd_bwd":
            out = {p: fwd_bwd(torch, capi, synthetic, p) for p in ("fp64", "mixed")}
        elif args.part == "autograd":
            out = {p: fwd_bwd_autograd(torch, synthetic, p) for p in ("fp64", "mixed")}
        elif args.part == "large300":
            out = large_rate(torch, capi, synthetic)
        else:
            w = synthetic.organic60()
            pts_host = np.ascontiguousarray(w.points[synthetic.morton_order(w.points)])
            out = host_api_rate(w, pts_host, ["rho", "grad", "lap", "tau_pd"])
        print(json.dumps({"part": args.part, **({"result": out} if not isinstance(out, dict) else out)}))
        return

    meta = json.load(open(os.path.join(ROOT, "BASELINE.json")))
    w = synthetic.WORKLOADS[args.workload]()
    order = synthetic.morton_order(w.points)
    pts_host = np.ascontiguousarray(w.points[order])
    n = pts_host.shape[0]
    M = w.basis.nbasis

    dev = torch.device("cuda:0")
    pts = torch.from_numpy(pts_host).to(dev)
    quantities = ["rho", "grad", "lap", "tau_pd"]

    def alloc():
        return {q: torch.empty((n, capi.WIDTH.get(q, 1)), dtype=torch.float64, device=dev) for q in quantities}

    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    if args.workspace_mb > 0:
        h.set_workspace_limit(args.workspace_mb << 20)
    if args.screening is not None:
        h.set_screening(args.screening)
    precision = args.precision
    if precision == "auto":
        precision = "mixed"

    route = {"auto": 0, "density": 1, "factor": 2}[args.route]

    def run(outs, prec):
        return h.eval(pts.data_ptr(), n, {q: t.data_ptr() for q, t in outs.items()}, prec, {"route": route})

    outs = alloc()
    for _ in range(args.warmup):
        run(outs, precision)
    torch.cuda.synchronize()
    h.enable_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep = None
    for _ in range(args.steps):
        rep = run(outs, precision)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stats = h.kernel_stats()
    h.enable_profiling(False)
    value = n * args.steps / dt

    ncol = 4
    flops_pt = 2.0 * M * M * ncol + 2.0 * M * 8
    contract_ms = stats["contract_ms"] / max(stats["contract_calls"], 1)
    pts_per_launch = n * args.steps / max(stats["contract_calls"], 1)
    achieved = flops_pt * pts_per_launch / (contract_ms * 1e-3)
    peak_key = "fp64" if precision == "fp64" else "bf16"
    rank = rep.get("factor_rank", -1) if rep else -1
    used_route = "factor" if rank >= 0 else "density"
    traffic_key = used_route + ("-unscreened" if used_route == "factor" and args.screening == 0 else "")
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK[peak_key] / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK[peak_key], "traffic": _traffic(w.name, precision, traffic_key, pts_per_launch),
                "kernel": "contract_" + ("f64" if precision == "fp64" else "mixed"),
                "avg_launch_ms": contract_ms, "points_per_launch": pts_per_launch,
                "flops_per_point": flops_pt, "route": used_route}
    if rank >= 0:
        # the factor route's own algorithmic count: psi = L^T [phi, d phi, lap phi]
        # (5 column types of rank r) and 8 dots of length r per point
        f_exec = 2.0 * M * rank * 5 + 2.0 * rank * 8
        a_exec = f_exec * pts_per_launch / (contract_ms * 1e-3)
        roofline["factor_rank"] = rank
        roofline["executed"] = {"flops_per_point": f_exec, "achieved": a_exec / 1e12, "frac": a_exec / PEAK[peak_key]}
        # spatial screening: fraction of (point tile, k-step) blocks contracted
        act = rep.get("active_fraction", 1.0) if rep else 1.0
        roofline["active_fraction"] = act
        if act < 1.0:
            f_scr = 2.0 * M * rank * 5 * act + 2.0 * rank * 8
            a_scr = f_scr * pts_per_launch / (contract_ms * 1e-3)
            roofline["executed_screened"] = {"flops_per_point": f_scr, "achieved": a_scr / 1e12,
                                             "frac": a_scr / PEAK[peak_key]}

    # what this device sustains for the same instruction mixes (clocks and
    # power as under load): the published peaks assume the top clock
    measured = capi.measure_peaks()
    m_key = "bf16_mfma_tflops" if precision != "fp64" else "f64_mfma_tflops"
    roofline["measured_peak"] = measured[m_key]
    roofline["frac_of_measured"] = roofline["achieved"] / measured[m_key]
    for sub in ("executed", "executed_screened"):
        if sub in roofline:
            roofline[sub]["frac_of_measured"] = roofline[sub]["achieved"] / measured[m_key]

    extra = {"measured_peaks": measured}
    if precision == "mixed" and not args.no_compare:
        # FP64 on the same route as the mixed run (AUTO: the screened factor
        # route when P has a low-rank factor): the like-for-like ratio
        ref = alloc()
        rep64 = run(ref, "fp64")  # warm-up (workspace growth)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rep64 = run(ref, "fp64")
        torch.cuda.synchronize()
        extra["fp64_points_per_s"] = n / (time.perf_counter() - t1)
        extra["fp64_route"] = "factor (screened)" if rep64.get("factor_rank", -1) >= 0 else "density"
        extra["mixed_over_fp64"] = value / extra["fp64_points_per_s"]
        # FP64 contracting with P itself (the reference's formulation): the
        # reference values for the error figures, and the ratio against them
        dens = alloc()
        h.eval(pts.data_ptr(), n, {q: t.data_ptr() for q, t in dens.items()}, "fp64", {"route": 1})
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        h.eval(pts.data_ptr(), n, {q: t.data_ptr() for q, t in dens.items()}, "fp64", {"route": 1})
        torch.cuda.synchronize()
        extra["fp64_density_route_points_per_s"] = n / (time.perf_counter() - t2)
        extra["mixed_over_fp64_density_route"] = value / extra["fp64_density_route_points_per_s"]
        for q, v in _max_errors(torch, outs, dens, quantities).items():
            extra[f"max_rel_err_{q}"] = v
        extra["max_rel_err_scales"] = ("over points with rho >= 1e-8, against the FP64 density route: rho and "
                                       "tau_pd relative to themselves; grad to max(|grad|, sqrt(8 rho tau)); "
                                       "lap to |lap - 4 tau| + 4 tau")
        extra["fp64_factor_vs_density_max_rel_err"] = max(_max_errors(torch, ref, dens, quantities).values())
        extra["n_reevaluated"] = rep.get("n_reevaluated") if rep else None
        del ref, dens

    if not args.no_extra:
        extra["host_arrays"] = host_api_rate(w, pts_host, quantities)
        fb, bw_ms = {}, {}
        for p in ("fp64", "mixed"):
            fb[p], bw_ms[p] = fwd_bwd(torch, capi, synthetic, p)
        extra["fwd_bwd_points_per_s"] = fb
        extra["fwd_bwd_backward_ms_per_call"] = bw_ms
        # the backward call (AO + Gram + Y = P phi per chunk): 4 M^2 flop per
        # point (dL/dP: 2 M^2; dL/dR: 2 M^2), 5e5 points per call
        m_d = synthetic.diff40(1).basis.nbasis
        f_bw = 4.0 * m_d * m_d * 500_000
        a_bw = f_bw / (bw_ms["mixed"] * 1e-3)
        extra["fwd_bwd_roofline"] = {"bound": "mfma", "achieved": a_bw / 1e12, "peak": PEAK["bf16"] / 1e12,
                                     "unit": "TFLOP/s", "frac": a_bw / PEAK["bf16"], "traffic": None,
                                     "kernel": "backward (gram_mixed + contract_adjoint_mixed + AO)",
                                     "avg_launch_ms": bw_ms["mixed"], "flops_per_point": 4.0 * m_d * m_d,
                                     "fp64": {"achieved": f_bw / (bw_ms["fp64"] * 1e-3) / 1e12,
                                              "peak": PEAK["fp64"] / 1e12,
                                              "frac": f_bw / (bw_ms["fp64"] * 1e-3) / PEAK["fp64"],
                                              "frac_of_measured": f_bw / (bw_ms["fp64"] * 1e-3) / 1e12 /
                                              measured["f64_mfma_tflops"]},
                                     "frac_of_measured": a_bw / 1e12 / measured["bf16_mfma_tflops"]}
        ag = {}
        for p in ("fp64", "mixed"):
            rate, fr = fwd_bwd_autograd(torch, synthetic, p)
            rate_mv, _ = fwd_bwd_autograd(torch, synthetic, p, move=True)
            ag[p] = {"points_per_s": rate, "points_per_s_moving_P_R": rate_mv, "reevaluated_fraction": fr}
        ag["mixed_over_fp64"] = ag["mixed"]["points_per_s"] / ag["fp64"]["points_per_s"]
        extra["fwd_bwd_autograd"] = ag
        extra["fwd_bwd_workload"] = ("diff40: 40 atoms, M=416, 500000 points, rho + dL/dP + dL/dR per step; "
                                     "autograd: L = sum w (rho - rho_t)^2, mixed forward held to 1e-6 of "
                                     "|rho - rho_t| (rho_target)")
        extra["esp"] = esp_rate(torch, capi, synthetic)
        extra["esp"]["roofline"]["frac_of_measured"] = (extra["esp"]["roofline"]["achieved"] /
                                                        measured["f64_valu_tflops"])
        extra["large300"] = large_rate(torch, capi, synthetic)

    kernel_ms = {k: stats[f"{k}_ms"] / args.steps for k in ("screen", "ao", "contract", "finalize", "reeval")}
    line = {
        "metric": meta["metric"],
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "fp32+bf16x3",
        "data": "synthetic (seeded generator, chemtoolscuda_amd/synthetic.py)",
        "config": {"workload": f"{w.name}: {w.basis.natom} atoms, M={M}, {n} grid points, rho+grad+lap+tau_PD",
                   "nbasis": M, "n_points": n, "precision": precision, "quantities": quantities},
        "roofline": roofline,
        "kernel_ms_per_step": kernel_ms,
        **extra,
    }
    if args.cpu_seconds > 0:
        line["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
