"""Benchmark: band-Hpsi/s of the plane-wave Kohn-Sham Hamiltonian at the
64^3 grid x 256 bands FP64 workload (2x2x2 diamond supercell, a = 20.52 bohr,
64 atoms x 4 projectors, Gamma point), on one MI355X.

A "step" is one Hpsi application to all 256 bands with psi resident in HBM.
Prints ONE JSON line on stdout; diagnostics go to stderr.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--cpu-seconds 20]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "inq-wannier-fork_amd"))

METRIC = ("band-Hψ/s at 64³×256 FP64, %HBM roofline; batched vs per-system; "
          "adjoint/fwd cost")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def algorithmic_bytes_per_band(N, nst, sum_k):
    """B_H = 80 N + 8 N/nst + 48 sum_K: two 3-D FFTs at one read + one write of
    16 B/point each, the V psi re-read of psi, V amortised over the bands, and
    16 B gather + 32 B scatter read-modify-write per sphere point."""
    return 80.0 * N + 8.0 * N / nst + 48.0 * sum_k


def kernel_design_bytes(N, nst, sum_k, nlm, chunks):
    """Bytes each kernel of one Hpsi must move by design (its own compulsory
    reads and writes), per launch: the projector's gather of psi and write of
    V_nl psi per sphere entry (plus its tables), each sweep's read and write of
    the band chunk, and the last sweep's extra reads of psi (V psi), V, the
    ownership map and the V_nl psi entries."""
    nb = nst / chunks
    return {
        "projector": 32.0 * sum_k * nst + (8.0 * nlm + 4.0 + 12.0 + 32.0) * sum_k,
        "z_fwd": 32.0 * N * nb,
        "xy_kin": 32.0 * N * nb,
        "z_bwd_epi": 48.0 * N * nb + 12.0 * N + 16.0 * sum_k * nb,
    }


def cpu_baseline(s, psi, budget_s):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O

    osys = O.System(s.n, s.lattice, s.vloc, s.tables)
    threads = O.get_threads()
    one = np.ascontiguousarray(psi[..., :1])
    t0 = time.perf_counter()
    osys.hamiltonian(one)
    t1 = time.perf_counter() - t0
    nb = int(max(1, min(psi.shape[-1], budget_s / max(t1, 1e-3))))
    sample = np.ascontiguousarray(psi[..., :nb])
    t0 = time.perf_counter()
    ref = osys.hamiltonian(sample)
    dt = time.perf_counter() - t0
    log(f"cpu oracle: {nb} bands in {dt:.2f} s with {threads} threads")
    return {
        "value": nb / dt, "unit": "band-Hpsi/s", "cores": threads, "kind": "port",
        "sample": f"{nb} of {psi.shape[-1]} bands of the same 64^3 system (all 64 atoms), "
                  f"C++ oracle -O3, OpenMP {threads} threads, one Hpsi call",
    }, ref


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also run 3 plain device copies of this size (calibrates rocprof "
                         "FETCH_SIZE/WRITE_SIZE against a kernel of known traffic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the bands shard with no communication inside "
              f"Hpsi); --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import inq_amd

    t0 = time.perf_counter()
    s = inq_amd.systems.diamond_system(2)
    nst = 256
    psi = inq_amd.systems.random_orbitals(s, nst)
    log(f"inputs built in {time.perf_counter() - t0:.1f} s: grid {s.n}, {s.natoms} atoms, "
        f"sphere sizes {int(s.tables['sphere_size'].min())}..{int(s.tables['sphere_size'].max())}")
    dev = torch.device("cuda:0")
    d_psi = torch.from_numpy(psi).to(dev)
    d_out = torch.empty_like(d_psi)
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    if args.copy_bytes > 0:
        src = torch.empty(args.copy_bytes // 8, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        ham(d_psi, out=d_out)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        ham(d_psi, out=d_out)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = ev0.elapsed_time(ev1)
    ms_per_step = ms / args.steps
    value = nst * args.steps / (ms / 1e3)
    log(f"{args.steps} steps: {ms:.2f} ms (wall {wall * 1e3:.2f} ms) -> {ms_per_step:.3f} ms/step")

    N = s.npoints
    sum_k = int(np.sum(s.tables["sphere_size"]))
    b_h = algorithmic_bytes_per_band(N, nst, sum_k)
    step_achieved = b_h * nst / (ms_per_step / 1e3) / 1e9
    # per-kernel timing: the same Hpsi with HIP events around every launch on
    # the handle's stream, averaged over the timed repetitions
    reps = max(1, min(args.steps, 10))
    tot = {}
    for _ in range(reps):
        for k, (ms_k, n_k) in ham.profile(d_psi, d_out).items():
            t, n = tot.get(k, (0.0, 0))
            tot[k] = (t + ms_k, n + n_k)
    chunks = tot["z_fwd"][1] // reps
    design = kernel_design_bytes(N, nst, sum_k, int(s.tables["max_nlm"]), chunks)
    kernels = {}
    for k, (t, n) in tot.items():
        avg_ms = t / max(n, 1)
        per_launch = design[k] if k != "projector" else design[k] / 2
        kernels[k] = {"ms_per_hpsi": t / reps, "launches_per_hpsi": n // reps,
                      "avg_launch_ms": avg_ms, "design_bytes_per_launch": per_launch,
                      "GBs": per_launch / (avg_ms / 1e3) / 1e9}
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_hpsi"])
    log("per-kernel: " + ", ".join(f"{k} {v['ms_per_hpsi']:.3f} ms ({v['GBs']:.0f} GB/s)"
                                   for k, v in kernels.items()))
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "hpsi_64_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_dispatch", {}).get(dom)
    achieved = kernels[dom]["GBs"]
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "avg_launch_ms": kernels[dom]["avg_launch_ms"],
                "design_bytes_per_launch": kernels[dom]["design_bytes_per_launch"],
                "kernels": kernels,
                "step": {"algorithmic_bytes_per_band": b_h, "achieved": step_achieved,
                         "frac": step_achieved / HBM_PEAK_GBS}}

    result_host = d_out.cpu().numpy()
    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        cpu, ref = cpu_baseline(s, psi, args.cpu_seconds)
        nb = ref.shape[-1]
        got = result_host[..., :nb]
        parity = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        log(f"parity vs oracle on {nb} bands: rel err {parity:.3e}")

    out = {
        "metric": METRIC, "value": value, "unit": "band-Hpsi/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "hpsi_64: 64^3 grid, 256 bands complex128, 2x2x2 diamond "
                               "supercell a=20.52 bohr, 64 atoms x 4 Gaussian projectors, "
                               "Gamma point, one ks_hamiltonian apply per step",
                   "global_batch": nst, "grid": list(s.n), "atoms": s.natoms,
                   "sum_sphere_points": sum_k, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "max_rel_err_vs_oracle": parity,
    }
    print(json.dumps(out), flush=True)
    ham.close()


if __name__ == "__main__":
    main()
