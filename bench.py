"""Benchmark of the plane-wave Kohn-Sham Hpsi and the real-time propagator on
one MI355X.  Prints ONE JSON line on stdout; diagnostics go to stderr.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--cpu-seconds 20]
                    [--workload hpsi_64|uneven_batch|adjoint_48|etrs_96|adjoint_96|gs_64|tddft_64]

hpsi_64 (default): band-Hpsi/s at the 64^3 grid x 256 bands FP64 workload
  (2x2x2 diamond supercell, a = 20.52 bohr, 64 atoms x 4 projectors, Gamma);
  a step is one Hpsi of all 256 bands with psi resident in HBM.  The same
  line carries the rest of the metric: the uneven batch against one launch
  per system (committed tuning, batched_speedup) and the adjoint/forward
  cost of adjoint_48 (adjoint_over_forward), each with its parity figures,
  and the oracle's rate at all host threads and at one (--hpsi-only skips
  the batch and adjoint parts).
uneven_batch: 24 seeded systems (grids 24..72, 8..256 bands, 2..128 atoms);
  a step is one Hpsi of every system, batched (one call) and per system.
adjoint_48: frozen-potential ETRS at 48^3 x 64 bands under a 32-term A(t),
  400 steps with checkpoints every 20; forward + adjoint cost.
etrs_96: the graph-captured ETRS step at 96^3 x 512 bands (216 atoms), and
  the adjoint of 100 steps with checkpoints every 10.
gs_64: one frozen-potential ground-state iteration (subspace diagonalisation,
  preconditioned steepest descent, orthogonalize) at 64^3 x 256 bands.
tddft_64: one interacting ETRS step (Hartree + PBE, self-consistent second
  half) at 64^3 x 256 bands.
autotune: the per-grid tiling table of the three-sweep Hpsi (inq_amd.tuning):
  every knob combination timed per grid length (--grids) and band class.
stream: the library's 16-byte streaming probe (read / write / copy) over
  1 GiB and over a band chunk of 64^3 x 256 rows: the achievable ceiling.
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


def vnl_from_coefficients(nlm):
    """Whether the three-sweep Hpsi expands V_nl psi from the projector
    coefficients in its last sweep (the library's default for nlm <= 4;
    INQAMD_VNL=entries keeps the per-entry V_nl psi round trip)."""
    return nlm <= 4 and os.environ.get("INQAMD_VNL", "") != "entries"


def kernel_design_bytes(N, nst, sum_k, nlm, chunks, coefficients=True, nproj=0):
    """Bytes each kernel of one Hpsi must move by design (its own compulsory
    reads and writes), per launch, counted once per launch whatever the
    caches do.  With the coefficient form: the projector gathers psi at the
    sphere points (once per 32-band tile its point, phase and nlm projector
    values) and writes P; each sweep reads and writes its band chunk; the last
    sweep also reads psi (V psi), V, the ownership word and sphere of each
    point and, once per 8-band tile, the entry tables (nlm values, phase).
    With per-entry V_nl psi the projector also writes 16 B per sphere point
    and band, and the last sweep reads them instead of the tables."""
    nb = nst / chunks
    if coefficients:
        return {
            "projector": 16.0 * sum_k * nst + (8.0 * nlm + 4.0 + 16.0) * sum_k * (nst / 32.0)
                         + 16.0 * nlm * nst * nproj,
            "z_fwd": 32.0 * N * nb,
            "xy_kin": 32.0 * N * nb,
            "z_bwd_epi": 48.0 * N * nb + 16.0 * N + (8.0 * 4 + 16.0 + 4.0) * sum_k * (nb / 8.0),
        }
    return {
        "projector": 32.0 * sum_k * nst + (8.0 * nlm + 4.0 + 12.0 + 32.0) * sum_k,
        "z_fwd": 32.0 * N * nb,
        "xy_kin": 32.0 * N * nb,
        "z_bwd_epi": 48.0 * N * nb + 12.0 * N + 16.0 * sum_k * nb,
    }


def cpu_baseline(s, psi, budget_s, threads=None):
    """The oracle's Hpsi on a bounded sample of bands of the same system, at
    `threads` OpenMP threads (None: all of them)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O

    osys = O.System(s.n, s.lattice, s.vloc, s.tables)
    default = O.get_threads()
    O.set_threads(threads if threads else 0)
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
    O.set_threads(default)
    return {
        "value": nb / dt, "unit": "band-Hpsi/s", "cores": threads, "kind": "port",
        "sample": f"{nb} of {psi.shape[-1]} bands of the same 64^3 system (all 64 atoms), "
                  f"C++ oracle -O3, OpenMP {threads} threads, one Hpsi call",
    }, ref


def common_fields(args, value, unit, ms_per_step, higher_is_better, config):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step,
            "higher_is_better": higher_is_better, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic", "config": config}


def timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def measure_uneven_batch(args, tune):
    """Config 2: 24 seeded systems; per-system launches against one batched
    call.  tune=False uses the committed tuning (inq_amd/tuning.json),
    tune=True re-runs the autotuner.  Returns a dict of the figures and the
    largest relative errors (batched vs per-system, one member vs oracle)."""
    import numpy as np
    import torch

    import inq_amd
    from inq_amd import batch as B

    t0 = time.perf_counter()
    lay = inq_amd.systems.uneven_batch_layout()
    dev = torch.device("cuda:0")
    sysl, hams, phis, outs = [], [], [], []
    for i, (L, nst, na) in enumerate(lay):
        s = inq_amd.systems.uneven_batch_system(i, L, na)
        psi = inq_amd.systems.random_orbitals(s, nst, seed=0, system_id=100 + i)
        sysl.append((s, psi))
        hams.append(inq_amd.ks_hamiltonian(s, max_bands=nst))
        phis.append(torch.from_numpy(psi).to(dev))
        outs.append(torch.empty_like(phis[-1]))
    log(f"uneven batch: {len(lay)} systems built in {time.perf_counter() - t0:.1f} s")
    bands = sum(nst for _, nst, _ in lay)

    def per_system():
        for h, p, o in zip(hams, phis, outs):
            h(p, out=o)

    ms_per = timed(per_system, args.steps, args.warmup)
    ref = [o.clone() for o in outs]
    bt = B.hamiltonian_batch(hams)
    if tune:
        t0 = time.perf_counter()
        best, table = B.autotune(bt, phis, outs, reps=3, chunk_mib=(128, 256, 512),
                                 ty_main=(8,), narrow=(True, False), group=(False, True),
                                 nstreams=(1, 2, 4))
        log(f"autotune {time.perf_counter() - t0:.1f} s: best {best}; " +
            ", ".join(f"{k}: {v:.3f} ms" for k, v in sorted(table.items(), key=lambda x: x[1])))
    else:
        bt.set_tuning(**B.committed_tuning("uneven_batch_24"))
        best = dict(bt.tuning)
    ms_bat = timed(lambda: bt(phis, outs), args.steps, args.warmup)
    torch.cuda.synchronize()
    batch_err = max(float((torch.linalg.vector_norm(o - r) / torch.linalg.vector_norm(r)).item())
                    for o, r in zip(outs, ref))
    info = bt.info()
    log(f"per-system {ms_per:.3f} ms, batched {ms_bat:.3f} ms -> {ms_per / ms_bat:.2f}x; {info}")
    total_bytes = 0.0
    for (s, _), (L, nst, _) in zip(sysl, lay):
        total_bytes += nst * algorithmic_bytes_per_band(s.npoints, nst,
                                                        int(np.sum(s.tables.get("sphere_size", 0))))
    achieved = total_bytes / (ms_bat / 1e3) / 1e9
    k = int(np.argmin([L ** 3 * nst for L, nst, _ in lay]))
    cpu = parity = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        s, psi = sysl[k]
        osys = O.System(s.n, s.lattice, s.vloc, s.tables)
        t0 = time.perf_counter()
        r = osys.hamiltonian(psi)
        dt = time.perf_counter() - t0
        parity = float(np.linalg.norm(outs[k].cpu().numpy() - r) / np.linalg.norm(r))
        cpu = {"value": lay[k][1] / dt, "unit": "band-Hpsi/s", "cores": O.get_threads(),
               "kind": "port",
               "sample": f"system {k} of the batch ({lay[k][0]}^3, {lay[k][1]} bands, "
                         f"{lay[k][2]} atoms), one Hpsi, C++ oracle OpenMP"}
        log(f"batch member {k} vs oracle: rel err {parity:.2e}")
    bt.close()
    for h in hams:
        h.close()
    return {"layout": [list(x) for x in lay], "bands": bands, "per_system_ms": ms_per,
            "batched_ms": ms_bat, "batched_speedup": ms_per / ms_bat, "tuning": best,
            "tuning_source": "autotune in this run" if tune else "inq_amd/tuning.json",
            "batch_info": info, "algorithmic_bytes_per_step": total_bytes,
            "achieved_GBs": achieved, "frac": achieved / HBM_PEAK_GBS,
            "batched_vs_per_system_max_rel_err": batch_err,
            "member_vs_oracle_rel_err": parity, "oracle_member": k, "cpu_baseline": cpu}


def run_uneven_batch(args):
    m = measure_uneven_batch(args, tune=args.autotune)
    cfg = {"workload": "uneven_batch: 24 seeded systems, grids 24..72 cubed, bands log-uniform "
                       "8..256, atoms 2..128 x 4 Gaussian projectors; a step = Hpsi of every "
                       "system in one batched call",
           "systems": m["layout"], "global_batch": m["bands"], "parallelism": "single GPU"}
    out = common_fields(args, m["bands"] / (m["batched_ms"] / 1e3), "band-Hpsi/s",
                        m["batched_ms"], True, cfg)
    out.update({"per_system_ms_per_step": m["per_system_ms"],
                "batched_speedup": m["batched_speedup"],
                "batched_vs_per_system_max_rel_err": m["batched_vs_per_system_max_rel_err"],
                "tuning": m["tuning"], "tuning_source": m["tuning_source"],
                "batch_info": m["batch_info"],
                "roofline": {"bound": "hbm", "achieved": m["achieved_GBs"], "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": m["frac"], "traffic": None,
                             "kernel": "batched Hpsi step (all launches of one call)",
                             "algorithmic_bytes_per_step": m["algorithmic_bytes_per_step"]},
                "cpu_baseline": m["cpu_baseline"],
                "max_rel_err_vs_oracle": m["member_vs_oracle_rel_err"]})
    return out


def _etrs_bytes(N, nst, sum_k):
    """B_step = nst (8 B_H + 256 N): eight Hpsi and two order-4 Taylor sums
    whose accumulations read and write the accumulator."""
    return nst * (8.0 * algorithmic_bytes_per_band(N, nst, sum_k) + 256.0 * N)


def _oracle_etrs_rate(s, psi, budget_s, dt, threads=None):
    """The oracle's frozen-potential ETRS step on a bounded sample of bands,
    at `threads` OpenMP threads (None: all of them)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    default = O.get_threads()
    O.set_threads(threads if threads else 0)
    osys = O.System(s.n, s.lattice, s.vloc, s.tables)
    one = np.ascontiguousarray(psi[..., :1])
    t0 = time.perf_counter()
    osys.etrs_step(one, dt, (0, 0, 0), (0, 0, 0))
    t1 = time.perf_counter() - t0
    nb = int(max(1, min(psi.shape[-1], budget_s / max(t1, 1e-3))))
    sample = np.ascontiguousarray(psi[..., :nb])
    t0 = time.perf_counter()
    osys.etrs_step(sample, dt, (0, 0, 0), (0, 0, 0))
    el = time.perf_counter() - t0
    cores = O.get_threads()
    O.set_threads(default)
    log(f"cpu oracle: {nb} bands, one ETRS step in {el:.2f} s with {cores} threads")
    return {"value": nb / el, "unit": "band-ETRS-steps/s", "cores": cores,
            "kind": "port", "sample": f"{nb} of {psi.shape[-1]} bands, one ETRS step "
                                      f"(two Taylor-4 half steps), C++ oracle OpenMP"}


def measure_adjoint(args, supercell, grid, nst, nsteps, ckpt, dt, fd_coefs=2, parity_bands=4,
                    parity_steps=5):
    """Forward propagation and its adjoint (loss J_x(T), 32 series
    coefficients): times, the adjoint/forward ratio, central differences of
    the GPU loss on the largest-gradient coefficients, and a short-horizon
    forward (parity_bands x parity_steps) against the oracle."""
    import numpy as np
    import torch

    import inq_amd
    from inq_amd import adjoint

    t0 = time.perf_counter()
    s = inq_amd.systems.diamond_system(supercell, grid=grid)
    psi = inq_amd.systems.random_orbitals(s, nst)
    coef = inq_amd.systems.rng(0, 0, 4).uniform(-0.05, 0.05, size=32)
    log(f"adjoint inputs built in {time.perf_counter() - t0:.1f} s: grid {s.n}, {nst} bands")
    dev = torch.device("cuda:0")
    d_psi = torch.from_numpy(psi).to(dev)
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    prop = adjoint.propagation_adjoint(ham, nst, nsteps, dt, 32, 0.1, (1.0, 0.0, 0.0),
                                       checkpoint_every=ckpt, band_slice=min(nst, 64))
    log(f"propagation memory {prop.memory_bytes / 2**30:.1f} GiB")
    torch.cuda.synchronize()
    w = time.perf_counter()
    loss, _ = prop.forward(coef, d_psi)
    t_fwd = time.perf_counter() - w
    w = time.perf_counter()
    gc, gV, _ = prop.backward(1.0)
    t_bwd = time.perf_counter() - w
    log(f"forward {t_fwd:.2f} s (loss {loss:.12e}), backward {t_bwd:.2f} s "
        f"-> ratio {t_bwd / t_fwd:.2f}")
    # central differences of the GPU forward on the largest-gradient coefficients
    h = 1e-5
    fd = {}
    for j in np.argsort(-np.abs(gc))[:fd_coefs]:
        e = np.zeros_like(coef)
        e[j] = h
        lp, _ = prop.forward(coef + e, d_psi)
        lm, _ = prop.forward(coef - e, d_psi)
        fd[int(j)] = ((lp - lm) / (2 * h), float(gc[j]))
    scale = float(np.abs(gc).max())
    fd_err = max(abs(a - b) for a, b in fd.values()) / scale
    log(f"FD check {fd} -> max rel err {fd_err:.2e}")
    memory = prop.memory_bytes / 2**30
    prop.close()
    # the forward arithmetic against the oracle on a short horizon
    parity = None
    if parity_bands > 0 and not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        few = np.ascontiguousarray(psi[..., :parity_bands])
        short = adjoint.propagation_adjoint(ham, parity_bands, parity_steps, dt, 32, 0.1,
                                            (1.0, 0.0, 0.0), checkpoint_every=ckpt)
        out = torch.empty_like(torch.from_numpy(few).to(dev))
        l_gpu, _ = short.forward(coef, torch.from_numpy(few).to(dev), out)
        short.close()
        osys = O.System(s.n, s.lattice, s.vloc, s.tables)
        t0 = time.perf_counter()
        l_ref, psi_ref = osys.td_current(few, coef, 0.1, (1.0, 0.0, 0.0), dt, parity_steps)
        log(f"oracle short forward ({parity_bands} bands x {parity_steps} steps) "
            f"{time.perf_counter() - t0:.1f} s")
        got = out.cpu().numpy()
        parity = {"bands": parity_bands, "steps": parity_steps,
                  "psi_rel_err": float(np.linalg.norm(got - psi_ref) / np.linalg.norm(psi_ref)),
                  "loss_rel_err": abs(l_gpu - l_ref) / max(abs(l_ref), 1e-300)}
        log(f"short forward vs oracle: {parity}")
    N = s.npoints
    sum_k = int(np.sum(s.tables["sphere_size"]))
    step_bytes = _etrs_bytes(N, nst, sum_k)
    fwd_ms_step = t_fwd * 1e3 / nsteps
    achieved = step_bytes / (fwd_ms_step / 1e3) / 1e9
    res = {"grid": list(s.n), "bands": nst, "atoms": s.natoms, "steps": nsteps,
           "checkpoint_every": ckpt, "dt": dt, "forward_s": t_fwd, "backward_s": t_bwd,
           "adjoint_over_forward": t_bwd / t_fwd, "loss": loss,
           "fd_check": {"h": h, "coefficients": fd, "max_rel_err": fd_err},
           "forward_vs_oracle": parity, "memory_GiB": memory,
           "forward_step_bytes": step_bytes, "forward_achieved_GBs": achieved,
           "forward_frac": achieved / HBM_PEAK_GBS, "sample_psi": psi, "system": s}
    ham.close()
    return res


def run_adjoint(args, supercell, grid, nst, nsteps, ckpt, dt, workload, fd_coefs=2):
    m = measure_adjoint(args, supercell, grid, nst, nsteps, ckpt, dt, fd_coefs=fd_coefs)
    cpu = None
    if not args.no_cpu_baseline:
        import inq_amd
        cpu = _oracle_etrs_rate(m["system"], inq_amd.systems.random_orbitals(m["system"], min(nst, 8)),
                                args.cpu_seconds, dt)
    cfg = {"workload": workload, "grid": m["grid"], "bands": nst, "atoms": m["atoms"],
           "steps": nsteps, "checkpoint_every": ckpt, "dt": dt, "taylor_order": 4,
           "coefficients": 32, "omega": 0.1, "parallelism": "single GPU"}
    out = common_fields(args, m["adjoint_over_forward"], "adjoint time / forward time",
                        m["forward_s"] * 1e3 / nsteps, False, cfg)
    out.update({"forward_s": m["forward_s"], "backward_s": m["backward_s"], "loss": m["loss"],
                "fd_check": m["fd_check"], "forward_vs_oracle": m["forward_vs_oracle"],
                "memory_GiB": m["memory_GiB"],
                "roofline": {"bound": "hbm", "achieved": m["forward_achieved_GBs"],
                             "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": m["forward_frac"],
                             "traffic": None, "kernel": "forward ETRS step (all launches)",
                             "algorithmic_bytes_per_step": m["forward_step_bytes"]},
                "cpu_baseline": cpu})
    return out


def run_etrs_96(args):
    import numpy as np
    import torch

    import inq_amd
    from inq_amd import real_time

    t0 = time.perf_counter()
    s = inq_amd.systems.diamond_system(3)
    nst = 512
    psi = inq_amd.systems.random_orbitals(s, nst)
    log(f"inputs built in {time.perf_counter() - t0:.1f} s: grid {s.n}, {s.natoms} atoms")
    dev = torch.device("cuda:0")
    d_psi = torch.from_numpy(psi).to(dev)
    # host bands kept for the parity check (first and last) and the CPU sample
    pbands = np.ascontiguousarray(psi[..., [0, nst - 1]])
    sample = np.ascontiguousarray(psi[..., :16])
    del psi
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    coef = inq_amd.systems.rng(0, 0, 4).uniform(-0.05, 0.05, size=32)
    pert = real_time.fourier_series_vector_potential(coef, 0.1)
    dt = 0.03
    d0 = d_psi.clone()
    g = real_time.etrs_graph(ham, d_psi, dt, 4, pert)
    ms = timed(g.launch, args.steps, args.warmup)
    g.close()
    N = s.npoints
    sum_k = int(np.sum(s.tables["sphere_size"]))
    step_bytes = _etrs_bytes(N, nst, sum_k)
    achieved = step_bytes / (ms / 1e3) / 1e9
    log(f"ETRS step {ms:.2f} ms ({achieved:.0f} GB/s algorithmic)")
    # one eager step under A(0), A(dt) from the same start, two bands against the oracle
    del d_psi
    real_time.etrs(ham, d0, 0.0, dt, pert)
    got = d0[..., [0, nst - 1]].cpu().numpy()
    del d0
    ham.close()
    parity = None
    cpu = cpu1 = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        lat = s.lattice
        t0 = time.perf_counter()
        ref = O.System(s.n, s.lattice, s.vloc, s.tables).etrs_step(
            pbands, dt, lat @ pert.uniform_vector_potential(0.0),
            lat @ pert.uniform_vector_potential(dt))
        parity = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        log(f"step vs oracle on bands 0 and {nst - 1}: rel err {parity:.2e} "
            f"({time.perf_counter() - t0:.1f} s)")
        cpu = _oracle_etrs_rate(s, sample, args.cpu_seconds, dt)
        cpu1 = _oracle_etrs_rate(s, sample, args.cpu_seconds / 2, dt, threads=1)
    cfg = {"workload": "etrs_96: 96^3 grid, 512 bands complex128, 3x3x3 diamond supercell "
                       "a=30.78 bohr, 216 atoms x 4 projectors, graph-captured ETRS step "
                       "(two Taylor-4 half steps) under a 32-term A(t), dt=0.03",
           "grid": list(s.n), "bands": nst, "atoms": s.natoms, "parallelism": "single GPU"}
    out = common_fields(args, 1e3 / ms, "ETRS steps/s", ms, True, cfg)
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                       "kernel": "graph-captured ETRS step (all launches)",
                       "algorithmic_bytes_per_step": step_bytes}
    out["cpu_baseline"] = cpu
    out["cpu_baseline_1thread"] = cpu1
    out["max_rel_err_vs_oracle"] = parity
    return out


def _gs_bytes(N, nst, sum_k):
    """Algorithmic bytes of one ground-state iteration (subspace
    diagonalisation, to_fourier, steepest descent, to_real) on nst bands:
    seven Hpsi at B_H (one in the subspace step, six G-space ones in steepest
    descent, each of those with two more transforms at 32N), the orbitals'
    two transforms, the line-search passes of five line minimisations (256N:
    16 B per array read or written; 224N in the last four, whose first pass
    comes with the previous step's update) and the two rotations (read +
    write, 32N each)."""
    b_h = algorithmic_bytes_per_band(N, nst, sum_k)
    return nst * (7.0 * b_h + 6 * 64.0 * N + 64.0 * N + (256.0 + 4 * 224.0) * N +
                  2 * 32.0 * N)


def run_gs_64(args):
    """One iteration of INQ's ground-state loop for a frozen potential
    (ground_state/calculator.hpp:150-180 without the density update) on the
    64^3 x 256 system: subspace diagonalisation, to_fourier, steepest descent
    (five preconditioned line minimisations + orthogonalize), to_real."""
    import numpy as np
    import torch

    import inq_amd
    from inq_amd import ground_state
    from inq_amd.operations import transform

    s = inq_amd.systems.diamond_system(2)
    nst = 256
    psi = inq_amd.systems.random_orbitals(s, nst)
    d_psi = torch.from_numpy(psi).to("cuda:0")
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    fphi = torch.empty_like(d_psi)

    def step():
        ground_state.subspace_diagonalization(ham, d_psi)
        transform.to_fourier(d_psi, ham, out=fphi)
        ground_state.steepest_descent(ham, fphi, True)
        transform.to_real(fphi, ham, out=d_psi)

    ms = timed(step, args.steps, args.warmup)
    eps, normres, _ = ham.energy(d_psi)
    log(f"GS iteration {ms:.2f} ms; eigenvalue sum {float(eps.sum()):.10f}, max normres "
        f"{float(normres.max()):.3e}")
    N = s.npoints
    sum_k = int(np.sum(s.tables["sphere_size"]))
    it_bytes = _gs_bytes(N, nst, sum_k)
    achieved = it_bytes / (ms / 1e3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        osys = O.System(s.n, s.lattice, s.vloc, s.tables)
        nb = 16
        sample = np.ascontiguousarray(psi[..., :nb])
        t0 = time.perf_counter()
        _, x = osys.subspace_diagonalization(sample)
        osys.steepest_descent(O.to_fourier(x), (0.0, 0.0, 0.0), True)
        dt = time.perf_counter() - t0
        cpu = {"value": nb / dt, "unit": "band-iterations/s", "cores": O.get_threads(),
               "kind": "port",
               "sample": f"{nb} of {nst} bands of the same 64^3 system, one iteration "
                         f"(subspace diagonalisation + steepest descent), C++ oracle with "
                         f"numpy eigh, OpenMP {O.get_threads()} threads"}
        log(f"cpu oracle: {nb} bands, one iteration in {dt:.2f} s")
    ham.close()
    cfg = {"workload": "gs_64: 64^3 grid, 256 bands complex128, 2x2x2 diamond supercell "
                       "a=20.52 bohr, 64 atoms x 4 projectors; one ground-state iteration "
                       "(subspace diagonalisation, steepest descent with the TPA "
                       "preconditioner, orthogonalize) for a frozen potential",
           "grid": list(s.n), "bands": nst, "atoms": s.natoms, "parallelism": "single GPU"}
    out = common_fields(args, nst * 1e3 / ms, "band-iterations/s", ms, True, cfg)
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                       "kernel": "one ground-state iteration (all launches)",
                       "algorithmic_bytes_per_step": it_bytes}
    out["cpu_baseline"] = cpu
    return out


def _tddft_bytes(N, nst, sum_k, nscf):
    """Algorithmic bytes of one interacting ETRS step on nst bands: the
    2-for-1 Taylor-4 step (four Hpsi; two accumulators read and written per
    order, except the first order's exp(phi), which is only written), then
    per SCF pass a Taylor-4 half step from the saved state (four Hpsi, one
    accumulator per order, the first only written) and a density pass (16N),
    plus the density of the full step (16N)."""
    b_h = algorithmic_bytes_per_band(N, nst, sum_k)
    two_for_one = 4 * b_h + 48.0 * N + 3 * 64.0 * N
    half = 4 * b_h + 16.0 * N + 3 * 32.0 * N
    per_band = two_for_one + 16.0 * N + nscf * (half + 16.0 * N)
    return nst * per_band


def run_tddft_64(args):
    """One interacting ETRS step (Hartree + PBE, INQ's default theory, or
    LDA with --xc lda; self-consistent second half) at 64^3 x 256 bands,
    from the KS potential of the initial density."""
    import numpy as np
    import torch

    import inq_amd
    from inq_amd import real_time, tddft

    s = inq_amd.systems.diamond_system(2)
    nst = 256
    psi = inq_amd.systems.random_orbitals(s, nst)
    d_psi = torch.from_numpy(psi).to("cuda:0")
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    vion = torch.from_numpy(s.vloc).to("cuda:0")
    coef = inq_amd.systems.rng(0, 0, 4).uniform(-0.05, 0.05, size=32)
    pert = real_time.fourier_series_vector_potential(coef, 0.1)
    dt = 0.015
    tddft.ks_potential(ham, vion, tddft.density(ham, d_psi, 2.0), xc=args.xc)
    state = {"t": 0.0, "nscf": []}

    def step():
        state["nscf"].append(tddft.etrs(ham, d_psi, vion, state["t"], dt, 2.0, pert, xc=args.xc))
        state["t"] += dt

    ms = timed(step, args.steps, args.warmup)
    nscf = float(np.mean(state["nscf"][args.warmup:])) if len(state["nscf"]) > args.warmup else 1.0
    log(f"interacting ETRS step {ms:.2f} ms, SCF passes per step {nscf:.2f}")
    N = s.npoints
    sum_k = int(np.sum(s.tables["sphere_size"]))
    step_bytes = _tddft_bytes(N, nst, sum_k, nscf)
    achieved = step_bytes / (ms / 1e3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        nb = 8
        sample = np.ascontiguousarray(psi[..., :nb])
        osys = O.System(s.n, s.lattice, s.vloc, s.tables)
        v0, _ = osys.ks_potential(s.vloc, O.density(sample, 2.0), xc=args.xc)
        t0 = time.perf_counter()
        _, _, n_ref = O.System(s.n, s.lattice, v0, s.tables).etrs_scf(sample, s.vloc, 2.0, dt,
                                                                     xc=args.xc)
        dtc = time.perf_counter() - t0
        cpu = {"value": nb / dtc, "unit": "band-steps/s", "cores": O.get_threads(), "kind": "port",
               "sample": f"{nb} of {nst} bands of the same 64^3 system (their own density), one "
                         f"interacting ETRS step ({n_ref} SCF passes), C++ oracle, OpenMP "
                         f"{O.get_threads()} threads"}
        log(f"cpu oracle: {nb} bands, one step in {dtc:.2f} s")
    ham.close()
    cfg = {"workload": "tddft_64: 64^3 grid, 256 bands complex128 (occupation 2), 2x2x2 diamond "
                       "supercell a=20.52 bohr, 64 atoms x 4 projectors; interacting ETRS step "
                       f"(Taylor-4, Hartree + {args.xc.upper()}, self-consistent second half) "
                       "under a 32-term A(t), dt=0.015",
           "grid": list(s.n), "bands": nst, "atoms": s.natoms, "parallelism": "single GPU",
           "scf_passes_per_step": nscf}
    out = common_fields(args, nst * 1e3 / ms, "band-steps/s", ms, True, cfg)
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                       "kernel": "one interacting ETRS step (all launches)",
                       "algorithmic_bytes_per_step": step_bytes}
    out["cpu_baseline"] = cpu
    return out


def run_stream(args):
    """The streaming ceiling: the library's own 16-byte vector read / write /
    copy kernel (inqamd_stream_probe) over 1 GiB contiguous, and over a
    48-band chunk of 64^3 x 256 rows (ld = 256), as a contiguous window and as
    interleaved 8-band groups, at 1, 2 and 4 workgroups per CU."""
    import ctypes
    import torch
    from inq_amd import _lib

    L = _lib.lib()
    dev = torch.device("cuda:0")
    N, ld, nb, nch = 64 ** 3, 256, 48, 6
    total = N * ld                      # complex128 elements (1 GiB)
    a = torch.ones(total, dtype=torch.complex128, device=dev)
    b = torch.zeros_like(a)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ms = ctypes.c_float(0.0)
    cases = [("contiguous", total, 1, 1, 0, 1)]
    cases += [("rows ld=256, 48-band window", N, ld, nb, 0, 1),
              ("rows ld=256, 48 bands in 8-band groups", N, ld, nb, 1, nch)]
    rows_out = []
    best = {}
    for name, rows, rld, rnb, grouped, nchunks in cases:
        for kind, kname in ((0, "read"), (1, "write"), (2, "copy")):
            for bpc in (1, 2, 4, 8):
                times = []
                for _ in range(args.warmup + args.steps):
                    _lib.check(L.inqamd_stream_probe(kind, ctypes.c_void_p(a.data_ptr()),
                                                     ctypes.c_void_p(b.data_ptr()), rows, rld,
                                                     rnb, grouped, nchunks, bpc, stream,
                                                     ctypes.byref(ms)))
                    times.append(ms.value)
                t = sorted(times[args.warmup:])[len(times[args.warmup:]) // 2]
                nbytes = rows * rnb * 16.0 * (2 if kind == 2 else 1)
                gbs = nbytes / (t / 1e3) / 1e9
                rows_out.append({"pattern": name, "kind": kname, "wg_per_cu": bpc,
                                 "bytes": nbytes, "median_ms": t, "GBs": gbs})
                key = (name, kname)
                best[key] = max(best.get(key, 0.0), gbs)
                log(f"{name:40s} {kname:5s} {bpc} WG/CU: {t:.3f} ms {gbs:7.0f} GB/s")
    cfg = {"workload": "stream: 16-byte vector read/write/copy, 1 GiB contiguous and a 48-band "
                       "chunk of 64^3 x 256 complex128 rows", "parallelism": "single GPU"}
    top = best[("contiguous", "copy")]
    out = common_fields(args, top, "GB/s", None, True, cfg)
    out["best"] = {f"{k[0]} {k[1]}": v for k, v in best.items()}
    out["cases"] = rows_out
    out["roofline"] = {"bound": "hbm", "achieved": top, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": top / HBM_PEAK_GBS, "traffic": None,
                       "kernel": "probe_kernel<2> (copy), 1 GiB contiguous"}
    out["cpu_baseline"] = None
    return out


def run_autotune(args):
    """The per-grid tiling table (inq_amd.tuning): every knob combination of
    the three-sweep Hpsi timed for each grid length of the workloads and each
    band class; prints the table (the "hpsi" section of tuning.json)."""
    from inq_amd import tuning
    lengths = [int(x) for x in args.grids.split(",")]
    t0 = time.perf_counter()
    table = tuning.grid_table(lengths, reps=args.steps, log=log)
    log(f"autotune of {len(lengths)} grid lengths in {time.perf_counter() - t0:.1f} s")
    cfg = {"workload": "autotune: three-sweep Hpsi tiling per grid length and band class",
           "grids": lengths, "parallelism": "single GPU"}
    out = common_fields(args, len(lengths), "grid lengths tuned", None, True, cfg)
    out["table"] = table
    out["roofline"] = None
    out["cpu_baseline"] = None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="hpsi_64",
                    choices=["hpsi_64", "uneven_batch", "adjoint_48", "etrs_96", "adjoint_96",
                             "gs_64", "tddft_64", "stream", "autotune"])
    ap.add_argument("--grids", default="24,30,32,36,40,48,60,64,72,96",
                    help="autotune: the grid lengths to tune")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--hpsi-only", action="store_true",
                    help="hpsi_64: skip the batched and adjoint parts of the metric")
    ap.add_argument("--autotune", action="store_true",
                    help="uneven_batch: re-run the autotuner instead of the committed tuning")
    ap.add_argument("--xc", choices=["pbe", "lda"], default="pbe",
                    help="exchange-correlation of the tddft_64 workload (PBE is INQ's default)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also run 3 plain device copies of this size (calibrates rocprof "
                         "FETCH_SIZE/WRITE_SIZE against a kernel of known traffic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the bands shard with no communication inside "
              f"Hpsi); --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    if args.workload == "uneven_batch":
        print(json.dumps(run_uneven_batch(args)), flush=True)
        return
    if args.workload == "stream":
        print(json.dumps(run_stream(args)), flush=True)
        return
    if args.workload == "autotune":
        print(json.dumps(run_autotune(args)), flush=True)
        return
    if args.workload == "adjoint_48":
        out = run_adjoint(args, 1, 48, 64, 400, 20, 0.015,
                          "adjoint_48: 48^3 grid (a0 cell, 8 atoms x 4 projectors), 64 bands, "
                          "velocity gauge A_x(t) = sum_j c_j sin(j 0.1 t), 32 seeded c_j, "
                          "Taylor-4 ETRS, 400 steps of dt=0.015, checkpoints every 20; "
                          "loss J_x(T)")
        print(json.dumps(out), flush=True)
        return
    if args.workload == "adjoint_96":
        out = run_adjoint(args, 3, None, 512, 100, 10, 0.03,
                          "adjoint_96: 96^3 grid, 512 bands, 216 atoms x 4 projectors, "
                          "Taylor-4 ETRS, 100 steps of dt=0.03, checkpoints every 10; "
                          "loss J_x(T)", fd_coefs=1)
        print(json.dumps(out), flush=True)
        return
    if args.workload == "tddft_64":
        print(json.dumps(run_tddft_64(args)))
        return
    if args.workload == "gs_64":
        print(json.dumps(run_gs_64(args)))
        return
    if args.workload == "etrs_96":
        print(json.dumps(run_etrs_96(args)), flush=True)
        return

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
    nlm = int(s.tables["max_nlm"])
    coefficients = vnl_from_coefficients(nlm)
    design = kernel_design_bytes(N, nst, sum_k, nlm, chunks, coefficients, s.natoms)
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
                "vnl": "expanded from the projector coefficients in z_bwd_epi" if coefficients
                       else "per-entry V_nl psi written by the projector, read by z_bwd_epi",
                "avg_launch_ms": kernels[dom]["avg_launch_ms"],
                "design_bytes_per_launch": kernels[dom]["design_bytes_per_launch"],
                "kernels": kernels,
                "step": {"algorithmic_bytes_per_band": b_h, "achieved": step_achieved,
                         "frac": step_achieved / HBM_PEAK_GBS}}

    result_host = d_out.cpu().numpy()
    ham.close()
    del d_psi, d_out
    cpu = cpu1 = None
    parity = None
    if not args.no_cpu_baseline:
        cpu, ref = cpu_baseline(s, psi, args.cpu_seconds)
        nb = ref.shape[-1]
        got = result_host[..., :nb]
        parity = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        log(f"parity vs oracle on {nb} bands: rel err {parity:.3e}")
        cpu1, ref1 = cpu_baseline(s, psi, args.cpu_seconds / 2, threads=1)
        p1 = float(np.linalg.norm(result_host[..., :ref1.shape[-1]] - ref1) / np.linalg.norm(ref1))
        cpu1["max_rel_err_vs_gpu"] = p1
    extras = {}
    if not args.hpsi_only:
        # the rest of the metric: batched vs per-system (config 2) with the
        # committed tuning, and the adjoint/forward cost (config 3)
        torch.cuda.empty_cache()
        try:
            extras["batched"] = measure_uneven_batch(args, tune=False)
        except Exception as e:  # the headline figure above stays valid
            extras["batched"] = {"error": repr(e)}
            log(f"uneven batch failed: {e!r}")
        torch.cuda.empty_cache()
        try:
            m = measure_adjoint(args, 1, 48, 64, 400, 20, 0.015)
            m.pop("sample_psi")
            m.pop("system")
            extras["adjoint"] = m
        except Exception as e:
            extras["adjoint"] = {"error": repr(e)}
            log(f"adjoint failed: {e!r}")

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
        "cpu_baseline_1thread": cpu1,
        "max_rel_err_vs_oracle": parity,
    }
    if extras:
        b = extras.get("batched", {})
        a = extras.get("adjoint", {})
        out["batched_speedup"] = b.get("batched_speedup")
        out["adjoint_over_forward"] = a.get("adjoint_over_forward")
        out["batched"] = b
        out["adjoint"] = a
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
