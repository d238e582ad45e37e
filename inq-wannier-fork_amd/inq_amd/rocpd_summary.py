"""Summarise rocprofv3 result databases (rocpd SQLite) as plain text.

    python -m inq_amd.rocpd_summary trace.db [--fetch fetch.db --write write.db]
        [--calib-bytes BYTES] [--per N] [--label TEXT]

--fetch/--write are databases of `rocprofv3 --pmc FETCH_SIZE` and
`--pmc WRITE_SIZE` runs (values in KiB per dispatch).  The counters are
scaled by what they read for the largest __amd_rocclr_copyBuffer dispatch,
a plain copy of --calib-bytes bytes (read and written once).  --per divides
totals by the number of Hpsi calls in the run.  --serial-calls K adds a
table over the dispatches of the run's last K Hpsi calls only (a bench run
ends with per-kernel timed calls that launch one kernel at a time on one
stream, so these durations are the kernels alone, comparable with the
bench's own event timings).
"""
import argparse
import collections
import sqlite3


def kernel_stats(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, total_calls, total_duration, average, percentage "
                       "from top_kernels").fetchall()
    con.close()
    return rows


def last_calls_stats(db, per, k):
    """Per kernel: dispatches and mean duration (us) over its dispatches in
    the last k of `per` Hpsi calls (each kernel launches n/per times per call)."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, start, duration from kernels order by start").fetchall()
    con.close()
    by = collections.defaultdict(list)
    for name, start, dur in rows:
        by[name].append(dur)
    out = []
    for name, durs in by.items():
        if len(durs) % per:
            continue
        take = len(durs) // per * k
        sel = durs[-take:]
        out.append((name, len(sel), sum(sel) / len(sel) / 1e3))
    return sorted(out, key=lambda x: -x[1] * x[2])


def pmc(db, counter):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, grid_size, value, duration from counters_collection "
                       "where counter_name = ?", (counter,)).fetchall()
    con.close()
    agg = collections.defaultdict(list)
    for name, grid, val, dur in rows:
        agg[(name, grid)].append((val, dur))
    return agg


def short(name, width=72):
    return name if len(name) <= width else name[: width - 3] + "..."


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--calib-bytes", type=float, default=float(1 << 30))
    ap.add_argument("--per", type=int, default=1, help="Hpsi calls in the traced run")
    ap.add_argument("--per-pmc", type=int, default=0,
                    help="Hpsi calls in each counter run (default: --per)")
    ap.add_argument("--serial-calls", type=int, default=0,
                    help="also summarise the dispatches of the last K Hpsi calls of the trace")
    ap.add_argument("--json", help="also write the per-dispatch bytes of each kernel here")
    ap.add_argument("--label", default="")
    args = ap.parse_args()

    print(f"# rocprofv3 kernel summary {args.label}".rstrip())
    print(f"{'kernel':74s} {'calls':>6s} {'avg_us':>9s} {'total_us':>10s} {'%':>6s}")
    for name, calls, total, avg, pct in kernel_stats(args.trace):
        # top_kernels reports microseconds
        print(f"{short(name):74s} {calls:6d} {avg:9.2f} {total:10.1f} {pct:6.2f}")

    if args.serial_calls:
        print()
        print(f"# the last {args.serial_calls} of {args.per} Hpsi calls (kernels one at a time)")
        print(f"{'kernel':74s} {'calls':>6s} {'avg_us':>9s}")
        for name, n, avg in last_calls_stats(args.trace, args.per, args.serial_calls):
            print(f"{short(name):74s} {n:6d} {avg:9.2f}")

    if args.fetch and args.write:
        f = pmc(args.fetch, "FETCH_SIZE")
        w = pmc(args.write, "WRITE_SIZE")

        def calib(agg):
            cands = [(g, v) for (n, g), v in agg.items() if "copyBuffer" in n]
            g, v = max(cands, key=lambda x: x[0])
            return args.calib_bytes / (sum(x[0] for x in v) / len(v) * 1024.0)

        cf, cw = calib(f), calib(w)
        print()
        print(f"# HBM-side bytes per dispatch (FETCH_SIZE x {cf:.3f}, WRITE_SIZE x {cw:.3f}; "
              f"factors from a {args.calib_bytes / 2**20:.0f} MiB copyBuffer)")
        print(f"{'kernel':74s} {'grid':>10s} {'read_MB':>9s} {'write_MB':>9s} {'avg_us':>8s} {'GB/s':>8s}")
        tot_r = tot_w = 0.0
        per_kernel = collections.defaultdict(lambda: [0.0, 0])
        for key in sorted(f, key=lambda k: -sum(x[1] for x in f[k])):
            name, grid = key
            if "copyBuffer" in name or "vectorized_elementwise" in name or key not in w:
                continue
            rd = sum(x[0] for x in f[key]) / len(f[key]) * 1024.0 * cf
            wr = sum(x[0] for x in w[key]) / len(w[key]) * 1024.0 * cw
            dur = sum(x[1] for x in f[key]) / len(f[key])
            n = len(f[key])
            tot_r += rd * n
            tot_w += wr * n
            base = name.split("(")[0].split("<")[0].split("::")[-1].replace("_kernel", "")
            per_kernel[base][0] += (rd + wr) * n
            per_kernel[base][1] += n
            print(f"{short(name):74s} {grid:10d} {rd / 1e6:9.1f} {wr / 1e6:9.1f} {dur / 1e3:8.1f} "
                  f"{(rd + wr) / max(dur, 1):8.0f}")
        per = args.per_pmc or args.per
        print(f"# total per Hpsi: read {tot_r / per / 1e9:.3f} GB, write "
              f"{tot_w / per / 1e9:.3f} GB, traffic {(tot_r + tot_w) / per / 1e9:.3f} GB")
        if args.json:
            import json
            out = {k: v[0] / max(v[1], 1) for k, v in per_kernel.items()}
            with open(args.json, "w") as fh:
                json.dump({"bytes_per_dispatch": out,
                           "bytes_per_hpsi": (tot_r + tot_w) / per}, fh, indent=1)


if __name__ == "__main__":
    main()
