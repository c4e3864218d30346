"""Summarise rocprofv3 result databases (rocpd SQLite) of the MD kernels.

Reads one kernel-trace database (--stats) and optionally two counter runs
(FETCH_SIZE, WRITE_SIZE) of the same command plus the same two counters for a
device-to-device copy of known size; the copy fixes the scale of each counter
(bytes moved / counter reading), which turns the MD kernel's readings into HBM
bytes per launch.  Prints a short markdown report, including the line

    traffic_bytes_per_launch: <int>

that bench.py reads for its roofline record.

    python -m moleculardynamics_amd.rocprof_summary --trace T.db \
        --fetch F.db --write W.db --copy-fetch CF.db --copy-write CW.db \
        --copy-bytes 268435456 --kernel "k_step_lds<false, false, false>" \
        --units 1048576 --bytes-per-unit 312 --command "..."

--halted-below FRAC sets apart the launches of the named kernel that did
next to no work (the single box's step kernels that found their list gate set
and returned): in a trace, launches shorter than FRAC x the kernel's median
duration; in a counter run, dispatches where any counter reads below FRAC x
its median over the kernel's dispatches (every pass holds a counter that
scales with the work).  They are counted on their own and left out of the
averages.

With --period-kernel K (and --period-range A:B), also reports the step
timeline: consecutive launches of K delimit the steps, and for steps A..B-1
the wall time per step is set against the kernels' busy time (kernels on one
stream do not overlap), giving the time between kernels -- what fusing the
step's kernels into one launch could at most save -- and each kernel's share.

With --counters, tabulates every counter of one or more --pmc runs for the
named kernels instead: per-instance rows summed per dispatch, then averaged
over the dispatches, plus the ratios the counters give directly (the trace,
if given, supplies the kernel duration for the VALU issue fraction).

    python -m moleculardynamics_amd.rocprof_summary --counters P1 P2 \
        --kernel k_nlist_lds --kernel "k_step_lds<false, false, false>" \
        [--trace T.db]
"""
from __future__ import annotations

import argparse
import glob
import sqlite3
from collections import defaultdict
from pathlib import Path


def _db(path: str) -> sqlite3.Connection:
    p = Path(path)
    if p.is_dir():
        hits = sorted(glob.glob(str(p / "**" / "*.db"), recursive=True))
        if not hits:
            raise FileNotFoundError(f"no .db under {p}")
        p = Path(hits[0])
    return sqlite3.connect(str(p))


def short(name: str) -> str:
    n = name.replace("void ", "").replace("mdx::", "").replace("(anonymous namespace)::", "")
    depth, out = 0, []
    for ch in n:   # drop the argument list, keep template arguments
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out)


def timeline_report(con: sqlite3.Connection, marker: str, lo: int, hi: int) -> None:
    ev = sorted((start, end, short(name)) for name, start, end in
                con.execute("select name, start, end from kernels"))
    marks = [s for s, _, n in ev if n == marker]
    if len(marks) < hi + 1:
        raise SystemExit(f"only {len(marks)} launches of {marker!r}; need {hi + 1}")
    t0, t1 = marks[lo], marks[hi]
    inside = [(s, e, n) for s, e, n in ev if t0 <= s < t1]
    wall = (t1 - t0) / (hi - lo)
    busy = sum(e - s for s, e, _ in inside) / (hi - lo)
    per = defaultdict(float)
    calls = defaultdict(int)
    for s, e, n in inside:
        per[n] += (e - s) / (hi - lo)
        calls[n] += 1
    print(f"\n## Step timeline (steps {lo}..{hi - 1}, delimited by `{marker}`)\n")
    print(f"- wall time per step: {wall / 1e3:.1f} us")
    print(f"- kernels busy per step: {busy / 1e3:.1f} us; between kernels: {(wall - busy) / 1e3:.1f} us "
          f"({100 * (wall - busy) / wall:.1f}% of the step)")
    print(f"- kernels per step: {sum(calls.values()) / (hi - lo):.1f}\n")
    print("| kernel | per step us | launches per step |")
    print("|---|---:|---:|")
    for n, t in sorted(per.items(), key=lambda kv: -kv[1]):
        print(f"| `{n}` | {t / 1e3:.1f} | {calls[n] / (hi - lo):.2f} |")


def kernel_stats(con: sqlite3.Connection):
    rows = defaultdict(list)
    meta = {}
    for name, dur, gx, wx, lds, vgpr, agpr, sgpr in con.execute(
            "select name, duration, grid_x, workgroup_x, lds_size, vgpr_count, accum_vgpr_count, sgpr_count "
            "from kernels"):
        rows[short(name)].append(dur)
        meta[short(name)] = (gx, wx, lds, vgpr, agpr, sgpr)
    return rows, meta


def _median(v):
    v = sorted(v)
    return v[len(v) // 2] if v else 0.0


def dispatch_sums(con: sqlite3.Connection):
    """{(dispatch id, counter): per-dispatch sum}, {dispatch id: kernel}"""
    per = defaultdict(float)
    names = {}
    for name, did, cn, v in con.execute("select name, dispatch_id, counter_name, counter_value from pmc_events"):
        per[(did, cn)] += v
        names[did] = short(name)
    return per, names


def halted_dispatches(per, names, kernel: str, frac: float | None) -> set:
    """Dispatches of `kernel` where any counter reads below frac x its median."""
    if not frac:
        return set()
    by_counter = defaultdict(dict)
    for (did, cn), v in per.items():
        if names[did] == kernel:
            by_counter[cn][did] = v
    out = set()
    for cn, d in by_counter.items():
        med = _median(list(d.values()))
        if med > 0:
            out |= {did for did, v in d.items() if v < frac * med}
    return out


def counter_means(con: sqlite3.Connection, counter: str, kernel: str | None = None, frac: float | None = None):
    per, names = dispatch_sums(con)
    skip = halted_dispatches(per, names, kernel, frac) if kernel else set()
    vals = defaultdict(list)
    for (did, cn), v in per.items():
        if cn == counter and did not in skip:
            vals[names[did]].append(v)
    return {k: sum(v) / len(v) for k, v in vals.items()}


def counter_totals(con: sqlite3.Connection, kernels=(), frac: float | None = None):
    """{kernel: {counter: mean over dispatches of the per-dispatch sum}}, the
    halted dispatches of the named kernels left out, and their counts"""
    per, names = dispatch_sums(con)
    skip = set()
    for k in kernels:
        skip |= halted_dispatches(per, names, k, frac)
    acc = defaultdict(lambda: defaultdict(list))
    for (did, cn), v in per.items():
        if did not in skip:
            acc[names[did]][cn].append(v)
    halted = defaultdict(int)
    for did in skip:
        halted[names[did]] += 1
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in acc.items()}, halted


# MI355X: 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies a 16-lane SIMD
# for 4 cycles; nominal engine clock 2.4 GHz
SIMDS, VALU_CYCLES, CLOCK_HZ = 1024, 4, 2.4e9


def counters_report(a) -> None:
    merged = defaultdict(dict)
    halted = defaultdict(int)
    for path in a.counters:
        totals, h = counter_totals(_db(path), a.kernel, a.halted_below)
        for k, cs in totals.items():
            merged[k].update(cs)
        for k, c in h.items():
            halted[k] = max(halted[k], c)
    durs = kernel_stats(_db(a.trace))[0] if a.trace else {}
    if a.halted_below:
        for k in list(durs):
            med = _median(durs[k])
            durs[k] = [d for d in durs[k] if d >= a.halted_below * med]
    print(f"# {a.title or 'Counters'}\n")
    if a.command:
        print("Commands (run from the repository root on one MI355X):\n")
        for c in a.command.split(" ;; "):
            print(f"    {c}")
        print()
    for k in a.kernel:
        cs = merged.get(k)
        if not cs:
            raise SystemExit(f"kernel {k!r} not in the counter runs")
        print(f"## `{k}`\n")
        print("| counter | per dispatch |")
        print("|---|---:|")
        for c in sorted(cs):
            print(f"| {c} | {cs[c]:.0f} |")
        waves = cs.get("SQ_WAVES")
        lines = []
        if waves:
            for c, what in (("SQ_INSTS_VALU", "VALU instructions"), ("SQ_INSTS_LDS", "LDS instructions"),
                            ("SQ_INSTS_VMEM_RD", "vector memory reads"), ("SQ_INSTS_SALU", "scalar instructions"),
                            ("TCP_TOTAL_CACHE_ACCESSES_sum", "vector L1 accesses")):
                if c in cs:
                    lines.append(f"- {what} per wave: {cs[c] / waves:.0f}")
        if "SQ_WAIT_ANY" in cs and "SQ_WAVE_CYCLES" in cs:
            lines.append(f"- waves waiting (SQ_WAIT_ANY / SQ_WAVE_CYCLES): {cs['SQ_WAIT_ANY'] / cs['SQ_WAVE_CYCLES']:.2f}")
        if "TCC_HIT_sum" in cs and "TCC_MISS_sum" in cs:
            lines.append(f"- L2 hit rate: {cs['TCC_HIT_sum'] / (cs['TCC_HIT_sum'] + cs['TCC_MISS_sum']):.2f}")
        if "SQ_LDS_BANK_CONFLICT" in cs and "SQ_ACTIVE_INST_LDS" in cs:
            lines.append(f"- LDS bank-conflict cycles per active LDS cycle: "
                         f"{cs['SQ_LDS_BANK_CONFLICT'] / cs['SQ_ACTIVE_INST_LDS']:.2f}")
        if "SQ_INSTS_VALU" in cs and k in durs:
            d = durs[k]
            dur_s = sum(d) / len(d) / 1e9
            frac = cs["SQ_INSTS_VALU"] / SIMDS * VALU_CYCLES / (dur_s * CLOCK_HZ)
            lines.append(f"- VALU issue fraction: SQ_INSTS_VALU / {SIMDS} SIMDs x {VALU_CYCLES} cycles / "
                         f"({dur_s * 1e6:.1f} us x {CLOCK_HZ / 1e9:.1f} GHz) = {frac:.2f}")
        if halted.get(k):
            lines.append(f"- halted dispatches left out: {halted[k]} (a counter below {a.halted_below} x its median)")
        if lines:
            print()
            print("\n".join(lines))
        print()


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--counters", nargs="+", help="--pmc result folders to tabulate (counters mode)")
    ap.add_argument("--trace", required=False)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--kernel", action="append", required=True,
                    help="short name of the dominant kernel (counters mode: repeat for several)")
    ap.add_argument("--units", type=int, help="work units (atoms) per launch")
    ap.add_argument("--bytes-per-unit", type=float)
    ap.add_argument("--command", default="")
    ap.add_argument("--title", default="")
    ap.add_argument("--period-kernel", default=None, help="kernel whose launches delimit the steps")
    ap.add_argument("--period-range", default=None, help="A:B, the steps A..B-1 of the timeline")
    ap.add_argument("--halted-below", type=float, default=None,
                    help="set apart launches of the named kernel below this fraction of its median")
    a = ap.parse_args()
    if a.counters:
        counters_report(a)
        return
    if not a.trace or a.units is None or a.bytes_per_unit is None:
        ap.error("--trace, --units and --bytes-per-unit are required without --counters")
    kernels = a.kernel            # one step: the sum of these kernels' mean launches
    a.kernel = " + ".join(kernels)

    rows, meta = kernel_stats(_db(a.trace))
    total = sum(sum(v) for v in rows.values())
    print(f"# {a.title or a.kernel}\n")
    if a.command:
        print("Commands (run from the repository root on one MI355X):\n")
        for c in a.command.split(" ;; "):
            print(f"    {c}")
        print()
    print("## Kernel trace (rocprofv3 --kernel-trace --stats)\n")
    print("| kernel | calls | total ms | avg us | % | grid | wg | LDS B | VGPR | SGPR |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|")
    for k, v in sorted(rows.items(), key=lambda kv: -sum(kv[1])):
        if sum(v) < 0.001 * total:
            continue
        gx, wx, lds, vg, ag, sg = meta[k]
        print(f"| `{k}` | {len(v)} | {sum(v) / 1e6:.3f} | {sum(v) / len(v) / 1e3:.1f} | "
              f"{100 * sum(v) / total:.1f} | {gx} | {wx} | {lds} | {vg} | {sg} |")
    avg_s, counts, short_ones = 0.0, [], []
    for k in kernels:
        durs = rows.get(k)
        if not durs:
            raise SystemExit(f"kernel {k!r} not in the trace")
        if a.halted_below:
            med = _median(durs)
            short_ones += [d for d in durs if d < a.halted_below * med]
            durs = [d for d in durs if d >= a.halted_below * med]
        avg_s += sum(durs) / len(durs) / 1e9
        counts.append(len(durs))
    alg = a.units * a.bytes_per_unit
    print(f"\n## Roofline of `{a.kernel}`\n")
    print(f"- average launch{'es (summed)' if len(kernels) > 1 else ''}: {avg_s * 1e6:.1f} us over "
          f"{' / '.join(map(str, counts))} launches")
    if short_ones:
        print(f"- halted launches (list gate set) left out: {len(short_ones)}, "
              f"{sum(short_ones) / len(short_ones) / 1e3:.1f} us each on average")
    print(f"- algorithmic bytes per launch: {a.units} units x {a.bytes_per_unit:g} B = {alg / 1e6:.1f} MB")
    print(f"- achieved (algorithmic / average launch): {alg / avg_s / 1e9:.0f} GB/s "
          f"= {alg / avg_s / 8e12:.3f} of 8 TB/s")
    if a.fetch and a.write and a.copy_fetch and a.copy_write and a.copy_bytes:
        cf = counter_means(_db(a.copy_fetch), "FETCH_SIZE")
        cw = counter_means(_db(a.copy_write), "WRITE_SIZE")
        ck = [k for k in cf if "copyBuffer" in k or "copy" in k.lower()]
        if not ck:
            raise SystemExit("no copy kernel in the calibration runs")
        f_copy, w_copy = cf[ck[0]], cw[ck[0]]
        kib = a.copy_bytes / 1024
        f_scale, w_scale = kib / f_copy, kib / w_copy
        fk = sum(counter_means(_db(a.fetch), "FETCH_SIZE", k, a.halted_below)[k] for k in kernels)
        wk = sum(counter_means(_db(a.write), "WRITE_SIZE", k, a.halted_below)[k] for k in kernels)
        traffic = (fk * f_scale + wk * w_scale) * 1024
        print("\n## HBM traffic (rocprofv3 --pmc, separate runs)\n")
        print(f"- copy calibration: a {a.copy_bytes} B device copy reads FETCH_SIZE {f_copy:.0f} KiB and "
              f"WRITE_SIZE {w_copy:.0f} KiB, so the scale factors are {f_scale:.3f} (fetch) and "
              f"{w_scale:.3f} (write)")
        print(f"- `{a.kernel}`: FETCH_SIZE {fk:.0f} KiB x {f_scale:.3f} + WRITE_SIZE {wk:.0f} KiB x "
              f"{w_scale:.3f} per launch")
        print(f"- traffic per launch: {traffic / 1e6:.1f} MB = {traffic / alg:.2f} x the algorithmic bytes")
        print(f"\ntraffic_bytes_per_launch: {int(round(traffic))}")
    if a.period_kernel:
        lo, hi = (int(x) for x in (a.period_range or "1:-1").split(":"))
        timeline_report(_db(a.trace), a.period_kernel, lo, hi)


if __name__ == "__main__":
    main()
