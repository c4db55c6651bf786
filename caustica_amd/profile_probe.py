"""Fixed-length run of one workload's time step, for hardware-counter passes.

``python -m caustica_amd.profile_probe --workload cfg1 --steps 30`` adds the workload
as one member, advances it ``--steps`` time steps (the last ``--record-steps``, default
5, record), then runs a
plain copy kernel over a buffer of known size (``cst_diag_copy``: 256 MiB read,
256 MiB written). Under ``rocprofv3 --pmc FETCH_SIZE`` (or ``WRITE_SIZE``)
the counter value of that calibration kernel converts the counter's units into bytes,
which scales the per-dispatch values of the k-space kernels into HBM bytes per launch.

``--summarize-stats DIR``, ``--summarize-traffic FETCH_DIR WRITE_DIR WORKLOAD`` and
``--summarize-sq A_DIR B_DIR`` turn rocprofv3 CSV output (``--output-format csv -o run``) into the short text tables kept
with the measurements; the traffic table ends with ``traffic_bytes_per_launch`` lines
(pass, workload, bytes) that bench.py reads.
"""

from __future__ import annotations

import argparse
import sys

CAL_BYTES = 256 << 20


# kernel-name prefix -> pass name used by the session's event timing and bench.py
PASS_OF = {"cst::k_xgrad": "xgrad", "cst::k_zu": "zu", "cst::k_xdiv": "xdiv", "cst::k_zp": "zp",
           "cst::k_ypass": "y", "cst::k_finalize": "finalize"}


def _short(name: str) -> str:
    return name.replace("(anonymous namespace)::", "cst::").split("(")[0].replace("void ", "")


def summarize_stats(d: str) -> str:
    """Markdown table of a rocprofv3 --stats kernel summary (run_kernel_stats.csv)."""
    import csv

    rows = list(csv.DictReader(open(f"{d}/run_kernel_stats.csv")))
    out = ["| kernel | calls | avg us | min us | max us | share % |", "|---|---:|---:|---:|---:|---:|"]
    for r in rows:
        out.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                   f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
                   f"{float(r['Percentage']):.2f} |")
    return "\n".join(out)


def summarize_traffic(fetch_dir: str, write_dir: str, workload: str) -> str:
    """Per-dispatch FETCH_SIZE / WRITE_SIZE scaled by the calibration copy kernel."""
    import csv
    from collections import defaultdict

    def load(d):
        agg = defaultdict(list)
        for r in csv.DictReader(open(f"{d}/run_counter_collection.csv")):
            agg[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        return {k: (sum(v) / len(v), len(v)) for k, v in agg.items()}

    f, w = load(fetch_dir), load(write_dir)
    cal = "cst::k_diag_copy"
    bf, bw = CAL_BYTES / f[cal][0], CAL_BYTES / w[cal][0]
    out = [f"calibration: {cal} over {CAL_BYTES} B reads FETCH_SIZE {f[cal][0]:.1f} and WRITE_SIZE "
           f"{w[cal][0]:.1f} -> {bf:.2f} and {bw:.2f} bytes per counter unit", "",
           "| kernel | dispatches | fetch MB | write MB | total MB |", "|---|---:|---:|---:|---:|"]
    lines = []
    for k, (fv, n) in f.items():
        if k == cal or not k.startswith("cst::"):
            continue
        wv = w.get(k, (0.0, 0))[0]
        tot = fv * bf + wv * bw
        out.append(f"| `{k}` | {n} | {fv * bf / 1e6:.1f} | {wv * bw / 1e6:.1f} | {tot / 1e6:.1f} |")
        name = next((v for p, v in PASS_OF.items() if k.startswith(p)), None)
        if name and name != "y":
            lines.append(f"traffic_bytes_per_launch {name} {workload} {tot:.0f}")
    return "\n".join(out + [""] + lines)


# MI355X: 256 CUs in 8 XCDs, 4 SIMDs per CU; a wave64 vector instruction occupies its
# SIMD for 4 cycles
N_CU, N_XCD, SIMD_CYCLES = 256, 8, 4


def summarize_sq(a_dir: str, b_dir: str) -> str:
    """Per-kernel shader-sequencer counters from two --pmc passes (SQ_* and GRBM_GUI_ACTIVE).

    rocprofv3 sums each counter over the 8 XCDs; GRBM_GUI_ACTIVE / 8 is the kernel's
    duration in cycles. VALU busy is the derived metric of the same name
    (SQ_ACTIVE_INST_VALU per CU over that duration); "VALU issue" is the share of the
    duration the kernel's vector instructions need at one wave64 instruction per SIMD every
    4 cycles; waves per CU is 4 SQ_WAVE_CYCLES (counted in quad cycles) per CU over the
    duration.
    """
    import csv
    from collections import defaultdict

    def load(d):
        agg = defaultdict(lambda: defaultdict(list))
        for r in csv.DictReader(open(f"{d}/run_counter_collection.csv")):
            agg[_short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
        return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in agg.items()}

    A, B = load(a_dir), load(b_dir)
    out = ["| kernel | waves | VALU instr / wave | VALU issue % | VALU busy % | waves / CU | "
           "LDS instr / wave | LDS bank-conflict cycles % of LDS active |",
           "|---|---:|---:|---:|---:|---:|---:|---:|"]
    for k, a in A.items():
        if not k.startswith("cst::"):
            continue
        b = B.get(k, {})
        cyc = a["GRBM_GUI_ACTIVE"] / N_XCD
        waves = a["SQ_WAVES"]
        vi = b.get("SQ_INSTS_VALU", 0.0)
        issue = 100 * vi * SIMD_CYCLES / (4 * N_CU) / cyc
        busy = 100 * a["SQ_ACTIVE_INST_VALU"] / N_CU / cyc
        occ = 4 * a["SQ_WAVE_CYCLES"] / N_CU / cyc
        conf = 100 * a["SQ_LDS_BANK_CONFLICT"] / max(a["SQ_ACTIVE_INST_LDS"], 1.0)
        out.append(f"| `{k}` | {waves:.0f} | {vi / waves:.0f} | {issue:.0f} | {busy:.0f} | {occ:.1f} | "
                   f"{a['SQ_INSTS_LDS'] / waves:.0f} | {conf:.0f} |")
    return "\n".join(out)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", default="cfg1", choices=["cfg0", "cfg1", "cfg2", "staggered", "pennes"],
                    help="k-space workload; staggered: the 256^3 medium under the opt-in "
                         "staggered power-law scheme (y = 1.1); or pennes: the thermal "
                         "benchmark's 256^3 tissue heated by a seeded random source, --steps "
                         "Pennes steps with the dose")
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--variant", type=int, default=0, help="z-pass thread layout (0 or 1)")
    ap.add_argument("--zu-variant", type=int, help="velocity-pass layout (0, 1 or 2; default --variant)")
    ap.add_argument("--zp-variant", type=int, help="pressure-pass layout (0 or 1; default --variant)")
    ap.add_argument("--record-steps", type=int, default=5, help="trailing steps that record")
    ap.add_argument("--summarize-stats", metavar="DIR")
    ap.add_argument("--summarize-traffic", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "WORKLOAD"))
    ap.add_argument("--summarize-sq", nargs=2, metavar=("A_DIR", "B_DIR"))
    args = ap.parse_args()
    if args.summarize_sq:
        print(summarize_sq(*args.summarize_sq))
        return 0
    if args.summarize_stats:
        print(summarize_stats(args.summarize_stats))
        return 0
    if args.summarize_traffic:
        print(summarize_traffic(*args.summarize_traffic))
        return 0

    from caustica_amd import _native as N
    from caustica_amd import workloads
    from caustica_amd.core.backend import get_backend
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.operators import plan_member

    get_backend("hip")
    if args.workload == "pennes":
        import numpy as np

        from caustica_amd.thermal import PennesSolver

        shape = (256, 256, 256)
        q = (1e6 * np.random.default_rng(1).random(shape)).astype(np.float32)
        print(f"[probe] pennes {shape}: {args.steps} steps", flush=True)
        r = PennesSolver("hip", finite_check_every=args.steps).solve(
            np.full(shape, 37.0, np.float32), q, workloads.pennes_tissue(shape, 3.75e-4), dt=0.05,
            n_steps=args.steps, dose=True)
        print(f"[probe] peak {float(r.temperature_max.max()):.3f} C", flush=True)
        return 0
    wl = {"cfg0": workloads.cfg0_water_bowl, "cfg1": workloads.cfg1_hetero_linear,
          "cfg2": workloads.cfg2_westervelt, "staggered": workloads.cfg1_hetero_linear}[args.workload]()
    if args.workload == "staggered":
        from caustica_amd.solvers.kspace.staggered import plan_staggered

        plan = plan_staggered(wl.grid, wl.medium, wl.source, wl.spec, harmonics=wl.harmonics,
                              reference_point=wl.reference_point, alpha_power=1.1)
    else:
        plan = plan_member(wl.grid, wl.medium, wl.source, wl.spec, harmonics=wl.harmonics,
                           reference_point=wl.reference_point, nonlinear=wl.solver == "westervelt")
    sess = KSpaceSession(plan.padded3, 1, scheme=plan.scheme)
    sess.set_graphs(False)  # counters are collected per dispatch of directly launched kernels
    zu = args.variant if args.zu_variant is None else args.zu_variant
    zp = args.variant if args.zp_variant is None else args.zp_variant
    sess.set_variant(2, zu)
    sess.set_variant(3, zp)
    slot = sess.add(plan, max_periods=args.steps // plan.spp + 4)
    if args.record_steps > 0:
        sess.schedule(slot, max(args.steps - args.record_steps, 0))
    print(f"[probe] {wl.name}: {args.steps} steps ({args.record_steps} recording), "
          f"layouts zu {zu} zp {zp}", flush=True)
    sess.advance([slot], args.steps)
    sess.sync()
    sess.close()

    print(f"[probe] calibration: copy kernel over {CAL_BYTES} B", flush=True)
    N.check(N.lib().cst_diag_copy(0, CAL_BYTES))
    print("[probe] done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
