"""Short text summaries of rocprofv3 runs of bench.py: per-kernel time, HBM traffic and shader
counters, and the measured-traffic entry bench.py reports beside its roofline.

Expects, under one directory, the CSV output (`--output-format csv -o run`) of
  kt_<workload>/                  rocprofv3 --kernel-trace --stats
  pmc_<workload>_FETCH_SIZE/      rocprofv3 --pmc FETCH_SIZE
  pmc_<workload>_WRITE_SIZE/      rocprofv3 --pmc WRITE_SIZE
  pmc_<workload>_SQ_WAVES/        rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU ... (optional)
  cal_FETCH_SIZE/, cal_WRITE_SIZE/  the same two counters over bench.py --calibrate-copy
each over the commands printed in the summary, and the bench JSON line of the same tree.

    python -m mmrecode_amd.rocprof_summary --workload mpeg2 --root <dir> --bench <bench.json> \\
        --out profiles/mpeg2_bench_rocprof.md --traffic profiles/traffic.json

FETCH_SIZE and WRITE_SIZE are scaled by what they read for a device-to-device copy of 1 GiB (the
HIP runtime's copy kernel reads and writes exactly that many bytes). Steps in a counter run are
counted by the fill kernel every submission launches once.
"""
import argparse
import collections
import csv
import json
import os

PREFIX = "cd /tmp && export TMPDIR=/tmp && "
HELPERS = ("__amd_rocclr", "fill_zero_kernel")
SQ = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_WAIT_INST_ANY",
      "SQ_INSTS_SALU", "SQ_INSTS_LDS")


def short(name):
    """Kernel name without its parameter list (`void mmr::mpeg2_kernel<32>` -> mpeg2_kernel)."""
    n = name.split("(")[0].replace("void ", "").strip()
    return n.split("::")[-1].split("<")[0]


def rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def counters(d):
    """kernel -> counter -> list of per-dispatch values; kernel -> dispatch count."""
    vals = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    for r in rows(os.path.join(d, "run_counter_collection.csv")):
        k = short(r["Kernel_Name"])
        vals[k][(r["Dispatch_Id"], r["Counter_Name"])] += float(r["Counter_Value"])
        disp[k].add(r["Dispatch_Id"])
    out = {}
    for k, v in vals.items():
        per = collections.defaultdict(list)
        for (_, c), x in v.items():
            per[c].append(x)
        out[k] = per
    return out, {k: len(v) for k, v in disp.items()}


def bench_cmd(workload, steps, warmup):
    return (f"python3 <repo>/bench.py --workload {workload} --steps {steps} --warmup {warmup} "
            "--no-plain --no-cpu --no-autotune")


def summarize(workload, root, bench_path, out_path, traffic_path):
    bench = json.loads(open(bench_path).read().strip().splitlines()[-1])
    roof = bench["roofline"]
    dominant = roof["kernel"]
    alg = {p["kernel"]: p.get("algorithmic_bytes") for p in roof.get("phases", [])}
    lines = [f"# rocprofv3 summary — {bench['config']['workload']}", ""]
    ph = ", ".join(f"`{p['kernel']}` {p['ms']:.4f} ms" for p in roof.get("phases", []))
    lines += [f"Bench line of the same tree: {bench['value']:.1f} {bench['unit']}, {bench['ms_per_step']:.4f} ms/step; "
              f"event-timed launch phases per step: {ph}. Roofline kernel `{dominant}`: "
              f"{roof['algorithmic_bytes_per_launch'] / 1e9:.4f} GB algorithmic per step in {roof['kernel_ms']:.4f} ms = "
              f"{roof['achieved']:.1f} GB/s ({roof['frac']:.3f} of 8000 GB/s).", ""]
    if roof.get("fp64_dense_tflops_equiv"):
        lines += [f"IDCT arithmetic: {roof['fp64_dense_tflops_equiv']:.2f} TFLOP/s of the dense f64 count "
                  f"(2048 per coded block) against {roof['fp64_peak_tflops']} TFLOP/s f64 peak.", ""]

    # ---- kernel trace
    kt = os.path.join(root, f"kt_{workload}")
    stats = rows(os.path.join(kt, "run_kernel_stats.csv"))
    lines += ["Kernel trace:", "", "```", PREFIX + "rocprofv3 --kernel-trace --stats --output-format csv -d <out> -o run -- "
              + bench_cmd(workload, 5, 2), "```", "",
              "| kernel | calls | avg µs | min µs | max µs | total ms | % |", "|---|---:|---:|---:|---:|---:|---:|"]
    avg_us = {}
    for r in stats:
        k = short(r["Name"])
        avg_us[k] = float(r["AverageNs"]) / 1e3
        lines.append(f"| `{k}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                     f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['TotalDurationNs']) / 1e6:.2f} | {float(r['Percentage']):.1f} |")
    lines.append("")

    # ---- calibration
    scale = {}
    cal_raw = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        v, _ = counters(os.path.join(root, f"cal_{c}"))
        copies = [x for k, per in v.items() if k.startswith("__amd_rocclr_copyBuffer") for x in per[c]]
        big = max(copies)
        cal_raw[c] = big
        scale[c] = (1 << 30) / (big * 1024)

    # ---- traffic per kernel
    fetch, nf = counters(os.path.join(root, f"pmc_{workload}_FETCH_SIZE"))
    write, nw = counters(os.path.join(root, f"pmc_{workload}_WRITE_SIZE"))
    subs = bench["config"].get("submissions_per_step", 1)
    steps_f, steps_w = nf.get("fill_zero_kernel", subs) / subs, nw.get("fill_zero_kernel", subs) / subs
    lines += [f"HBM traffic, two counter passes (each over every submission of the run; {subs} submission(s) per step, "
              "per step = total / steps):",
              "", "```", PREFIX + "rocprofv3 --pmc FETCH_SIZE --output-format csv -d <out> -o run -- " + bench_cmd(workload, 2, 1),
              PREFIX + "rocprofv3 --pmc WRITE_SIZE --output-format csv -d <out> -o run -- " + bench_cmd(workload, 2, 1), "```", "",
              "| kernel | dispatches / step | FETCH KiB / step (raw) | WRITE KiB / step (raw) | HBM bytes / step | algorithmic | ratio |",
              "|---|---:|---:|---:|---:|---:|---:|"]
    kernels = {}
    for k in fetch:
        if k.startswith(HELPERS) or k not in write:
            continue
        f_kib = sum(fetch[k]["FETCH_SIZE"]) / steps_f
        w_kib = sum(write[k]["WRITE_SIZE"]) / steps_w
        fb, wb = f_kib * 1024 * scale["FETCH_SIZE"], w_kib * 1024 * scale["WRITE_SIZE"]
        a = alg.get(k)
        kernels[k] = {"dispatches_per_step": round(nf[k] / steps_f, 2), "fetch_bytes_per_step": int(fb),
                      "write_bytes_per_step": int(wb), "hbm_bytes_per_step": int(fb + wb),
                      "avg_dispatch_us": round(avg_us.get(k, 0.0), 2)}
        lines.append(f"| `{k}` | {nf[k] / steps_f:g} | {f_kib:.0f} | {w_kib:.0f} | {fb + wb:.4g} | "
                     f"{a:.4g} | {(fb + wb) / a:.2f} |" if a else
                     f"| `{k}` | {nf[k] / steps_f:g} | {f_kib:.0f} | {w_kib:.0f} | {fb + wb:.4g} | — | — |")
    if dominant not in kernels and roof.get("phases"):
        # the roofline covers every kernel of the step: its traffic is the sum over them
        tot = {c: sum(v[c] for v in kernels.values()) for c in ("fetch_bytes_per_step", "write_bytes_per_step", "hbm_bytes_per_step")}
        tot["dispatches_per_step"] = round(sum(v["dispatches_per_step"] for v in kernels.values()), 2)
        kernels[dominant] = tot
        a = roof["algorithmic_bytes_per_launch"]
        lines.append(f"| all kernels | {tot['dispatches_per_step']:g} | | | {tot['hbm_bytes_per_step']:.4g} | {a:.4g} | "
                     f"{tot['hbm_bytes_per_step'] / a:.2f} |")
    if dominant in kernels and "avg_dispatch_us" in kernels[dominant]:
        d = kernels[dominant]
        ev = roof["kernel_ms"] * 1e3 / d["dispatches_per_step"]
        lines += ["", f"`{dominant}`: {d['dispatches_per_step']:g} dispatch(es) per step; the bench's event time per step / "
                  f"dispatches = {ev:.1f} µs per dispatch, the trace's average {d['avg_dispatch_us']:.1f} µs "
                  f"({(d['avg_dispatch_us'] / ev - 1) * 100:+.1f} % under the tracer)."]
    lines += ["", f"Scale factors from a 1 GiB device-to-device copy: FETCH_SIZE reads {cal_raw['FETCH_SIZE']:.0f} KiB "
              f"(x{scale['FETCH_SIZE']:.4f}), WRITE_SIZE {cal_raw['WRITE_SIZE']:.0f} KiB (x{scale['WRITE_SIZE']:.4f}) per copy:",
              "", "```", PREFIX + "rocprofv3 --pmc FETCH_SIZE --output-format csv -d <out> -o run -- python3 <repo>/bench.py --calibrate-copy",
              PREFIX + "rocprofv3 --pmc WRITE_SIZE --output-format csv -d <out> -o run -- python3 <repo>/bench.py --calibrate-copy",
              "```", ""]

    # ---- shader counters
    sq_dir = os.path.join(root, f"pmc_{workload}_SQ_WAVES")
    if os.path.isdir(sq_dir):
        sq, nsq = counters(sq_dir)
        ks = [k for k in sq if not k.startswith(HELPERS)]
        lines += ["Shader counters per wave (sums over every dispatch of the run divided by SQ_WAVES):", "", "```",
                  PREFIX + f"rocprofv3 --pmc {' '.join(SQ)} --output-format csv -d <out> -o run -- " + bench_cmd(workload, 2, 1),
                  "```", "", "| counter | " + " | ".join(f"`{k}`" for k in ks) + " |", "|---|" + "---:|" * len(ks)]
        for c in SQ:
            cells = []
            for k in ks:
                waves = sum(sq[k]["SQ_WAVES"])
                tot = sum(sq[k][c])
                cells.append(f"{tot / nsq[k]:.0f} per dispatch" if c == "SQ_WAVES" else f"{tot / waves:.1f}")
            lines.append(f"| {c} | " + " | ".join(cells) + " |")
        lines.append("")
        for k in ks:
            if k in kernels:
                waves = sum(sq[k]["SQ_WAVES"])
                kernels[k]["valu_per_wave"] = round(sum(sq[k]["SQ_INSTS_VALU"]) / waves, 1)
                kernels[k]["salu_per_wave"] = round(sum(sq[k]["SQ_INSTS_SALU"]) / waves, 1)

    with open(out_path, "w") as f:
        f.write("\n".join(lines))
    if traffic_path and dominant in kernels:
        try:
            traffic = json.load(open(traffic_path))
        except (OSError, ValueError):
            traffic = {}
        traffic["factors"] = {"FETCH_SIZE": scale["FETCH_SIZE"], "WRITE_SIZE": scale["WRITE_SIZE"],
                              "calibration": "1 GiB device-to-device hipMemcpy (bench.py --calibrate-copy): FETCH_SIZE "
                                             f"{cal_raw['FETCH_SIZE']:.0f} KiB, WRITE_SIZE {cal_raw['WRITE_SIZE']:.0f} KiB per copy"}
        entry = dict(kernels[dominant])
        entry["kernel"] = dominant
        entry["source"] = f"{os.path.relpath(out_path, os.path.dirname(os.path.dirname(os.path.abspath(out_path))))}: " \
                          "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE x calibration factors"
        entry["kernels"] = kernels
        traffic[workload] = entry
        with open(traffic_path, "w") as f:
            json.dump(traffic, f, indent=1)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", required=True)
    ap.add_argument("--root", required=True)
    ap.add_argument("--bench", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--traffic")
    a = ap.parse_args()
    summarize(a.workload, a.root, a.bench, a.out, a.traffic)


if __name__ == "__main__":
    main()
