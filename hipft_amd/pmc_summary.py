"""HBM bytes per kernel launch from two rocprofv3 counter runs.

Usage (after the two `rocprofv3 --pmc FETCH_SIZE ...` / `--pmc WRITE_SIZE ...`
runs of `bench.py --pmc-probe`, each with `--output-format csv`):

    python -m hipft_amd.pmc_summary FETCH_DIR WRITE_DIR OUT_JSON OUT_MD [--copy-bytes N]
                                    [--probe PROBE_LOG]

FETCH_SIZE and WRITE_SIZE are reported in KiB per dispatch.  The probe's
first large device-to-device copy (N bytes, default 256 MiB; the largest
`copyBuffer` dispatch) calibrates them: scale = N / (counter x 1024), so a
copy of N bytes reads and writes N bytes after scaling.  The JSON feeds
bench.py's `roofline.traffic`; the markdown table is the readable summary.
PROBE_LOG is the output of one of the probe runs: its `pmc_probe` line
(the algorithmic bytes per fused launch of the probed step) is kept in the
JSON as `probe`, so measured and algorithmic bytes refer to the same
launches.
"""
import csv
import json
import os
import sys
from collections import defaultdict


def _load(d, counter):
    path = os.path.join(d, "run_counter_collection.csv")
    per = defaultdict(list)
    with open(path) as fh:
        for r in csv.DictReader(fh):
            if r["Counter_Name"] == counter:
                per[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return per


def summarize(fetch_dir, write_dir, copy_bytes=256 << 20):
    fe, wr = _load(fetch_dir, "FETCH_SIZE"), _load(write_dir, "WRITE_SIZE")
    copies = [k for k in fe if "copyBuffer" in k]
    if not copies:
        raise ValueError("no copyBuffer dispatch to calibrate against")
    fcal = max(max(fe[k]) for k in copies)
    wcal = max(max(wr.get(k, [0.0])) for k in copies)
    fscale = copy_bytes / (fcal * 1024.0)
    wscale = copy_bytes / (wcal * 1024.0) if wcal > 0 else 1.0
    kernels = {}
    for k, v in fe.items():
        if "rocclr" in k:
            continue
        w = wr.get(k, [0.0] * len(v))
        rd = 1024.0 * fscale * sum(v) / len(v)
        wb = 1024.0 * wscale * sum(w) / max(len(w), 1)
        kernels[k] = dict(dispatches=len(v), read_bytes_avg=round(rd),
                          write_bytes_avg=round(wb), hbm_bytes_per_launch=round(rd + wb))
    return dict(fetch_scale=round(fscale, 4), write_scale=round(wscale, 4),
                calibration=f"device-to-device copy of {copy_bytes} bytes: FETCH_SIZE "
                            f"{fcal:.0f} KiB, WRITE_SIZE {wcal:.0f} KiB",
                kernels=dict(sorted(kernels.items(),
                                    key=lambda kv: -kv[1]["hbm_bytes_per_launch"] *
                                    kv[1]["dispatches"])))


def markdown(s):
    lines = ["| kernel | dispatches | read MB/dispatch | write MB/dispatch |",
             "|---|---|---|---|"]
    for k, v in s["kernels"].items():
        lines.append(f"| `{k}` | {v['dispatches']} | {v['read_bytes_avg'] / 1e6:.1f} | "
                     f"{v['write_bytes_avg'] / 1e6:.1f} |")
    return "\n".join(lines)


def probe_line(path):
    """The last JSON line of a probe log that carries "pmc_probe"."""
    found = None
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if line.startswith("{") and '"pmc_probe"' in line:
                found = json.loads(line)
    if found is None:
        raise ValueError(f"no pmc_probe line in {path}")
    return found


def main(argv):
    opts = {}
    args = []
    it = iter(argv)
    for a in it:
        if a in ("--copy-bytes", "--probe"):
            opts[a] = next(it)
        else:
            args.append(a)
    copy_bytes = int(opts.get("--copy-bytes", 256 << 20))
    fetch_dir, write_dir, out_json, out_md = args[:4]
    s = summarize(fetch_dir, write_dir, copy_bytes)
    if "--probe" in opts:
        s["probe"] = probe_line(opts["--probe"])
    with open(out_json, "w") as fh:
        json.dump(s, fh, indent=1)
    with open(out_md, "w") as fh:
        fh.write(markdown(s) + "\n")
    print(markdown(s))


if __name__ == "__main__":
    main(sys.argv[1:])
