"""Summaries of a rocprofv3 kernel trace (``--kernel-trace --output-format
csv``, the ``*_kernel_trace.csv`` file).

Per kernel name: dispatches, summed and mean dispatch duration, and the
*busy time*, the union of the dispatches' [start, end) intervals.  When the
engine runs a batch's shots on several streams (lanes), dispatches of one
kernel overlap in time; their mean duration is then that of concurrent
kernels, while the busy time is the time the GPU spent running that kernel
at all.  Busy time divided by the number of fine steps is the figure the
engine's pass-level timing reports per launch.

    python -m acoustic_isotropic_operators_amd.profiling TRACE.csv [--match k_fwd_pair k_step]
"""
from __future__ import annotations

import argparse
import csv
from collections import defaultdict


def _col(header, *names):
    for n in names:
        if n in header:
            return header.index(n)
    raise KeyError(f"none of {names} in the trace header")


def summarize(path, match=None):
    """{kernel name: {"dispatches", "sum_ns", "mean_ns", "busy_ns", "span_ns"}}
    over the dispatches whose name contains one of `match` (all if None)."""
    iv = defaultdict(list)
    with open(path, newline="") as f:
        r = csv.reader(f)
        header = next(r)
        kn = _col(header, "Kernel_Name", "KernelName", "Name")
        ks = _col(header, "Start_Timestamp", "BeginNs", "Start")
        ke = _col(header, "End_Timestamp", "EndNs", "End")
        for row in r:
            name = row[kn]
            if match and not any(m in name for m in match):
                continue
            iv[name].append((int(row[ks]), int(row[ke])))
    out = {}
    for name, lst in iv.items():
        lst.sort()
        busy, cs, ce = 0, None, None
        for s, e in lst:
            if ce is None or s > ce:
                if ce is not None:
                    busy += ce - cs
                cs, ce = s, e
            else:
                ce = max(ce, e)
        busy += ce - cs
        tot = sum(e - s for s, e in lst)
        out[name] = {"dispatches": len(lst), "sum_ns": tot, "mean_ns": tot / len(lst), "busy_ns": busy,
                     "span_ns": lst[-1][1] - lst[0][0]}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("trace")
    ap.add_argument("--match", nargs="*", default=None, help="keep kernels whose name contains one of these")
    a = ap.parse_args()
    s = summarize(a.trace, a.match)
    print("| kernel | dispatches | mean duration (us) | summed (ms) | busy (ms) | summed / busy |")
    print("|---|---:|---:|---:|---:|---:|")
    for name, v in sorted(s.items(), key=lambda kv: -kv[1]["busy_ns"]):
        print(f"| `{name[:70]}` | {v['dispatches']} | {v['mean_ns'] / 1e3:.2f} | {v['sum_ns'] / 1e6:.1f} | "
              f"{v['busy_ns'] / 1e6:.1f} | {v['sum_ns'] / max(v['busy_ns'], 1):.2f} |")


if __name__ == "__main__":
    main()
