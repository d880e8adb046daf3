"""Summaries of rocprofv3 results (rocpd SQLite databases) for the solver's kernels.

    python -m persiansph_amd.profiling trace RESULTS.db [--match k_]
        kernel table: calls, total, mean duration (from the database's top_kernels view)
    python -m persiansph_amd.profiling counters RESULTS.db [RESULTS.db ...] [--match force]
        per-kernel mean of every collected counter per dispatch
    python -m persiansph_amd.profiling traffic --fetch F.db --write W.db --copy-fetch CF.db
            --copy-write CW.db --copy-bytes BYTES --kernels k_nlist,k_force11_nl [--out T.json]
        HBM bytes per dispatch of the named kernels: FETCH_SIZE and WRITE_SIZE (KiB per
        dispatch) each scaled by bytes / counter of a plain device copy of known size; per
        pass, a kernel dispatched on some passes only counts by its share of the passes

FETCH_SIZE/WRITE_SIZE count the L2's traffic to memory; a copy of B bytes moves B in and B
out, so B / (counter of the copy in bytes) calibrates each counter.
"""
import argparse
import json
import sqlite3
import sys


def kernel_table(db_path):
    db = sqlite3.connect(db_path)
    rows = list(db.execute("select name, total_calls, total_duration, average, percentage from top_kernels"))
    return [{"kernel": n, "calls": int(c), "total_us": t, "mean_us": a, "percent": p}
            for n, c, t, a, p in rows]


def counters(db_paths):
    """{kernel: {counter: (mean per dispatch, dispatches)}} over all databases."""
    out = {}
    for path in db_paths:
        db = sqlite3.connect(path)
        for k, c, v, n in db.execute("select kernel_name, counter_name, sum(value), count(*) "
                                     "from counters_collection group by kernel_name, counter_name"):
            out.setdefault(k, {})[c] = (v/n, n)
    return out


def _counter_mean(path, counter, match):
    """{kernel: (mean per dispatch, dispatches)} of one counter."""
    got = {}
    for k, cs in counters([path]).items():
        if match(k) and counter in cs:
            got[k] = cs[counter]
    return got


def _is_copy(k):
    return "copyBuffer" in k or "elementwise" in k or "copy" in k.lower()


def traffic(fetch_db, write_db, copy_fetch_db, copy_write_db, copy_bytes, kernels):
    """HBM bytes per dispatch of each kernel whose name contains one of `kernels`, and per pass:
    kernels dispatched on some passes only (list builds) count in proportion to their
    dispatches against the most frequent kernel's."""
    cf = _counter_mean(copy_fetch_db, "FETCH_SIZE", _is_copy)
    cw = _counter_mean(copy_write_db, "WRITE_SIZE", _is_copy)
    # the copy kernel is the one with the largest counter (the probe's own launch)
    f_copy = max(v[0] for v in cf.values())*1024.0
    w_copy = max(v[0] for v in cw.values())*1024.0
    f_scale = copy_bytes/f_copy
    w_scale = copy_bytes/w_copy
    want = lambda k: any(s in k for s in kernels)
    fe = _counter_mean(fetch_db, "FETCH_SIZE", want)
    wr = _counter_mean(write_db, "WRITE_SIZE", want)
    per = {}
    for k in sorted(set(fe) | set(wr)):
        rb = fe.get(k, (0.0, 0))[0]*1024.0*f_scale
        wb = wr.get(k, (0.0, 0))[0]*1024.0*w_scale
        n = max(fe.get(k, (0.0, 0))[1], wr.get(k, (0.0, 0))[1])
        per[k] = {"read_bytes": rb, "write_bytes": wb, "bytes": rb + wb, "dispatches": n}
    nmax = max((v["dispatches"] for v in per.values()), default=1) or 1
    return {
        "fetch_scale": f_scale, "write_scale": w_scale,
        "copy_bytes": copy_bytes,
        "kernels": per,
        "passes": nmax,
        "bytes_per_pass": sum(v["bytes"]*v["dispatches"]/nmax for v in per.values()),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m persiansph_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--match", default="")
    c = sub.add_parser("counters")
    c.add_argument("db", nargs="+")
    c.add_argument("--match", default="")
    r = sub.add_parser("traffic")
    r.add_argument("--fetch", required=True)
    r.add_argument("--write", required=True)
    r.add_argument("--copy-fetch", required=True)
    r.add_argument("--copy-write", required=True)
    r.add_argument("--copy-bytes", type=float, required=True)
    r.add_argument("--kernels", required=True, help="comma-separated name fragments")
    r.add_argument("--out", default="")
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        print("| kernel | calls | total (us) | mean (us) | % |")
        print("|---|---|---|---|---|")
        for row in kernel_table(a.db):
            if a.match in row["kernel"]:
                print(f"| `{row['kernel']}` | {row['calls']} | {row['total_us']:.1f} | "
                      f"{row['mean_us']:.1f} | {row['percent']:.1f} |")
    elif a.cmd == "counters":
        for k, cs in sorted(counters(a.db).items()):
            if a.match not in k:
                continue
            print(k)
            for name, (v, n) in sorted(cs.items()):
                print(f"    {name:28s} {v:14.6g}  (mean of {n} dispatches)")
    else:
        res = traffic(a.fetch, a.write, a.copy_fetch, a.copy_write, a.copy_bytes, a.kernels.split(","))
        js = json.dumps(res, indent=1)
        if a.out:
            with open(a.out, "w") as f:
                f.write(js + "\n")
        print(js)
    return 0


if __name__ == "__main__":
    sys.exit(main())
