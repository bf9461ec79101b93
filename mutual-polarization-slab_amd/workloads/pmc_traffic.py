"""HBM traffic per kernel launch from two rocprofv3 counter runs (one with
FETCH_SIZE, one with WRITE_SIZE) of workloads.operator_loop --copy-mb M.

The counters (KiB per dispatch) are scaled by what they read for the
device-to-device copy of M MiB in the same run (known to move M MiB each
way), then averaged over the launches of each kernel.

    python -m workloads.pmc_traffic FETCH_DB WRITE_DB COPY_MB OUT_JSON
"""
from __future__ import annotations

import collections
import json
import sqlite3
import sys


def _counter(db, name):
    cur = sqlite3.connect(db).cursor()
    rows = cur.execute("select kernel_name, value from counters_collection where counter_name = ?", (name,)).fetchall()
    per = collections.defaultdict(list)
    for k, v in rows:
        per[k].append(float(v))
    return per


def _short(name, templated=False):
    base = name.split("(")[0]
    if not templated:
        base = base.split("<")[0]
    return base.replace("void ", "").replace("mpol::", "").strip()


def _keys(names):
    """Kernel name -> key: the bare name where one instance of the template
    ran, the name with its template arguments where several did (the forward
    and inverse instances of one y pass, say), so no instance hides another."""
    count = collections.Counter(_short(k) for k in names)
    return {k: _short(k, templated=count[_short(k)] > 1) for k in names}


def main(argv=None):
    fetch_db, write_db, copy_mb, out = (argv or sys.argv[1:])[:4]
    copy_bytes = float(copy_mb) * 2 ** 20
    fetch = _counter(fetch_db, "FETCH_SIZE")
    write = _counter(write_db, "WRITE_SIZE")
    copy_f = [v for k, vs in fetch.items() if "copy" in k.lower() for v in vs if v * 1024 > 0.5 * copy_bytes]
    copy_w = [v for k, vs in write.items() if "copy" in k.lower() for v in vs if v * 1024 > 0.5 * copy_bytes]
    if not copy_f or not copy_w:
        raise SystemExit("no copy kernel of the calibration size in the counter runs")
    sf = copy_bytes / (1024 * max(copy_f))
    sw = copy_bytes / (1024 * max(copy_w))
    res = {"copy_mb": float(copy_mb), "fetch_scale": sf, "write_scale": sw, "kernels": {}}
    common = sorted(k for k in set(fetch) & set(write) if "copy" not in k.lower())
    keys = _keys(common)
    for k in common:
        f = 1024 * sf * sum(fetch[k]) / len(fetch[k])
        w = 1024 * sw * sum(write[k]) / len(write[k])
        res["kernels"][keys[k]] = {"fetch_bytes": round(f), "write_bytes": round(w), "bytes_per_launch": round(f + w),
                                     "launches": len(fetch[k]), "full_name": k[:160]}
    with open(out, "w") as fh:
        json.dump(res, fh, indent=1)
    for k, v in res["kernels"].items():
        print(f"{k:48s} fetch {v['fetch_bytes'] / 1e6:9.2f} MB  write {v['write_bytes'] / 1e6:9.2f} MB")


if __name__ == "__main__":
    main()
