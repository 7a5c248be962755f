"""CPU calibration of the oracle against the compiled reference (test infrastructure).

Times the reference's own CPU path (oracle/_ref/ref_driver: doCalculation of the reference
engine, calc time as the engine reports it, median of 3) and the oracle's image
(oracle/liboracle.so, median of 3) on the same parameter sets and thread counts, on this
machine, and writes oracle/cpu_calibration.json.  bench.py's CPU baseline times the oracle
on the GPU host and reports this ratio beside it, so that its number can be read as the
reference's.

    make -C oracle all ref && python oracle/calibrate.py
"""
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

import oracle  # noqa: E402
from helpers import CFG1, CFG3  # noqa: E402
from make_golden import run_ref  # noqa: E402


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    cases = [("config 3, 1024x1024 FCC spheres", CFG3, 8), ("config 3 workload at 256x256", dict(CFG3, GridPoints=128), 1),
             ("config 1, 128x128 spheres", CFG1, 1)]
    rows = []
    for name, cfg, th in cases:
        r = run_ref(cfg, {"threads": th, "repeat": 3})
        ref_ms = float(np.median(r["calc_ms"]))
        o = oracle.OracleEngine(cfg)
        o.prepare()
        ts = []
        for _ in range(3):
            t = time.perf_counter()
            o.image(nthreads=th)
            ts.append(time.perf_counter() - t)
        orc_ms = 1e3 * float(np.median(ts))
        rows.append(dict(case=name, threads=th, reference_calc_ms=ref_ms, oracle_ms=orc_ms,
                         oracle_over_reference=orc_ms / ref_ms))
        print(rows[-1], flush=True)
    out = dict(machine=cpu_model(), cpus=os.cpu_count(), compiler="g++ -O2 (both)", rows=rows,
               oracle_over_reference=float(np.median([r["oracle_over_reference"] for r in rows])))
    with open(os.path.join(HERE, "cpu_calibration.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
