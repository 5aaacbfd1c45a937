"""Script-level drop-in for the reference drivers
(Correlation_in_CUDA/test_cuda.py, Correlation_in_MultiCUDA/test_multicuda.py,
Correlation_in_OMP/test_c.py): loads the events file, runs the chosen
library's ``compute_correlation`` on one MI355X, prints the timing lines and
writes the out_*.npz results.

  python -m seismic_amd.run_correlation cuda EVENTS.npy --shift 256 --threshold 0.2
  python -m seismic_amd.run_correlation omp EVENTS.npy --shift 250 --threshold 0.2 --classes 20
"""
from __future__ import annotations

import argparse
import time

from . import io


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("library", choices=["cuda", "multicuda", "omp"])
    ap.add_argument("events")
    ap.add_argument("--shift", type=int, required=True)
    ap.add_argument("--threshold", type=float, default=0.0)
    ap.add_argument("--threads", type=int, default=1, help="num_threads (accepted, unused on the GPU)")
    ap.add_argument("--chunk-size", type=int, default=1, help="OMP chunk_size (accepted, unused)")
    ap.add_argument("--classes", type=int, default=20, help="OMP n_classes_hist (< 2 -> 20, as test_c.py)")
    ap.add_argument("--gpus", type=int, default=1, help="MultiCUDA nGPUS (the job runs on one GPU)")
    ap.add_argument("--out", default=".")
    ap.add_argument("--allow-pickle", action="store_true",
                    help="load object arrays (pickled traces) -- only for trusted files")
    a = ap.parse_args(argv)
    events = io.load_events(a.events, allow_pickle=a.allow_pickle)
    start = time.time()
    if a.library == "cuda":
        from . import correlation_lib as lib
        lib.initialiseRunTime()
        res = lib.compute_correlation(events, a.shift, a.threads, a.threshold)
    elif a.library == "multicuda":
        from . import multicuda_lib as lib
        lib.initialiseRunTime(a.gpus)
        res = lib.compute_correlation(events, a.shift, a.gpus, a.threshold)
    else:
        from . import omp_lib as lib
        classes = a.classes if a.classes >= 2 else 20
        res = lib.compute_correlation(events, a.shift, a.threads, a.chunk_size, a.threshold, classes)
    diff = time.time() - start
    print("Execution finished in: " + str(diff))
    for p in io.save_results(a.out, *res):
        print("wrote " + p)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
