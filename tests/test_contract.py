"""CPU checks of the measurement and loading contracts: bench.py refuses
more than one GPU with a message, its FP64 accuracy reference agrees with the
oracle, and the package fails loudly (no CPU fallback) when the native
library is missing."""
import importlib.util
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr and r.stdout.strip() == ""


def test_missing_library_fails_loudly(tmp_path):
    code = ("import seismic_amd._lib as L\n"
            "try:\n"
            "    L.load_engine()\n"
            "except RuntimeError as e:\n"
            "    print('raised:', e)\n")
    env = dict(os.environ, SEISMIC_AMD_LIB_DIR=str(tmp_path))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, env=env,
                       timeout=120)
    assert "raised:" in r.stdout and "is missing" in r.stdout


def test_bench_fp64_reference_matches_oracle(oracle):
    """The numpy FP64 reference bench.py states every form's error against
    equals the oracle's FP64 direct circular sums (full and windowed lags)."""
    from seismic_amd import synth
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    ev, _ = synth.make_events(20, 3, 300, seed=5)
    P, n, seed = 512, 50, 7
    idx, ref = bench.fp64_reference(ev, P, n, seed, [(-P // 2 + 1, P // 2), (-40, 40)])
    rng = np.random.Generator(np.random.PCG64(seed))
    a, b = rng.integers(0, 20, size=n), rng.integers(0, 20, size=n)
    pi, pj = np.minimum(a, b).astype(np.int32), np.maximum(a, b).astype(np.int32)
    assert np.array_equal(idx, pi.astype(np.int64) * 20 - pi.astype(np.int64) * (pi - 1) // 2 + (pj - pi))
    for lo, hi in [(-P // 2 + 1, P // 2), (-40, 40)]:
        o = oracle.pairs_f64(ev, pi, pj, lo, hi, P=P, circular=True)
        r = ref[(lo, hi)]
        assert np.abs(o["pos"] - r["pos"]).max() < 1e-12 and np.abs(o["neg"] - r["neg"]).max() < 1e-12
        assert np.array_equal(o["lag_pos"], r["lag_pos"]) and np.array_equal(o["lag_neg"], r["lag_neg"])
