"""CPU checks of the measurement and loading contracts: bench.py refuses
more than one GPU with a message, and the package fails loudly (no CPU
fallback) when the native library is missing."""
import os
import subprocess
import sys

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
