"""Script-level drop-in (seismic_amd.run_correlation) on the GPU: events
file in, out_*.npz files out, equal to calling the binding directly."""
import numpy as np
import pytest

from seismic_amd import correlation_lib, io, omp_lib, run_correlation, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("library", ["cuda", "multicuda", "omp"])
def test_driver_writes_reference_files(tmp_path, library):
    ev, _ = synth.make_events(30, 1, 700, seed=81)
    path = tmp_path / "events.npy"
    np.save(path, ev[:, 0])
    print(f"case driver {library}", flush=True)
    rc = run_correlation.main([library, str(path), "--shift", "256", "--threshold", "0.1",
                               "--classes", "10", "--out", str(tmp_path)])
    assert rc == 0
    got = io.load_results(str(tmp_path))
    if library == "omp":
        want = omp_lib.compute_correlation(ev[:, 0], 256, 1, 1, 0.1, 10)
        keys = ("xcm_pos", "xclags_pos", "xcm_neg", "xclags_neg", "max_hist", "min_hist")
    else:
        want = correlation_lib.compute_correlation(ev[:, 0], 256, 1, 0.1)
        keys = ("xcm_pos", "xclags_pos", "xcm_neg", "xclags_neg")
    for k, w in zip(keys, want):
        assert np.array_equal(got[k], w), k
