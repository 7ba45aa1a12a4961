"""The msamd.mp processor package executable: spec, requirements, parameter
handling and .mda I/O paths that run without a GPU
(packages/ms3/mountainsort3_main.cpp:385-436,547-591,812-820)."""
import json
import os
import subprocess

import numpy as np
import pytest

from mountainsort_amd import mdaio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mountainsort_amd", "bin", "msamd.mp")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="msamd.mp not built")


def run(*args, check=False):
    r = subprocess.run([EXE, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    if check:
        assert r.returncode == 0, r.stderr
    return r


# (name, version, required params, {optional: default}) of the reference's processors
REFERENCE_SPECS = [
    ("ms3.bandpass_filter", "0.20", ["samplerate", "freq_min", "freq_max"],
     {"freq_wid": 1000, "quantization_unit": 0, "subsample_factor": 1}),
    ("ms3.whiten", "0.1", [], {"quantization_unit": 0}),
    ("ms3.compute_whitening_matrix", "0.11", [], {"channels": None}),
    ("ms3.apply_whitening_matrix", "0.1", [], {"quantization_unit": 0}),
    ("ms3.whiten_clips", "0.1", [], {"quantization_unit": 0}),
    ("ms3.extract_clips", "0.11", ["clip_size"], {"channels": None}),
    ("ms3.mask_out_artifacts", "0.1", [], {"threshold": 6, "interval_size": 2000}),
    ("ms3.compute_templates", "0.11", ["clip_size"], {"clusters": ""}),
]


def test_spec_lists_reference_processors():
    spec = json.loads(run("spec", check=True).stdout)
    by_name = {p["name"]: p for p in spec["processors"]}
    for name, version, req, opt in REFERENCE_SPECS:
        p = by_name[name]
        assert p["version"] == version
        params = {q["name"]: q for q in p["parameters"]}
        for r in req:
            assert params[r]["optional"] is False
        for k, d in opt.items():
            assert params[k]["optional"] is True
            if d is None:
                assert "default_value" not in params[k]
            else:
                assert params[k]["default_value"] == d
        assert p["exe_command"] == f"{EXE} {name} $(arguments)"
    assert "requirements_command" in by_name["ms3.bandpass_filter"]
    assert "requirements_command" in by_name["ms3.whiten"]
    assert "requirements_command" not in by_name["ms3.mask_out_artifacts"]
    # stage processors use the ms3alg parameter names and defaults
    det = {q["name"]: q.get("default_value") for q in by_name["msamd.detect_events"]["parameters"]}
    assert det["detect_threshold"] == 3 and det["detect_interval"] == 10
    assert det["detect_sign"] == 0 and det["t1"] == 0 and det["t2"] == -1
    pca = {q["name"]: q.get("default_value") for q in by_name["msamd.pca_features"]["parameters"]}
    assert pca == {"adjacency_radius": -1, "clip_size": 50, "num_features": 10,
                   "num_features_per_channel": 10, "max_pca_samples": 10000}
    inputs = {f["name"]: f["optional"] for f in by_name["msamd.pca_features"]["inputs"]}
    assert inputs == {"timeseries": False, "event_times": False, "geom": True}
    fit = {q["name"]: q.get("default_value") for q in by_name["msamd.fit_stage"]["parameters"]}
    assert fit == {"clip_size": 50, "reference_num_threads": 0}


def test_spec_of_one_processor_and_unknown_name():
    p = json.loads(run("spec", "ms3.whiten", check=True).stdout)
    assert p["name"] == "ms3.whiten"
    r = run("spec", "ms3.nonexistent")
    assert r.returncode == 255 and json.loads(r.stdout) == {}
    r = run("ms3.nonexistent")
    assert r.returncode == 255 and "Unexpected processor name" in r.stderr


def test_requirements():
    r = run("requirements", "ms3.bandpass_filter", "--timeseries=x.mda", "--timeseries_out=y.mda",
            "--samplerate=30000", "--freq_min=300", "--freq_max=6000", check=True)
    assert json.loads(r.stdout)["peak_ram_mb"] > 0
    r = run("requirements", "ms3.mask_out_artifacts")
    assert r.returncode == 255


def test_missing_required_parameter():
    r = run("ms3.bandpass_filter", "--timeseries=x.mda", "--timeseries_out=y.mda",
            "--samplerate=30000", "--freq_min=300")
    assert r.returncode == 255 and "freq_max" in r.stderr


@pytest.mark.parametrize("args", [["--threshold=0"], ["--interval_size", "0"], ["--threshold=0.0"]])
def test_mask_rejects_zero_parameters(tmp_path, args):
    X = np.zeros((10, 2), np.float32)
    mdaio.writemda(tmp_path / "x.mda", X)
    r = run("ms3.mask_out_artifacts", f"--timeseries={tmp_path / 'x.mda'}",
            f"--timeseries_out={tmp_path / 'y.mda'}", *args)
    assert r.returncode == 255
    assert "Either threshold or interval_size is zero" in r.stderr
    assert not (tmp_path / "y.mda").exists()


def test_extract_clips_rejects_matrix_event_times(tmp_path):
    mdaio.writemda(tmp_path / "x.mda", np.zeros((100, 3), np.float32))
    mdaio.writemda(tmp_path / "et.mda", np.ones((4, 2), np.float64))
    r = run("ms3.extract_clips", f"--timeseries={tmp_path / 'x.mda'}",
            f"--event_times={tmp_path / 'et.mda'}", f"--clips_out={tmp_path / 'c.mda'}",
            "--clip_size=10")
    assert r.returncode == 255 and "must be a vector" in r.stderr
    mdaio.writemda(tmp_path / "et.mda", np.ones(4, np.float64))
    r = run("ms3.extract_clips", f"--timeseries={tmp_path / 'x.mda'}",
            f"--event_times={tmp_path / 'et.mda'}", f"--clips_out={tmp_path / 'c.mda'}",
            "--clip_size=0")
    assert r.returncode == 255 and "Clip size is zero" in r.stderr


def test_bandpass_with_freq_max_zero_copies_the_file(tmp_path):
    X = np.arange(40, dtype=np.int16).reshape(10, 4)
    mdaio.writemda(tmp_path / "x.mda", X)
    run("ms3.bandpass_filter", f"--timeseries={tmp_path / 'x.mda'}",
        f"--timeseries_out={tmp_path / 'y.mda'}", "--samplerate=30000", "--freq_min=300",
        "--freq_max=0", check=True)
    assert (tmp_path / "y.mda").read_bytes() == (tmp_path / "x.mda").read_bytes()
    assert not (tmp_path / "y.mda.tmp").exists()


def test_value_after_space_and_repeated_names(tmp_path):
    # "--name value" form, and a value token that is not consumed by a flag
    X = np.arange(40, dtype=np.int16).reshape(10, 4)
    mdaio.writemda(tmp_path / "x.mda", X)
    run("ms3.bandpass_filter", "--timeseries", str(tmp_path / "x.mda"),
        "--timeseries_out", str(tmp_path / "y.mda"), "--samplerate", "30000",
        "--freq_min", "300", "--freq_max", "0", check=True)
    assert (tmp_path / "y.mda").exists()


def test_mdaio_roundtrip(tmp_path):
    for dt in [np.float32, np.int16, np.float64, np.uint8, np.int32, np.uint16, np.uint32]:
        a = (np.arange(24) % 7).astype(dt).reshape(2, 3, 4)
        mdaio.writemda(tmp_path / "a.mda", a)
        b = mdaio.readmda(tmp_path / "a.mda")
        assert b.dtype == a.dtype and b.shape == a.shape and np.array_equal(a, b)
    raw = (tmp_path / "a.mda").read_bytes()
    assert np.frombuffer(raw[:20], "<i4").tolist() == [-8, 4, 3, 4, 3]
