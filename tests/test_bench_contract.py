"""bench.py's contract, checked on the CPU: a run on more than one GPU is
refused with a message, and the table of measured counters covers the
default Cornell step, so the default run reports the roofline's issue
fraction, lane utilization and HBM traffic."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", BENCH)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_counter_table_covers_default_step():
    b = load_bench()
    # the default step: 1024^2, 32 spp, the whole pass in the persistent tail
    # over the flat leaf array (the shape the autotuner picks on this scene)
    key = ("k_tail", 1024, 32, b.TAIL_ALWAYS, 2)
    assert key in b.PMC and b.PMC[key]
    row = b.PMC[key]
    assert row["hbm_bytes"] > 0 and row["valu_insts"] > 0
    # rocprof's VALUBusy assumes 4 cycles per wave64 instruction, so paired
    # issue reads up to 200
    assert 0 < row["valu_util"] <= 100 and 0 < row["valu_busy"] <= 201
    assert b.HBM_PEAK_GBS == 8000.0
    # one kind of VALU instruction: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles;
    # two kinds paired: twice that
    assert abs(b.VALU_SINGLE_GINST - 614.4) < 1e-9
    assert abs(b.VALU_PEAK_GINST - 1228.8) < 1e-9
    # the counted instructions cannot issue faster than the paired peak at the
    # kernel time the counters were taken with (7.6-8.0 ms per launch)
    assert row["valu_insts"] / 7.6e-3 / 1e9 < b.VALU_PEAK_GINST
