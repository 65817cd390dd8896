"""CPU checks of bench.py's contract: more than one GPU is refused with a
message, and the algorithmic byte counts behind the roofline match the
per-product figures of the three phases (and the HBM counter summary's
algorithmic column for a batch of 128)."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_phase_bytes_per_product():
    bench = load_bench()
    n, k = 1 << 15, 59
    per = bench.phase_bytes(n, k)
    # conversion: a and b in (112 B per coefficient), 2k residue rows out
    assert per["conversion"] == 2 * n * 112 + 2 * k * n * 4
    # tile: both operands' residues in, the product's residues out
    assert per["tile"] == 3 * k * n * 4 == 23_199_744
    # CRT: k residue rows in, 112-byte coefficients out
    assert per["crt"] == k * n * 4 + n * 112
    assert bench.algorithmic_bytes(n, k) == sum(per.values())


def test_traffic_summary_matches_the_formula():
    bench = load_bench()
    path = os.path.join(ROOT, "profiles", "traffic_batched_n32768_b128.json")
    phases = json.load(open(path))["phases"]
    per = bench.phase_bytes(1 << 15, 59)
    for name in ("conversion", "tile", "crt"):
        assert phases[name]["algorithmic_bytes_per_launch"] == per[name] * 128, name
        # measured HBM bytes within 15 % of the compulsory traffic (the CRT's
        # reads run 10-15 % over, the other phases 1-2 %)
        assert 1.0 <= phases[name]["traffic_over_algorithmic"] < 1.15, name


def test_roofline_names_the_binding_roof():
    """With the dominant kernel's VALU instruction count the record's bound is
    the VALU issue rate (frac = busy fraction) with the HBM fraction beside
    it; without it, the HBM roof."""
    bench = load_bench()
    per = bench.phase_bytes(1 << 15, 59)
    ms = {"conversion": 0.6, "tile_2^13": 1.2, "crt": 0.4}
    gbs = {k: per["tile" if k.startswith("tile") else k] * 128 / (v * 1e-3) / 1e9 for k, v in ms.items()}
    key = lambda name: "tile" if name.startswith("tile") else name  # noqa: E731
    instr = 600_000_000
    frac = instr * 4 / (bench.VALU_SIMDS * bench.VALU_CLOCK_HZ * 1.2e-3)
    rec = bench.roofline_record("tile_2^13", ms, gbs, per, key, 128, "phased", 3.0e9, instr, frac,
                                bench.algorithmic_bytes(1 << 15, 59), 2.5)
    assert rec["bound"] == "valu" and abs(rec["frac"] - round(frac, 4)) < 1e-9
    assert abs(rec["achieved"] / rec["peak"] - frac) < 1e-3  # lane-ops/s over the peak issue rate
    assert abs(rec["hbm_frac"] - gbs["tile_2^13"] / bench.HBM_PEAK_GBS) < 1e-4
    rec = bench.roofline_record("tile_2^13", ms, gbs, per, key, 128, "phased", None, None, None,
                                bench.algorithmic_bytes(1 << 15, 59), 2.5)
    assert rec["bound"] == "hbm" and rec["unit"] == "GB/s" and rec["frac"] == rec["hbm_frac"]
