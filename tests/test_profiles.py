"""The committed HBM traffic summaries that bench.py reports as `roofline.traffic`
(CPU only): each must hold a plausible copy calibration and a per-step byte
count between the algorithmic bytes and the line traffic of one random read
per edge, and the file names bench.py looks up must exist."""
import json
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILES = os.path.join(ROOT, "profiles")


def traffic_files():
    return sorted(f for f in os.listdir(PROFILES) if f.endswith("_traffic.json"))


def test_bench_traffic_files_exist():
    src = open(os.path.join(ROOT, "bench.py")).read()
    named = set(re.findall(r'"(k_[A-Za-z0-9_]+_traffic\.json)"', src))
    assert named, "bench.py names no traffic summary"
    for name in named:
        assert os.path.exists(os.path.join(PROFILES, name)), name


def test_traffic_summaries_plausible():
    files = traffic_files()
    assert files
    for name in files:
        d = json.load(open(os.path.join(PROFILES, name)))
        # FETCH_SIZE counts a 256 MiB copy's reads at about half: the factor is ~2
        assert 1.5 < d["fetch_factor"] < 2.5, (name, d["fetch_factor"])
        assert d.get("write_factor", 1.0) == 1.0, name
        alg = d["algorithmic_bytes_per_launch"]
        got = d["bytes_per_launch"]
        # at least the compulsory bytes; at most a 128-byte line per 8-byte
        # gather (16x) plus the streams
        assert alg <= got <= 20 * alg, (name, got, alg)
        for k, v in d.get("per_kernel", {}).items():
            assert 0 < v["bytes_per_dispatch"] <= got, (name, k)
        assert d.get("commands"), name
        for c in d["commands"]:
            assert "bench.py" in c and "rocprofv3 --pmc" in c, (name, c)
