"""The benchmark's HBM traffic comes from the committed counter profile
(profiles/pmc_traffic_colonnade_trace_and_gi.md): every block parses, and the
launch configurations a new context starts with (as include/lux.h states
them) have a block of their own, so a default run reports measured traffic."""
import re
from pathlib import Path

import bench

ROOT = Path(__file__).resolve().parents[1]
FIELDS = ("waves_per_cu", "lds_stack_depth", "regroup_threshold", "steps_per_refill")


def header_default(pattern: str) -> dict:
    text = (ROOT / "include" / "lux.h").read_text()
    m = re.search(pattern + r"\s*\n?\s*\*?\s*\{(\d+), (\d+), (\d+), (\d+)\}", text)
    assert m, pattern
    return dict(zip(FIELDS, (int(v) for v in m.groups())))


def test_every_traffic_block_parses():
    blocks, traffic = 0, 0
    for line in bench.TRAFFIC_PROFILE.read_text().splitlines():
        parts = line.split()
        if parts and parts[0] in ("config", "gi_config"):
            kv = dict(p.split("=") for p in parts[1:])
            assert set(FIELDS) <= set(kv), line
            blocks += 1
        elif parts and parts[0] == "traffic":
            assert len(parts) == 4, line
            kv = dict(p.split("=") for p in parts[2:])
            assert int(kv["fetch_bytes"]) > 0 and int(kv["write_bytes"]) > 0, line
            traffic += 1
    assert blocks >= 2 and traffic >= 8


def test_default_configurations_have_measured_traffic():
    trace = header_default(r"A new context starts at")
    gi = header_default(r"a new context starts at")
    for kernel, cfg, is_gi in (("k_trace_regroup<%d>" % trace["lds_stack_depth"], trace, False),
                               ("k_gi_primary<%d>" % gi["lds_stack_depth"], gi, True)):
        want = dict(cfg, blas="sah")
        value, measured = bench.measured_traffic(kernel, want, gi=is_gi)
        assert value is not None and value > 0, (kernel, want)
        assert measured == want, (kernel, want, measured)


def test_traffic_falls_back_to_the_same_kernel():
    cfg = {"waves_per_cu": 17, "lds_stack_depth": 12, "regroup_threshold": 3, "steps_per_refill": 5, "blas": "sah"}
    value, measured = bench.measured_traffic("k_trace_regroup<12>", cfg)
    assert value is not None and measured["lds_stack_depth"] == 12 and measured["blas"] == "sah"
    assert bench.measured_traffic("k_trace_regroup<5>", cfg) == (None, None)
