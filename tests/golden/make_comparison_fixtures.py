"""Writes comparison_fixtures.json: the fixtures of the reference's external benchmark harness
(scripts/benchmark/comparison/fixtures.json) that list their operations explicitly (the ideal and noise
categories).  The scalability fixtures are generated by the harness with CPython's random module and are
left out.  Run from the repository root in a checkout next to the reference tree:

    python tests/golden/make_comparison_fixtures.py /root/reference/scripts/benchmark/comparison/fixtures.json
"""
import json
import sys
from pathlib import Path

src = Path(sys.argv[1])
data = json.loads(src.read_text(encoding="utf-8"))
keep = [f for f in data["fixtures"] if f.get("operations")]
out = {"schema_version": data["schema_version"], "defaults": data["defaults"], "fixtures": keep}
Path(__file__).with_name("comparison_fixtures.json").write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
print(f"{len(keep)} fixtures")
