"""Self-generated regression pins: seeded inputs and oracle outputs must keep their SHA-256.

The fixture (tests/golden/oracle_sha256.json) was written by tests/golden/make_golden.py; a
mismatch means the generator or the oracle changed.
"""
import json
import os
import sys

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)

import make_golden  # noqa: E402


def test_oracle_outputs_match_committed_digests():
    want = json.load(open(os.path.join(HERE, "oracle_sha256.json")))["cases"]
    got = make_golden.cases()
    assert sorted(got) == sorted(want)
    for name in want:
        assert got[name]["inputs"] == want[name]["inputs"], f"{name}: generator output changed"
        assert got[name]["outputs"] == want[name]["outputs"], f"{name}: oracle output changed"
