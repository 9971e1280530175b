"""The benchmark workloads regenerate bit for bit: SHA-256 of their FP64
bytes against the fixtures written by tests/golden/make_workload_sha256.py."""
import json
import os

from tests.golden.make_workload_sha256 import digests

HERE = os.path.dirname(os.path.abspath(__file__))


def test_workloads_match_committed_digests():
    with open(os.path.join(HERE, "golden", "workload_sha256.json")) as f:
        want = json.load(f)
    assert digests() == want
