"""Every seeded workload (the 10k x 50, 1M x 1k, 2M x 500 temporal, the 256
batched databases, the served pool and request stream) regenerates to the
SHA-256 committed in tests/golden/datagen_sha256.json."""

import json
from pathlib import Path

import make_datagen_sha256

GOLDEN = Path(__file__).resolve().parent / "golden" / "datagen_sha256.json"


def test_workload_digests():
    assert make_datagen_sha256.digests() == json.loads(GOLDEN.read_text())
