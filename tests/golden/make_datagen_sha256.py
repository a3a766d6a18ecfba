"""Writes tests/golden/datagen_sha256.json: the SHA-256 of every seeded
workload's COO (rows u64, items u32, timestamps i64 when present, little
endian), so a change to the generators or to numpy's streams shows up as a
test failure instead of silently different inputs.

    python tests/golden/make_datagen_sha256.py
"""

import hashlib
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from cupami_amd import datagen  # noqa: E402


def combined(dbs):
    h = hashlib.sha256()
    for db in dbs:
        h.update(bytes.fromhex(db.sha256()))
    return h.hexdigest()


def digests():
    batch = datagen.batch_256()
    out = {
        "small_uniform_101": datagen.small_uniform(101).sha256(),
        "eclat_1m_1k_202": datagen.eclat_1m_1k().sha256(),
        "temporal_2m_500_404": datagen.temporal_2m_500().sha256(),
        "batch_256_303": combined(db for db, _ in batch),
        "batch_256_303_minsups": hashlib.sha256(json.dumps([ms for _, ms in batch]).encode()).hexdigest(),
        "serve_pool_48_505": combined(datagen.build_pool(datagen.serve_pool(48))),
        "serve_stream_10000_1000qps_505": hashlib.sha256(
            json.dumps(datagen.serve_stream(10_000, 1000.0, 48)).encode()).hexdigest(),
    }
    return out


if __name__ == "__main__":
    d = digests()
    (HERE / "datagen_sha256.json").write_text(json.dumps(d, indent=1) + "\n")
    print(json.dumps(d, indent=1))
