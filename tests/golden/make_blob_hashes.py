"""Regenerates tests/golden/blob_hashes.json (SHA-256 of the flattened blobs of two fixed scenes).

Run from the repository root: python -m tests.golden.make_blob_hashes
These hashes guard the builder restatement against unintended changes; they come from this
repository's own builder, not from the reference (whose builder cannot be compiled here).
"""
import json
import os

from tests.test_builder import _hashes

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "blob_hashes.json")
    with open(out, "w") as fh:
        json.dump(_hashes(), fh, indent=1, sort_keys=True)
    print("wrote", out)
