"""Writes the reference's committed golden DATA as trimmed JSON fixtures.

Run in a container where the reference checkout is mounted (default
/root/reference). Copies only data: inputs and expected outputs from
test/golden.json, oracle/fixtures/*.json and oracle/golden/*.json. Fields the
tests here do not consume (the 3V x 3V expansion A3 = I3 (x) A, the stage-2
Barnes-Hut/multigrid payloads, prose "conventions") are dropped to keep the
fixtures small; every retained number is the reference's value, untouched.

Usage: python tests/golden/extract_reference_goldens.py [/root/reference]
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

STAGE1 = ["crossing", "junction-y", "helix", "linked-rings", "knot"]
SETS = [
    "crossing-length", "crossing-length-frozen", "linked-rings-length", "linked-rings-length-frozen",
    "knot-length-frozen", "crossing-point", "crossing-point-frozen", "linked-rings-point",
    "crossing-edgelengths", "crossing-edgelengths-frozen", "junction-y-edgelengths",
    "junction-y-edgelengths-frozen", "crossing-bary-pen-length", "helix-bary-pen-field",
    "junction-y-bary-pen-diff", "crossing-length-pen-combo",
]
DROP = {"A3", "conventions", "orders"}


def main(ref: str) -> None:
    out = {"energy_gradient": json.load(open(os.path.join(ref, "test", "golden.json")))}
    fixtures = {}
    for name in STAGE1:
        fixtures[name] = json.load(open(os.path.join(ref, "oracle", "fixtures", name + ".json")))
    out["fixtures"] = fixtures
    goldens = {}
    for name in STAGE1 + SETS:
        g = json.load(open(os.path.join(ref, "oracle", "golden", name + ".json")))
        conv = g.get("conventions", {})
        kept = {k: v for k, v in g.items() if k not in DROP}
        kept["projection_mode"] = conv.get("projection", "reassemble")
        goldens[name] = kept
    out["goldens"] = goldens
    path = os.path.join(HERE, "reference_goldens.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
