"""Writes tests/golden/c0_digests.json: SHA-256 digests of the seeded C0 scene (10k soup
triangles, 512x512, seed 1000) and of the oracle's outputs on it. Regenerate only when the
scene generator or the oracle changes on purpose."""

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
from lucid_amd import scenes  # noqa: E402


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    sc = scenes.config0()
    res = oracle.render(sc)
    out = {
        "scene": sc.digest(),
        "frag_counts": digest(res["frag_counts"]),
        "info_counts": digest(res["info"][:oracle.INFO_U32 + 6 * res["bin_count"]]),
        "num_fragments": int(res["num_fragments"]),
        "image": digest(res["image"]),
    }
    with open(os.path.join(HERE, "c0_digests.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
