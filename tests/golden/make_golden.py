"""Writes the golden fixtures of this directory from the oracle (run from
the repository root: python tests/golden/make_golden.py).

c1_pyramid.shpd      dumps (shapes_amd.wire format) of the 55-box pyramid
                     after steps 1, 2, 10, 100 and 300, and of a contact
                     probe of the world at step 300
c2_polygon_pile.json sha256 of the 8192-polygon pile's dump after steps
                     1, 50, 100, ..., 600, with its body, pair and contact
                     counts
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.dirname(HERE)):
    sys.path.insert(0, p)

import oracle  # noqa: E402
import golden_util  # noqa: E402


def main():
    oracle.build()
    with open(os.path.join(HERE, "c1_pyramid.shpd"), "wb") as f:
        f.write(golden_util.c1_oracle_dumps(oracle))
    with open(os.path.join(HERE, "c2_polygon_pile.json"), "w") as f:
        json.dump(golden_util.c2_oracle_digests(oracle), f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
