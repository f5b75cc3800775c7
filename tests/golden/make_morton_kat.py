"""Write morton_lut_kat.json: known-answer vectors sampled from the reference's
Morton lookup tables (src/libs/libmorton/include/morton3D_LUTs.h:7-224).

Run in a container that has the reference at /root/reference; the output is a small
table of (index, value) pairs per table, committed as a fixture.
"""
import json
import os
import re

SRC = "/root/reference/src/libs/libmorton/include/morton3D_LUTs.h"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "morton_lut_kat.json")
TABLES = ["Morton3D_encode_x_256", "Morton3D_encode_y_256", "Morton3D_encode_z_256",
          "Morton3D_decode_x_512", "Morton3D_decode_y_512", "Morton3D_decode_z_512"]


def main():
    text = open(SRC).read()
    out = {"source": "src/libs/libmorton/include/morton3D_LUTs.h", "tables": {}}
    for name in TABLES:
        m = re.search(name + r"\[\d+\]\s*=\s*\{(.*?)\}", text, re.S)
        vals = [int(v, 0) for v in re.findall(r"0x[0-9A-Fa-f]+|\d+", m.group(1))]
        idx = sorted(set(list(range(0, len(vals), 7)) + [1, 2, 3, 255, len(vals) - 1]))
        out["tables"][name] = {"size": len(vals), "samples": [[i, vals[i]] for i in idx]}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
