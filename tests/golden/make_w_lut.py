"""Extract a subset of the reference's correctly rounded root-of-unity tables as a fixture.

Reads src/base/root_of_unity.hpp of the reference tree (hex-float data of the tables
w_lut_N[i] = {cos(2 pi i / N), sin(2 pi i / N)}, i = 0..N/2) and writes w_lut.json with the
values for the N listed below. Run in a container that has the reference tree:
    python tests/golden/make_w_lut.py /root/reference
"""
import json
import os
import re
import sys

NS = [3, 5, 6, 7, 8, 9, 10, 12, 15, 16, 24, 25, 27, 32, 60, 64]


def main(ref):
    text = open(os.path.join(ref, "src/base/root_of_unity.hpp")).read()
    out = {}
    for n in NS:
        m = re.search(r"w_lut_%d\s*=\s*\{\s*\{(.*?)\}\};" % n, text, re.S)
        if not m:
            raise SystemExit(f"table w_lut_{n} not found")
        pairs = re.findall(r"\{\s*([-+0-9a-fx.p]+)\s*,\s*([-+0-9a-fx.p]+)\s*\}", m.group(1))
        out[str(n)] = [[float.fromhex(a), float.fromhex(b)] for a, b in pairs]
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "w_lut.json"), "w") as f:
        json.dump({"source": "src/base/root_of_unity.hpp (w_lut_N)",
                   "values": {k: [[v[0].hex(), v[1].hex()] for v in vals]
                              for k, vals in out.items()}}, f, indent=0)
    print({k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
