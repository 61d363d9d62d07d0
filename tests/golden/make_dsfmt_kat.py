"""Extract the dSFMT-19937 known-answer vectors that the reference ships
(dSFMT/dSFMT.19937.out.txt) into tests/golden/dsfmt_19937_kat.json.

Run in a container that has the reference checkout:
    python tests/golden/make_dsfmt_kat.py /root/reference
The output is data only: seed, range and the printed values.
"""
import json
import os
import re
import sys

RANGES = {"[1, 2)": "close1_open2", "[0, 1)": "close_open", "(0, 1]": "open_close",
          "(0, 1)": "open_open"}


def main(ref_root: str) -> None:
    path = os.path.join(ref_root, "dSFMT", "dSFMT.19937.out.txt")
    blocks, cur = [], None
    with open(path) as f:
        for line in f:
            m = re.match(r"init_gen_rand\((\d+)\) generated randoms (.*)$", line.strip())
            if m:
                cur = {"seed": int(m.group(1)), "kind": RANGES[m.group(2)], "values": []}
                blocks.append(cur)
            elif line.startswith("init_"):
                cur = None  # init_by_array streams: not used by the nbody client
            elif cur is not None and line.strip():
                cur["values"].extend(float(x) for x in line.split())
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dsfmt_19937_kat.json")
    with open(out, "w") as f:
        json.dump({"source": "dSFMT/dSFMT.19937.out.txt", "blocks": blocks}, f)
    print(out, len(blocks), sum(len(b["values"]) for b in blocks))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
