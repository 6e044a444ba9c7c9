"""Transcribe the ELKI OPTICS reference vectors held by fuel's own test
(src/cluster/optics.rs, optics_matches_elki_reference_ordering_and_reachability:
60 points in 2-d, eps = inf, min_points = 5) into a JSON fixture.  Reads the
reference test source as text; nothing of the reference is run."""
import json
import os
import re
import sys

SRC = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", "src/cluster/optics.rs")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "optics_elki_60.json")


def main():
    text = open(SRC).read()
    start = text.index("fn optics_matches_elki_reference_ordering_and_reachability")
    body = text[start:]
    pts_block = body[body.index("let points = vec!["):body.index("let expected_ordering")]
    points = [[float(a), float(b)] for a, b in re.findall(r"vec!\[(-?[0-9.]+), (-?[0-9.]+)\]", pts_block)]
    ord_block = body[body.index("let expected_ordering"):body.index("let expected_reachability_by_order")]
    ordering = [int(v) for v in re.findall(r"\b(\d+)\b", ord_block[ord_block.index("["):])]
    r_block = body[body.index("let expected_reachability_by_order"):body.index("let data =")]
    reach = []
    for tok in re.findall(r"(f64::INFINITY|-?[0-9]+\.[0-9]+)", r_block):
        reach.append(float("inf") if tok == "f64::INFINITY" else float(tok))
    assert len(points) == 60 and len(ordering) == 60 and len(reach) == 60
    with open(OUT, "w") as f:
        json.dump({"source": "src/cluster/optics.rs optics_matches_elki_reference_ordering_and_reachability",
                   "eps": "inf", "min_points": 5, "points": points, "expected_ordering": ordering,
                   "expected_reachability_by_order": [None if r == float("inf") else r for r in reach]}, f,
                  indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
