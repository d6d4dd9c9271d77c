"""Write tdm_h6.json: the transition-density slices the reference's
test/test_TDMs.jl:55-170 asserts (|Γ| to 1e-5 for h6 with clusters
(1:4), (5:6), complete bases), as data: cluster, operator, (bra, ket)
sectors, the fixed leading indices and the expected matrix.  Reads the
reference test file as text:

    python3 tests/golden/extract_tdm_values.py /root/reference/test/test_TDMs.jl
"""
import json
import os
import re
import sys


def extract(text):
    out = []
    # each assertion: a matrix literal assigned to ref/ref1/ref2, then the
    # slice it is compared with
    pat = re.compile(r"ref\d?\s*=\s*\[(?P<body>[^\]]*)\]"
                     r"(?P<gap>(?:(?!ref\d?\s*=).)*?)"
                     r"test\s*=\s*cluster_ops\[(?P<c>\d+)\]\[\"(?P<op>\w+)\"\]"
                     r"\[\(\((?P<b0>\d+),\s*(?P<b1>\d+)\),\s*\((?P<k0>\d+),\s*(?P<k1>\d+)\)\)\]"
                     r"\[(?P<idx>[^\]]*)\]", re.S)
    for m in pat.finditer(text):
        rows = [[float(x) for x in r.split()] for r in m.group("body").strip().split("\n")
                if r.strip()]
        lead = [int(x) for x in m.group("idx").split(",") if x.strip() not in (":", "")]
        out.append({"cluster": int(m.group("c")) - 1, "op": m.group("op"),
                    "bra": [int(m.group("b0")), int(m.group("b1"))],
                    "ket": [int(m.group("k0")), int(m.group("k1"))],
                    "lead": [i - 1 for i in lead], "ref": rows})
    return out


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test/test_TDMs.jl"
    recs = extract(open(src).read())
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "tdm_h6.json"), "w") as f:
        json.dump(recs, f, indent=1)
    for r in recs:
        print(r["cluster"], r["op"], r["bra"], r["ket"], r["lead"], len(r["ref"]), len(r["ref"][0]))


if __name__ == "__main__":
    main()
