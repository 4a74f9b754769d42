"""Writes rust_sci.json: numbers the reference printed with Rust's
`{:.17e}` (18 significant digits) and `{:.9e}` (10), collected from the
result tables and logs the reference publishes under docs/, evidence/ and
validation/ (its examples write them with these format specifiers).  Only
strings in Rust's exponent form are kept (no '+', no leading zero in the
exponent), so a C or Python "%e" output cannot slip in.  The test re-formats
each parsed value and must get the string back (tests/test_dense3d.py).

Run in the build container (the reference tree is at /root/reference):
    python tests/golden/make_rust_sci.py
"""
import json
import os
import re

REF = "/root/reference"
ROOTS = ["docs", "evidence", "validation"]
PAT = {18: re.compile(r"(?<![0-9.+\-eE])-?[0-9]\.[0-9]{17}e(?:0|-?[1-9][0-9]{0,2})(?![0-9.])"),
       10: re.compile(r"(?<![0-9.+\-eE])-?[0-9]\.[0-9]{9}e(?:0|-?[1-9][0-9]{0,2})(?![0-9.])")}


def main():
    found = {18: set(), 10: set()}
    files = {18: set(), 10: set()}
    for root in ROOTS:
        for dirpath, _, names in os.walk(os.path.join(REF, root)):
            for name in names:
                path = os.path.join(dirpath, name)
                try:
                    text = open(path, encoding="utf-8", errors="ignore").read()
                except OSError:
                    continue
                for digits, pat in PAT.items():
                    hits = pat.findall(text)
                    if hits:
                        found[digits].update(hits)
                        files[digits].add(os.path.relpath(path, REF))
    out = {
        "note": "strings printed by the reference's Rust code with {:.17e} / {:.9e}",
        "sci18": sorted(found[18]),
        "sci10": sorted(found[10]),
        "files18": len(files[18]),
        "files10": sorted(files[10]),
    }
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "rust_sci.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(len(out["sci18"]), "x 18 digits,", len(out["sci10"]), "x 10 digits")


if __name__ == "__main__":
    main()
