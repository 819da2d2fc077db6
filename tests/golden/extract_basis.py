"""Data recipe (not shipped code): writes the GENBAS-format basis-set files
aquarius_amd/basis/{cc-pVDZ,DZ}.gbs, the entries for H, C, N and O of cc-pVDZ
and DZ (Dunning), taken verbatim from the EMSL-derived library files that
Aquarius distributes under basis/.  Run once, in a checkout that has the
reference tree at /root/reference:

    python tests/golden/extract_basis.py
"""

import os
import sys

SRC = "/root/reference/basis"
KEEP = ("H", "C", "N", "O")
SETS = ("cc-pVDZ", "DZ")


def entries(text):
    """Split a GENBAS file into (header comment lines, {element: entry text})."""
    lines = text.splitlines()
    head, out, cur, name = [], {}, None, None
    for line in lines:
        if line[:1] == "!" and cur is None:
            head.append(line)
            continue
        sep = line.find(":")
        if 1 <= sep <= 2 and line[:sep].strip().isalpha() and not line[0].isspace():
            name = line[:sep].strip().upper()
            cur = out.setdefault(name, [])
        if cur is not None:
            cur.append(line)
    return head, {k: "\n".join(v).rstrip() + "\n" for k, v in out.items()}


def main():
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "aquarius_amd", "basis")
    for s in SETS:
        with open(os.path.join(SRC, s)) as f:
            head, ent = entries(f.read())
        missing = [e for e in KEEP if e not in ent]
        if missing:
            sys.exit(f"{s}: no entry for {missing}")
        with open(os.path.join(here, s + ".gbs"), "w") as f:
            f.write("\n".join(head) + "\n\n")
            for e in KEEP:
                f.write(ent[e] + "\n")


if __name__ == "__main__":
    main()
