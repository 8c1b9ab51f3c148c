"""Extract the H, C and O entries of four Psi4-format (.gbs) basis sets.

The basis-set data (Basis Set Exchange, public) are read from the hfscf
checkout (``basis/<name>.gbs``) and written to the package's basis folder
(``hfscf_amd/basis``) with the
header line (``spherical``/``cartesian``) and the element blocks kept exactly
as published, so shell order, SP blocks and contraction coefficients are the
ones hfscf itself reads (hfscf: src/basis/hfscf_basis.cpp:19-242).

Usage: python tests/golden/extract_basis.py /path/to/hfscf
"""
import os
import sys

NAMES = ["sto-3g", "cc-pvdz", "6-31g_st_", "6-31g_st__st_"]
KEEP = {"H", "C", "O"}


def extract(src, dst):
    with open(src) as f:
        lines = f.read().splitlines()
    out = [lines[0].strip(), "", "****"]
    i = 1
    while i < len(lines):
        toks = lines[i].split()
        if len(toks) == 2 and toks[0] in KEEP and toks[1] == "0":
            block = [lines[i]]
            i += 1
            while i < len(lines) and not lines[i].startswith("****"):
                block.append(lines[i])
                i += 1
            out.extend(block)
            out.append("****")
        i += 1
    with open(dst, "w") as f:
        f.write("\n".join(out) + "\n")


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    dest = os.path.join(os.path.dirname(os.path.dirname(here)), "hfscf_amd", "basis")
    for n in NAMES:
        extract(os.path.join(root, "basis", n + ".gbs"), os.path.join(dest, n + ".gbs"))


if __name__ == "__main__":
    main()
