"""Writes reference_exports.txt: the names of the entry points the reference
exports from its flat C-ABI (every non-commented `__stdcall` declaration of
src/JMultiSliceLib.h:111-380).  A list of interface names, used by
tests/test_product_host.py to check that the library resolves each of them.
Run in a container where the reference tree is mounted:
    python tests/golden/make_reference_exports.py /root/reference/src/JMultiSliceLib.h
"""
import os
import re
import sys


def main(path: str) -> None:
    names = []
    for line in open(path, encoding="latin-1"):
        s = line.strip()
        if s.startswith("//") or "__stdcall" not in s:
            continue
        m = re.search(r"__stdcall\s+(\w+)\s*\(", s)
        if m:
            names.append(m.group(1))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_exports.txt")
    with open(out, "w") as f:
        f.write("\n".join(names) + "\n")
    print(f"{len(names)} names -> {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/JMultiSliceLib.h")
