"""Writes reference_layout.json: the line count of every source file of the
reference's time-step path (Main_Code/src) and of its remesh tool
(tools/Re_Mesh/src), and the first and last line of every function defined in
them (brace-matched).  Only names and line numbers are recorded, no source
text.  tests/test_citations.py checks the repository's file:line citations
against this table, so it runs without the reference present.

Run from the repository root in a checkout next to the reference:
    python tests/golden/make_reference_layout.py [/root/reference]
"""
import json
import os
import re
import sys

DIRS = {"Main_Code/src": "", "tools/Re_Mesh/src": "Re_Mesh/"}
DEF = re.compile(r'^\s*(?:static\s+|__global__\s+|__device__\s+|inline\s+)*[\w<>:\*]+[\s\*]+\**(\w+)\s*\(')


def functions(lines):
    out = []
    i, n = 0, len(lines)
    while i < n:
        line = lines[i]
        m = DEF.match(line)
        s = line.strip()
        if m and not s.startswith(("//", "#", "/*")) and not s.endswith(";"):
            depth, started, end, j = 0, False, None, i
            while j < n:
                for ch in lines[j].split("//")[0]:
                    if ch == "{":
                        depth += 1
                        started = True
                    elif ch == "}":
                        depth -= 1
                if started and depth == 0:
                    end = j
                    break
                if not started and lines[j].rstrip().endswith(";"):
                    break
                j += 1
            if end is not None:
                out.append([m.group(1), i + 1, end + 1])
                i = end + 1
                continue
        i += 1
    return out


def main(root):
    layout = {}
    for d, prefix in DIRS.items():
        for f in sorted(os.listdir(os.path.join(root, d))):
            text = open(os.path.join(root, d, f), errors="replace").read()
            lines = text.split("\n")
            if lines and lines[-1] == "":
                lines.pop()
            layout[prefix + f] = {"lines": len(lines), "functions": functions(lines)}
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "reference_layout.json"), "w") as fh:
        json.dump(layout, fh, indent=1, sort_keys=True)
        fh.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
