"""Citations of the reference (file:line) in the repository's sources and
documents point at lines that exist and, where a reference function is named
next to the citation, at that function.

The reference's file lengths and function line ranges come from
tests/golden/reference_layout.json (written by make_reference_layout.py from
the reference's sources), so this runs without the reference present.  Where
the reference is present, the fixture itself is checked against it.
"""
import json
import os
import re

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LAYOUT = json.load(open(os.path.join(HERE, "golden", "reference_layout.json")))
REF = "/root/reference"

SCAN_DIRS = ["oracle", "include", "nspipe-gpu_amd/csrc", "nspipe-gpu_amd/nsp", "tests"]
SCAN_FILES = ["DESIGN.md", "INTEGRATION.md", "README.md", "bench.py", "__graft_entry__.py"]
EXTS = (".cpp", ".hip", ".h", ".py", ".md")

FILE_RE = re.compile(r"((?:tools/Re_Mesh/src/|Main_Code/src/)?(?:mod_\w+\.cu|head\.h|main\.cu))"
                     r"((?::\d+(?:-\d+)?)(?:,\s*:?\d+(?:-\d+)?)*)")
CONT_RE = re.compile(r"\b(\w+)\s+:(\d+)(?:-(\d+))?")
GAP_RE = re.compile(r"^[\s(),;`*]*$")


def sources():
    paths = [os.path.join(ROOT, f) for f in SCAN_FILES]
    for d in SCAN_DIRS:
        for dp, _, fs in os.walk(os.path.join(ROOT, d)):
            if "__pycache__" in dp or "golden" in dp:
                continue
            paths += [os.path.join(dp, f) for f in fs if f.endswith(EXTS)]
    return sorted(p for p in paths if os.path.isfile(p) and not p.endswith("test_citations.py"))


def layout_key(name, line):
    base = name.split("/")[-1]
    remesh = "Re_Mesh" in name or base in ("mod_remesh.cu", "mod_io.cu") or \
        (base in ("main.cu", "head.h") and "Re_Mesh" in line)
    return ("Re_Mesh/" + base) if remesh else base


def citations():
    """(source, line number, layout key, first, last, text before the citation)."""
    out = []
    for path in sources():
        last_key = None
        for n, line in enumerate(open(path, errors="replace"), 1):
            spans = []
            for m in FILE_RE.finditer(line):
                key = layout_key(m.group(1), line)
                last_key = key
                for a, b in re.findall(r"(\d+)(?:-(\d+))?", m.group(2)):
                    out.append((path, n, key, int(a), int(b or a), line[:m.start()]))
                spans.append((m.start(), m.end()))
            for m in CONT_RE.finditer(line):
                if last_key is None or any(s <= m.start() < e for s, e in spans):
                    continue
                if m.group(1) in {f[0] for f in LAYOUT[last_key]["functions"]}:
                    out.append((path, n, last_key, int(m.group(2)), int(m.group(3) or m.group(2)),
                                line[:m.start() + len(m.group(1))]))
    return out


CITES = citations()


def test_citations_found():
    assert len(CITES) > 150
    assert {c[2] for c in CITES} >= {"mod_linear.cu", "mod_radialFD.cu", "mod_boundary.cu",
                                     "mod_integrate.cu", "mod_io_nonlinear.cu", "mod_cublasflux.cu",
                                     "Re_Mesh/mod_remesh.cu", "head.h", "main.cu"}


def test_citations_within_files():
    bad = [f"{os.path.relpath(p, ROOT)}:{n} cites {k}:{a}-{b} (file has {LAYOUT[k]['lines']} lines)"
           for p, n, k, a, b, _ in CITES if not (1 <= a <= b <= LAYOUT[k]["lines"])]
    assert not bad, "\n".join(bad)


def named_range(key, before):
    """Line range of the reference function(s) named directly before a
    citation: "name (file:N)", "name, file:N", "name :N", or a run of names
    joined by "/" or "," ("rBufferBinary/wBufferBinary, file:N-M"), whose
    ranges are united.  None when no function name directly precedes it."""
    funcs = {f[0]: (f[1], f[2]) for f in LAYOUT[key]["functions"]}
    rest, lo, hi, names = before, None, None, []
    while True:
        m = re.search(r"(\w+)([\s(),;`*/]*)$", rest)
        if not m or m.group(1) not in funcs or (names and not m.group(2).strip(" `*")):
            break
        if not names and not GAP_RE.match(m.group(2)):
            break
        s, e = funcs[m.group(1)]
        lo = s if lo is None else min(lo, s)
        hi = e if hi is None else max(hi, e)
        names.append(m.group(1))
        rest = rest[:m.start()]
    return (names, lo, hi) if names else None


def test_citations_hit_named_function():
    bad = []
    for p, n, k, a, b, before in CITES:
        hit = named_range(k, before)
        if hit is not None:
            names, s, e = hit
            if a < s or b > e:
                bad.append(f"{os.path.relpath(p, ROOT)}:{n} cites {k}:{a}-{b} next to {'/'.join(names)} ({s}-{e})")
    assert not bad, "\n".join(bad)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
def test_layout_matches_reference():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", os.path.join(HERE, "golden", "make_reference_layout.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    for d, prefix in mk.DIRS.items():
        for f in sorted(os.listdir(os.path.join(REF, d))):
            lines = open(os.path.join(REF, d, f), errors="replace").read().split("\n")
            if lines and lines[-1] == "":
                lines.pop()
            assert LAYOUT[prefix + f]["lines"] == len(lines)
            assert LAYOUT[prefix + f]["functions"] == mk.functions(lines)
