"""Every reference citation (``file.py:a-b``, ``file.h:a``) in the product,
the oracle, the tests and the design notes points inside the cited file of
the reference checkout, and where the text just before a citation names a
function, kernel or class that the cited file defines, the cited range
holds its first line or lies within it (a C definition ends at its closing
brace, a Python one at the next definition at the same or a lower indent). Skipped where the reference checkout is absent."""
import collections
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
SCAN = ["include", "oracle", "svirl_amd", "tests", "DESIGN.md", "INTEGRATION.md", "README.md", "bench.py",
        "__graft_entry__.py"]
CITE = re.compile(r"((?:[\w.-]+/)*[\w-]+\.(?:py|h)):(\d+)(?:-(\d+))?")
DEF = re.compile(r"^\s*(?:def|class)\s+(\w+)|^\s*(?:__global__\s+)?(?:void|int32_t|uint32_t|real_t|complex_t)\s+(\w+)\s*\(")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "svirl")), reason="reference checkout absent")


def _reference():
    files, defs = {}, {}
    for root, _, names in os.walk(REF):
        if "/.git" in root or "/docs" in root:
            continue
        for f in names:
            if not f.endswith((".py", ".h")):
                continue
            p = os.path.join(root, f)
            lines = open(p, errors="replace").read().split("\n")
            rel = os.path.relpath(p, REF)
            files[rel] = len(lines)
            starts = []
            for k, line in enumerate(lines, 1):
                m = DEF.match(line)
                if m:
                    starts.append((k, len(line) - len(line.lstrip()), m.group(1) or m.group(2)))
            # a C definition extends to its closing brace in column 0, a
            # Python one to the next definition at the same or a lower indent
            d = collections.defaultdict(list)
            for idx, (k, ind, name) in enumerate(starts):
                if p.endswith(".h"):
                    end = next((k2 for k2 in range(k, len(lines) + 1) if lines[k2 - 1].startswith("}")), len(lines))
                else:
                    end = next((k2 - 1 for k2, ind2, _ in starts[idx + 1:] if ind2 <= ind), len(lines))
                d[name].append((k, end))
            defs[rel] = d
    return files, defs


def _resolve(name, files):
    """Reference files a citation can mean: the path itself, or every file
    with that base name for a bare file name."""
    if "/" in name:
        return [r for r in files if r == name or r.endswith("/" + name)]
    return [r for r in files if os.path.basename(r) == name]


def _citations():
    for entry in SCAN:
        path = os.path.join(ROOT, entry)
        walk = os.walk(path) if os.path.isdir(path) else [(ROOT, None, [entry])]
        for root, _, names in walk:
            if "__pycache__" in root or "_lib" in root:
                continue
            for f in names:
                if not f.endswith((".py", ".h", ".hpp", ".hip", ".md")) or f == "test_citations.py":
                    continue
                p = os.path.join(root, f)
                for ln, line in enumerate(open(p, errors="replace"), 1):
                    for m in CITE.finditer(line):
                        yield os.path.relpath(p, ROOT), ln, line, m


def test_citations_point_inside_the_reference():
    files, defs = _reference()
    out_of_range, misplaced, n = [], [], 0
    for where, ln, line, m in _citations():
        cands = _resolve(m.group(1), files)
        if not cands:
            continue  # a file of this repository, not of the reference
        a = int(m.group(2))
        b = int(m.group(3) or a)
        n += 1
        if a > b or b > max(files[c] for c in cands):
            out_of_range.append(f"{where}:{ln}: {m.group(0)}")
            continue
        # names the cited file defines, mentioned just before the citation
        window = line[max(0, m.start() - 90):m.start()]
        for c in cands:
            named = [w for w in set(re.findall(r"\w+", window)) if w in defs[c] and len(w) > 6]
            if len(named) == 1 and not any(a <= k <= b or k <= a <= b <= e for k, e in defs[c][named[0]]):
                misplaced.append(f"{where}:{ln}: {named[0]} is defined at {c}:{defs[c][named[0]]}, "
                                 f"cited as {m.group(0)}")
    assert n > 200
    assert not out_of_range, out_of_range
    assert not misplaced, misplaced
