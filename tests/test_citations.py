"""Every `file:N` / `file:N-M` citation of a reference file in this
repository's code and documents points inside that file.

Line counts come from tests/golden/reference_files.json (written by
tests/golden/make_refindex.py from the reference checkout), so the test
needs no checkout; where the checkout is present, the index is checked
against it too.
"""
import json
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INDEX = os.path.join(ROOT, "tests", "golden", "reference_files.json")
CITE = re.compile(r"([A-Za-z0-9_./-]+\.(?:py|cu|cpp|h|txt|toml|cfg|md)):(\d+)(?:-(\d+))?")
# files of third-party projects cited for behaviour the reference relies on
EXTERNAL = {"_morphology.py"}
SCAN_DIRS = ("include", "torchmorph_amd", "oracle", "tests")
SCAN_FILES = ("DESIGN.md", "INTEGRATION.md", "README.md", "bench.py", "__graft_entry__.py")
SCAN_EXTS = (".py", ".h", ".hip", ".c", ".md")


def _index():
    with open(INDEX) as fh:
        return json.load(fh)


def _sources():
    for name in SCAN_FILES:
        path = os.path.join(ROOT, name)
        if os.path.exists(path):
            yield path
    for d in SCAN_DIRS:
        for dirpath, dirnames, filenames in os.walk(os.path.join(ROOT, d)):
            dirnames[:] = [x for x in dirnames if x not in ("_build", "__pycache__", "_ref")]
            for name in filenames:
                if name.endswith(SCAN_EXTS) or name == "Makefile":
                    yield os.path.join(dirpath, name)


def _resolve(cited, index):
    """The reference file a cited path names (suffix match on whole path
    components; a bare name shared by benchmark/ and the package means the
    package file), or None for a file the reference does not hold."""
    cited = cited.lstrip("./")
    hits = [p for p in index if p == cited or p.endswith("/" + cited)]
    if len(hits) > 1:
        hits = [p for p in hits if p.startswith("torchmorph/")] or hits
    if len(hits) > 1:
        hits = [p for p in hits if p.count("/") == min(h.count("/") for h in hits)]
    return hits


def test_index_matches_checkout():
    ref = "/root/reference"
    if not os.path.isdir(ref):
        pytest.skip("reference checkout not present")
    index = _index()
    for rel, n in index.items():
        with open(os.path.join(ref, rel), "rb") as fh:
            data = fh.read()
        assert data.count(b"\n") + (1 if data and not data.endswith(b"\n") else 0) == n, rel


def test_reference_citations_in_range():
    index = _index()
    bad, seen = [], 0
    for path in _sources():
        if os.path.abspath(path) == os.path.abspath(__file__):
            continue
        with open(path, encoding="utf-8", errors="replace") as fh:
            for lineno, line in enumerate(fh, 1):
                for m in CITE.finditer(line):
                    cited, lo = m.group(1), int(m.group(2))
                    hi = int(m.group(3)) if m.group(3) else lo
                    base = cited.rsplit("/", 1)[-1]
                    if base in EXTERNAL or cited.startswith(("torchmorph_amd/", "tests/",
                                                             "oracle/", "include/")):
                        continue
                    hits = _resolve(cited, index)
                    if not hits:
                        continue  # not a reference file (this repository's own)
                    seen += 1
                    where = f"{os.path.relpath(path, ROOT)}:{lineno}: {m.group(0)}"
                    if len(hits) != 1:
                        bad.append(f"{where} is ambiguous ({hits})")
                    elif not (1 <= lo <= hi <= index[hits[0]]):
                        bad.append(f"{where} outside {hits[0]} ({index[hits[0]]} lines)")
    assert seen > 50, f"only {seen} citations found: the scanner is broken"
    assert not bad, "\n".join(bad)
