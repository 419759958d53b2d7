"""Writes tests/golden/reference_files.json: the line count of every text
source file of the torchmorph reference checkout (path relative to its root).

The citation test (tests/test_citations.py) checks every `file:N-M`
reference citation in this repository against these counts, so it runs
without the reference checkout present.  Run from the repository root:

    python tests/golden/make_refindex.py /root/reference
"""
import json
import os
import sys

EXTS = (".py", ".cu", ".cpp", ".h", ".txt", ".md", ".toml", ".cfg")


def main(root: str) -> None:
    out = {}
    for dirpath, dirnames, filenames in os.walk(root):
        dirnames[:] = sorted(d for d in dirnames if not d.startswith("."))
        for name in sorted(filenames):
            if name.endswith(EXTS):
                path = os.path.join(dirpath, name)
                with open(path, "rb") as fh:
                    data = fh.read()
                out[os.path.relpath(path, root)] = (
                    data.count(b"\n") + (1 if data and not data.endswith(b"\n") else 0))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_files.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
        fh.write("\n")
    print(f"wrote {len(out)} entries to {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
