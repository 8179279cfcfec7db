"""Every reference citation `path.jl:a-b` (also .md, .toml, .txt) in this repository's sources, tests and documents
names lines that exist in the cited reference file.

A citation resolves by its path suffix to the reference's files; one with a bare file name that several reference
files share passes when the range fits one of them. Skipped where the reference checkout is absent."""
import os
import re
import subprocess

import pytest

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# inputs of the build written by others (the survey, the reviews, the baseline), not this repository's text
NOT_OURS = {"SURVEY.md", "VERDICT.md", "ADVICE.md", "BASELINE.md", "BASELINE.json"}
CITE = re.compile(r"([A-Za-z0-9_./-]+\.(?:jl|md|toml|txt))((?::\d+(?:-\d+)?(?:,\d+(?:-\d+)?)*)+)")


def _our_files():
    try:
        out = subprocess.run(["git", "ls-files"], cwd=ROOT, capture_output=True, text=True, check=True).stdout.split()
    except (OSError, subprocess.CalledProcessError):
        out = [os.path.relpath(os.path.join(dp, f), ROOT) for dp, _, fs in os.walk(ROOT) for f in fs]
    keep = (".py", ".hip", ".hpp", ".h", ".cpp", ".md", "Makefile")
    return [f for f in out if f.endswith(keep) and os.path.basename(f) not in NOT_OURS and not f.startswith("_work/")]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference checkout absent")
def test_reference_citations_in_range():
    index = {}
    for dp, dn, fn in os.walk(REF):
        dn[:] = [d for d in dn if d != ".git"]
        for f in fn:
            index.setdefault(f, []).append(os.path.relpath(os.path.join(dp, f), REF))
    nlines = {}

    def length(p):
        if p not in nlines:
            with open(os.path.join(REF, p), "rb") as fh:
                nlines[p] = len(fh.read().splitlines())
        return nlines[p]

    bad, checked = [], 0
    for f in _our_files():
        try:
            with open(os.path.join(ROOT, f), encoding="utf-8") as fh:
                text = fh.read()
        except (OSError, UnicodeDecodeError):
            continue
        for ln, line in enumerate(text.splitlines(), 1):
            for m in CITE.finditer(line):
                name, rng = m.group(1), m.group(2)
                base = os.path.basename(name)
                if name.startswith("LOG/"):  # the verify-log directory abbreviation
                    cands = index.get(base, [])
                else:
                    cands = [p for p in index.get(base, []) if p.endswith(name.lstrip("./"))]
                if not cands:
                    continue  # not a reference file (one of this repository's own names)
                top = max(int(x) for x in re.findall(r"\d+", rng))
                checked += 1
                if not any(top <= length(c) for c in cands):
                    bad.append(f"{f}:{ln}: {name}{rng} (files: {[(c, length(c)) for c in cands]})")
    assert checked > 300
    assert not bad, "\n".join(bad)
