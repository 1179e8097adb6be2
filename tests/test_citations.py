"""Every reference citation (`file:line`, `file:a-b`) in the sources, the
boundary header, the oracle, the tests and the docs resolves to an existing
reference file, lies inside it, and contains the symbol the comment names
(tests/citations.py).  Needs the reference's sources; skipped where they are
absent."""
from __future__ import annotations

import os

import pytest

import citations


@pytest.mark.skipif(not os.path.isdir(citations.REF), reason="the reference's sources are not present")
def test_reference_citations_resolve():
    res = citations.scan()
    bad = [f"{r['src']}:{r['line']}: {r['cite']}: {'; '.join(r['problems'])}" for r in res if r["problems"]]
    assert len(res) > 200
    assert not bad, "\n".join(bad)
