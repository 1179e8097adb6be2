"""Scanner for the reference citations (`file:line` and `file:a-b`) in the
repository's sources, resolved against the reference's own files.

A citation is `path.ext:R` where R is a comma-separated list of lines or
ranges; `.ext:R` right after a citation names the same file stem with
another extension (`PerfectHashFunction.h:53, .cu:302-576`), and a bare
`:R` continues the last file named on the same line or in the same comment.

Checks: the file exists, every range lies inside the file, and when the text
just before the citation names a symbol of the cited file (`Foo::bar`,
`doBezierClipping`), that symbol occurs inside one of the cited ranges."""
from __future__ import annotations

import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
SCAN = ["include", "cgspline_amd/csrc", "cgspline_amd", "oracle", "tests", "bench.py", "__graft_entry__.py",
        "DESIGN.md", "INTEGRATION.md", "README.md"]
EXTS = (".h", ".hpp", ".hip", ".cpp", ".py", ".cu")

RANGES = r"\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*"
CITE = re.compile(r"(?<![\w.])((?:[A-Za-z_]+/)*[A-Za-z_][A-Za-z_0-9]*\.(?:h|cu|cpp))?(\.(?:h|cu|cpp))?:(" + RANGES + r")(?![\d])")
IDENT = re.compile(r"[A-Za-z_][A-Za-z_0-9]*(?:::[A-Za-z_~][A-Za-z_0-9]*)*")
COMMON = {"the", "and", "for", "with", "from", "this", "that", "where", "when", "each", "into", "over", "only",
          "file", "line", "lines", "see", "per", "its", "not", "are", "one", "two", "all", "both", "same",
          "reference", "restates", "restated", "follows", "as", "in", "of", "on", "at", "by", "to", "or", "is",
          "device", "host", "kernel", "kernels", "path", "loop", "step", "steps", "call", "calls", "test",
          "function", "functions", "variant", "form", "order", "value", "values", "else", "then", "true",
          "false", "return", "returns", "float", "uint32_t", "int", "bool", "void", "const", "auto"}


def ref_files():
    out = {}
    for d, _, fs in os.walk(REF):
        for f in fs:
            if f.endswith((".h", ".cu", ".cpp")):
                rel = os.path.relpath(os.path.join(d, f), REF)
                out.setdefault(f, []).append(rel)
    return out


def source_files():
    for p in SCAN:
        full = os.path.join(ROOT, p)
        if os.path.isfile(full):
            yield full
            continue
        for d, dirs, fs in os.walk(full):
            dirs[:] = [x for x in dirs if x not in ("__pycache__", "_build", "golden")]
            if full.endswith("cgspline_amd") and d != full:
                continue
            for f in sorted(fs):
                if f.endswith(EXTS) and f != "citations.py":
                    yield os.path.join(d, f)


def parse_ranges(r):
    out = []
    for part in r.replace(" ", "").split(","):
        a, _, b = part.partition("-")
        out.append((int(a), int(b) if b else int(a)))
    return out


def subject(text):
    """The symbol named just before a citation: the last identifier that
    looks like code (has ::, _, an inner capital, or is followed by '(')."""
    cands = []
    for m in IDENT.finditer(text):
        w = m.group(0)
        last = w.split("::")[-1]
        if len(last) < 3 or last.lower() in COMMON or last.isupper() and len(last) < 4:
            continue
        codey = "::" in w or "_" in last or re.search(r"[a-z][A-Z]", last) or (m.end() < len(text) and text[m.end()] == "(")
        if codey:
            cands.append(last)
    return cands[-1] if cands else None


def scan():
    files = ref_files()
    cache = {}
    results = []
    for path in sorted(set(source_files())):
        rel = os.path.relpath(path, ROOT)
        with open(path, encoding="utf-8", errors="replace") as f:
            lines = f.read().split("\n")
        last_file = None
        for ln, line in enumerate(lines, 1):
            if not line.strip():
                last_file = None
            pos = 0
            for m in CITE.finditer(line):
                name, ext, rng = m.group(1), m.group(2), m.group(3)
                if name is None and ext is None:
                    # bare ":R" continues the last file: after a backtick, '(' or
                    # a space, and not a Python slice such as a[..., :3]
                    after = line[m.end():m.end() + 1]
                    if last_file is None or m.start() == 0 or line[m.start() - 1] not in "`( ," or after == "]":
                        continue
                    target = last_file
                elif name is None:
                    if last_file is None:
                        continue
                    target = os.path.splitext(last_file)[0] + ext
                else:
                    target = name
                base = os.path.basename(target)
                cands = files.get(base, [])
                if "/" in target:
                    cands = [c for c in cands if c.endswith(target)]
                text_before = line[pos:m.start()]
                pos = m.end()
                rec = dict(src=rel, line=ln, cite=m.group(0), file=target, ranges=parse_ranges(rng), problems=[])
                if not cands:
                    rec["problems"].append("no such reference file")
                    results.append(rec)
                    continue
                ref = cands[0]
                last_file = target
                if ref not in cache:
                    with open(os.path.join(REF, ref), encoding="utf-8", errors="replace") as f:
                        cache[ref] = f.read().split("\n")
                body = cache[ref]
                n = len(body) - (1 if body and body[-1] == "" else 0)
                rec["ref"] = ref
                for a, b in rec["ranges"]:
                    if a < 1 or b < a or b > n:
                        rec["problems"].append(f"range {a}-{b} outside 1-{n}")
                # the symbol check reads code comments; prose tables name
                # enclosing functions whose bodies span the cited lines
                sym = None if rel.endswith(".md") else subject(text_before)
                rec["symbol"] = sym
                if sym and not rec["problems"]:
                    occurs = [i + 1 for i, l in enumerate(body) if re.search(r"\b" + re.escape(sym) + r"\b", l)]
                    if occurs and not any(a - 3 <= o <= b + 3 for a, b in rec["ranges"] for o in occurs):
                        rec["problems"].append(f"'{sym}' occurs at lines {occurs[:8]}, not in the cited ranges")
                results.append(rec)
    return results


if __name__ == "__main__":
    res = scan()
    bad = [r for r in res if r["problems"]]
    for r in bad:
        print(f"{r['src']}:{r['line']}: {r['cite']} -> {r['file']}: {'; '.join(r['problems'])}")
    print(f"{len(res)} citations, {len(bad)} with problems", file=sys.stderr)
