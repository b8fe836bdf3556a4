"""Every citation of the reference's sources (``src/<file>.py:a-b``) in the
boundary header, the package, the oracle, the tests and the bench must name
lines that exist, and a citation written next to a reference function's name
must overlap that function's lines.  Reads the reference's files as text when
they are present (skipped otherwise)."""
import ast
import pathlib
import re

import pytest

from tests.conftest import ROOT

REF = pathlib.Path("/root/reference")
REF_FILES = ["microfacet.py", "svbrdf.py", "optimization.py", "imageio.py", "scripts.py", "materialgan.py",
             "descriptor.py", "capture.py", "globalvar.py", "mitsubarender.py"]

# src/<file>.py:<ranges>, or <file>.py:<ranges> for the reference's own
# file names (citations inside a sentence that already named src/)
CITE = re.compile(r"(?:src/)?(?P<file>(?:higan_models/)?[a-z_0-9]+\.py):(?P<ranges>\d+(?:-\d+)?(?:,\d+(?:-\d+)?)*)")
# "name (src/file.py:a-b" or "name`` (file.py:a-b": a function named right before its citation
NAMED = re.compile(r"(?P<name>[A-Za-z_][A-Za-z_0-9]*)(?:``|`|\(\))?\s*\((?:src/)?(?P<file>[a-z_0-9]+\.py):(?P<a>\d+)(?:-(?P<b>\d+))?")

pytestmark = pytest.mark.skipif(not (REF / "src" / "microfacet.py").exists(),
                                reason="the reference's sources are not present")


def _sources():
    pats = ["include/*.h", "svbrdf-diff-renderer_amd/*.py", "svbrdf-diff-renderer_amd/csrc/*.hip",
            "svbrdf-diff-renderer_amd/csrc/*.hpp", "oracle/*.py", "tests/*.py", "tests/golden/*.py", "bench.py",
            "__graft_entry__.py"]
    out = []
    for p in pats:
        out.extend(sorted(ROOT.glob(p)))
    return [p for p in out if p.name != "test_citations.py"]


def _ref_path(name):
    for base in (REF / "src", REF / "src" / "higan_models"):
        if (base / name).exists():
            return base / name
    if (REF / "src" / name).exists():
        return REF / "src" / name
    return None


def _spans(path):
    """name -> (first line, last line) of every function and class of a
    reference file."""
    tree = ast.parse(path.read_text())
    spans = {}
    for node in ast.walk(tree):
        if isinstance(node, (ast.FunctionDef, ast.ClassDef)):
            spans.setdefault(node.name, []).append((node.lineno, node.end_lineno))
    return spans


def test_cited_lines_exist():
    lengths = {}
    bad = []
    n = 0
    for src in _sources():
        for lineno, line in enumerate(src.read_text().splitlines(), 1):
            for m in CITE.finditer(line):
                f = m.group("file")
                if f.split("/")[-1] not in REF_FILES and not f.startswith("higan_models/"):
                    continue   # torch/..., this repository's own files
                if "torch/" in line[max(0, m.start() - 6):m.start()]:
                    continue
                ref = _ref_path(f)
                if ref is None:
                    bad.append(f"{src.relative_to(ROOT)}:{lineno}: {m.group(0)} names no reference file")
                    continue
                size = lengths.setdefault(ref, len(ref.read_text().splitlines()))
                for r in m.group("ranges").split(","):
                    a, _, b = r.partition("-")
                    a, b = int(a), int(b or a)
                    n += 1
                    if not (1 <= a <= b <= size):
                        bad.append(f"{src.relative_to(ROOT)}:{lineno}: {m.group(0)} ({f} has {size} lines)")
    assert n > 100, n
    assert not bad, "\n".join(bad)


def test_named_citations_overlap_the_function():
    spans = {}
    bad = []
    checked = 0
    for src in _sources():
        for lineno, line in enumerate(src.read_text().splitlines(), 1):
            for m in NAMED.finditer(line):
                f = m.group("file")
                if f not in REF_FILES:
                    continue
                ref = _ref_path(f)
                table = spans.setdefault(ref, _spans(ref))
                name = m.group("name")
                if name not in table:
                    continue
                a = int(m.group("a"))
                b = int(m.group("b") or a)
                checked += 1
                if not any(a <= e and s <= b for s, e in table[name]):
                    bad.append(f"{src.relative_to(ROOT)}:{lineno}: {name} cited at {f}:{a}-{b}, "
                               f"defined at {table[name]}")
    assert checked >= 10, checked
    assert not bad, "\n".join(bad)


def test_boundary_header_cites_the_io_methods_exactly():
    """The codec entry points of the header replace SvbrdfIO methods; their
    citations must be those methods' own line ranges."""
    header = (ROOT / "include" / "svbrdf_hip.h").read_text()
    table = _spans(REF / "src" / "svbrdf.py")
    for name in ("load_textures_th", "save_textures_th", "load_images_th", "save_images_th"):
        (s, e), = table[name]
        assert re.search(rf"{name}[,\s(*]*src/svbrdf\.py:{s}-{e}", header), (name, s, e)
