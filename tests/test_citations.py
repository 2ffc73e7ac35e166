"""Every `file:line` citation of the reference in the boundary headers, the package and the oracle
must name a line that exists in the reference and is not blank; the method citations of the
host classes must land on the method they name.  Skipped where the reference is not present
(it is read here only as text)."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")

pytestmark = pytest.mark.skipif(not REF.is_dir(), reason="reference tree not present")

CITE = re.compile(r"((?:[\w-]+/)*[\w-]+\.(?:h|hpp|cpp|comp|vert|frag)):(\d+(?:-\d+)?(?:(?:,\s*|\s+and\s+)\d+(?:-\d+)?)*)")
SOURCES = ["include/*.h", "include/*.hpp", "watercube_amd/*.py", "watercube_amd/csrc/*", "oracle/*.c",
           "oracle/*.py", "bench.py", "__graft_entry__.py", "DESIGN.md", "INTEGRATION.md", "README.md", "tests/*.py"]


def _ref_files():
    return [p for p in REF.rglob("*") if p.is_file() and p.suffix in (".h", ".hpp", ".cpp", ".comp", ".vert", ".frag")]


def _resolve(name: str, files):
    if (REF / name).is_file():
        return REF / name
    hits = [p for p in files if p.as_posix().endswith("/" + name)]
    return hits[0] if len(hits) == 1 else None


def _citations():
    files = _ref_files()
    out = []
    for pattern in SOURCES:
        for src in sorted(ROOT.glob(pattern)):
            if src.suffix in (".so", ".o") or not src.is_file():
                continue
            text = src.read_text(errors="replace")
            for ln, line in enumerate(text.splitlines(), 1):
                for m in CITE.finditer(line):
                    name = m.group(1)
                    if name.startswith(("include/", "watercube_amd/", "tests/", "oracle/")):
                        continue  # this repository's own files
                    target = _resolve(name, files)
                    spans = [tuple(int(x) for x in s.split("-")) for s in re.split(r",\s*|\s+and\s+", m.group(2))]
                    out.append((f"{src.relative_to(ROOT)}:{ln}", name, target, spans))
    return out


def test_citations_name_existing_lines():
    cites = _citations()
    assert len(cites) > 100, "the citation scan found too few citations"
    bad = []
    for where, name, target, spans in cites:
        if target is None:
            bad.append(f"{where}: {name} is not a unique file of the reference")
            continue
        lines = target.read_text(errors="replace").splitlines()
        for span in spans:
            a, b = span[0], span[-1]
            if not (1 <= a <= b <= len(lines)):
                bad.append(f"{where}: {name}:{a}-{b} outside its {len(lines)} lines")
            elif all(not lines[i - 1].strip() for i in range(a, b + 1)):
                bad.append(f"{where}: {name}:{a}-{b} is blank")
    assert not bad, "\n".join(bad)


# (citing file, cited reference line, text that must appear on that line)
METHOD_CITES = [
    ("include/watercube.hpp", "src/core/Fluid.h:32", "class Fluid"),
    ("include/watercube.hpp", "src/core/Fluid.h:141", "};"),
    ("include/watercube.hpp", "src/core/Fluid.h:85", "thisRef"),
    ("include/watercube.hpp", "src/core/Fluid.h:52", "setMouseRay"),
    ("include/watercube.hpp", "src/core/Sort.h:20", "class Sort"),
    ("include/watercube.h", "src/core/Sort.h:35", "getCountBuffer"),
    ("include/watercube.h", "src/core/Sort.h:36", "getOffsetBuffer"),
    ("include/watercube.h", "Sort.h:37", "getSortedBuffer"),
    ("include/watercube.h", "src/core/util.h:18", "WORK_GROUP_SIZE"),
    ("watercube_amd/fluid.py", "src/core/Fluid.h:32", "class Fluid"),
    ("watercube_amd/fluid.py", "src/core/Sort.h:20", "class Sort"),
    ("DESIGN.md", "src/core/Fluid.h:78", "runDensityProg"),
    ("DESIGN.md", "src/core/Fluid.h:79", "runUpdateProg"),
    ("watercube_amd/csrc/wc_internal.hpp", "update.comp:78", "size - kernelRadius"),
    ("watercube_amd/csrc/wc_internal.hpp", "update.comp:208", "size - border"),
    ("watercube_amd/csrc/wc_internal.hpp", "density.comp:54", "kernelRadius * kernelRadius"),
    ("watercube_amd/csrc/wc_kernels.hip", "advect.comp:30", "wallDamping = 0.3"),
    ("oracle/wc_oracle.c", "reorder.comp:42", "atomicAdd"),
    ("include/watercube.h", "update.comp:202", "wallDamping = 0.3"),
    ("include/watercube.h", "update.comp:203", "border = 0.001"),
    ("include/watercube.h", "update.comp:5", "MAX_SPEED"),
    ("include/watercube.hpp", "Fluid.cpp:308", "time_scale_"),
    ("DESIGN.md", "density.comp:117", "dist >= kernelRadius"),
    ("DESIGN.md", "update.comp:173", "dist >= kernelRadius"),
]


@pytest.mark.parametrize("src,cite,token", METHOD_CITES)
def test_method_citations_land_on_the_method(src, cite, token):
    name, line = cite.rsplit(":", 1)
    text = (ROOT / src).read_text()
    # the line is cited alone, or as the first or last line of a range
    assert re.search(re.escape(name) + r":(?:\d+-)?" + line + r"(?:-\d+)?\b", text), f"{src} no longer cites {cite}"
    target = _resolve(name, _ref_files())
    assert target is not None
    assert token in target.read_text().splitlines()[int(line) - 1], f"{cite} does not hold {token!r}"
