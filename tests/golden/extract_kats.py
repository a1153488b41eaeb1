"""Extract known-answer arrays from the reference's own tests into JSON.

Run in a checkout that has the reference at /root/reference (it is not
needed afterwards; the JSON written next to this script is the fixture):

    python tests/golden/extract_kats.py

For every listed test file, each array initialised with a braced list —
`xarray<T> name = {...};`, `xtensor<T, N> name = {...};`, `xarray<T> name{...};`
— becomes one record keyed "<file>:<TEST name>:<variable>" holding the C++
element type, the shape implied by the nesting, the values (row-major) and
the line number; numpy-generated scalars `double py_name = v;` become records
of shape [].  Only numbers are kept; the surrounding C++ is not copied.
"""

from __future__ import annotations

import json
import os
import re
import sys

REF = "/root/reference/test"
FILES = [
    "test_xaccumulator.cpp",
    "test_extended_broadcast_view.cpp",
    "test_xreducer.cpp",
    "test_xnorm.cpp",
    "test_xmath.cpp",
    "test_xbroadcast.cpp",
    "test_strided_assign.cpp",
    "test_extended_xmath_reducers.cpp",
    "test_xnpy.cpp",
    "test_xindex_view.cpp",
]
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat_arrays.json")

DECL = re.compile(
    r"(?:xt::)?(xarray|xtensor)\s*<\s*([A-Za-z_:0-9 ]+?)\s*(?:,\s*\d+\s*)?>\s+([A-Za-z_][A-Za-z_0-9]*)\s*(=\s*)?\{")
TEST = re.compile(r"TEST\s*\(\s*([A-Za-z_0-9]+)\s*,\s*([A-Za-z_0-9]+)\s*\)")
NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")
# numpy-generated scalar results: `double py_name = 0.123;`
SCALAR = re.compile(r"\bdouble\s+(py_[A-Za-z_0-9]*)\s*=\s*([-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?)\s*;")


def braced(text: str, start: int) -> tuple[str, int]:
    """The balanced {...} block starting at text[start] == '{'."""
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return text[start:i + 1], i + 1
    raise ValueError("unbalanced braces")


def parse(block: str):
    """Nested brace list of numbers -> nested python lists."""
    pos = 0

    def value():
        nonlocal pos
        while block[pos] in " \n\t\r,":
            pos += 1
        if block[pos] == "{":
            pos += 1
            items = []
            while True:
                while block[pos] in " \n\t\r,":
                    pos += 1
                if block[pos] == "}":
                    pos += 1
                    return items
                items.append(value())
        m = NUM.match(block, pos)
        if not m:
            raise ValueError(f"not a number at {block[pos:pos + 20]!r}")
        pos = m.end()
        # skip literal suffixes / casts like 1.f, 2u, short(3)
        while pos < len(block) and block[pos] in "fFuUlL":
            pos += 1
        txt = m.group(0)
        return float(txt) if any(c in txt for c in ".eE") else int(txt)

    return value()


def shape_of(v):
    s = []
    while isinstance(v, list):
        s.append(len(v))
        if not v:
            break
        v = v[0]
    return s


def flatten(v):
    if isinstance(v, list):
        out = []
        for x in v:
            out.extend(flatten(x))
        return out
    return [v]


def main() -> int:
    if not os.path.isdir(REF):
        print("reference tests not found at", REF, file=sys.stderr)
        return 1
    records = {}
    for fname in FILES:
        text = open(os.path.join(REF, fname)).read()
        # drop short(...) wrappers so the numbers parse
        text_n = re.sub(r"short\(([^()]*)\)", r"\1", text)
        tests = [(m.start(), f"{m.group(1)}.{m.group(2)}") for m in TEST.finditer(text_n)]
        for m in DECL.finditer(text_n):
            brace = m.end() - 1
            try:
                block, _ = braced(text_n, brace)
                val = parse(block)
            except (ValueError, IndexError):
                continue
            if not isinstance(val, list):
                continue
            tname = ""
            for pos, name in tests:
                if pos < m.start():
                    tname = name
            line = text_n.count("\n", 0, m.start()) + 1
            key = f"{fname}:{tname}:{m.group(3)}"
            records[key] = {
                "container": m.group(1),
                "ctype": m.group(2).replace("std::", "").strip(),
                "shape": shape_of(val),
                "values": flatten(val),
                "line": line,
            }
        for m in SCALAR.finditer(text_n):
            tname = ""
            for pos, name in tests:
                if pos < m.start():
                    tname = name
            line = text_n.count("\n", 0, m.start()) + 1
            records[f"{fname}:{tname}:{m.group(1)}"] = {
                "container": "scalar", "ctype": "double", "shape": [],
                "values": [float(m.group(2))], "line": line,
            }
    with open(OUT, "w") as f:
        json.dump(records, f, indent=0, sort_keys=True)
    print(f"wrote {len(records)} arrays to {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
