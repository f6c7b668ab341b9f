"""Turn the reference's golden test vectors into JSON fixtures.

Reads the array literals of the 3D section of
  ExaHyPE/exahype/tests/kernels/testdata/generic_euler_testdata.cpp  (#ifdef Dim3, :1348-6199)
and of
  ExaHyPE/exahype/tests/kernels/testdata/limiter_testdata.cpp
and writes them, keyed "<namespace>/<array>", to tests/golden/*.json.  Only
data (inputs and expected outputs) is kept.  Run in a container that has the
reference mounted:  python tests/golden/extract_golden.py
"""
import json
import os
import re
import sys

REF = "/root/reference/ExaHyPE/exahype/tests/kernels/testdata"
HERE = os.path.dirname(os.path.abspath(__file__))

_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def parse_arrays(text):
    """Yield (namespace path, name, values) for 'const double name[n] = {...};'."""
    out = {}
    stack = []
    pos = 0
    tok = re.compile(r"namespace\s+(\w+)\s*\{|\}\s*(?://[^\n]*)?|const\s+double\s+(\w+)\s*\[\s*(\d+)\s*\]\s*=?\s*\{")
    while True:
        m = tok.search(text, pos)
        if not m:
            break
        if m.group(1):
            stack.append(m.group(1))
            pos = m.end()
        elif m.group(2):
            name, n = m.group(2), int(m.group(3))
            end = text.index("}", m.end())
            body = text[m.end():end]
            vals = [float(x) for x in _NUM.findall(body)]
            if not vals:  # empty placeholder arrays hold no data
                pos = end + 1
                continue
            if len(vals) != n:
                raise SystemExit("%s/%s: expected %d values, got %d" % ("/".join(stack), name, n, len(vals)))
            out["/".join(stack[-1:] + [name])] = vals
            pos = end + 1
        else:
            if stack:
                stack.pop()
            pos = m.end()
    return out


def section(text, begin, end):
    i = text.index(begin)
    j = text.index(end, i)
    return text[i + len(begin):j]


def main():
    src = open(os.path.join(REF, "generic_euler_testdata.cpp")).read()
    dim3 = _strip_comments(section(src, "#ifdef Dim3", "#endif  // Dim3"))
    arrays = parse_arrays(dim3)
    with open(os.path.join(HERE, "generic_euler_3d.json"), "w") as f:
        json.dump(arrays, f)
    print("generic_euler_3d.json:", len(arrays), "arrays")
    lim = open(os.path.join(REF, "limiter_testdata.cpp")).read()
    if "#ifdef Dim3" in lim:
        lim = section(lim, "#ifdef Dim3", "#endif // Dim3")
    arrays = parse_arrays(_strip_comments(lim))
    with open(os.path.join(HERE, "limiter_3d.json"), "w") as f:
        json.dump(arrays, f)
    print("limiter_3d.json:", len(arrays), "arrays", sorted(arrays)[:20])


if __name__ == "__main__":
    sys.exit(main())
