"""Extract the 57 w(z) known-answer points of the reference's Faddeeva
self-test (src/Faddeeva.cpp:1922-2090, values from WolframAlpha/Maple) into
faddeeva_kat.json.  Run in a container that has the reference tree:

    python tests/golden/make_faddeeva_kat.py /root/reference/src/Faddeeva.cpp

The fixture holds numbers only: inputs z and expected w(z), as strings so
that Inf/NaN survive JSON.
"""
import json
import os
import re
import sys


def parse_block(text, start_marker):
    i = text.index(start_marker)
    j = text.index("};", i)
    body = text[text.index("{", i) + 1:j]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    vals = re.findall(r"C\(\s*([^,()]+?)\s*,\s*([^,()]+?)\s*\)", body)
    return [(a.strip(), b.strip()) for a, b in vals]


def main(src):
    text = open(src).read()
    seg = text[text.index("#define NTST 57"):]
    z = parse_block(seg, "cmplx z[NTST]")
    w = parse_block(seg, "cmplx w[NTST]")
    assert len(z) == 57 and len(w) == 57, (len(z), len(w))
    out = {"source": "src/Faddeeva.cpp:1922-2090", "tolerance": 1e-13,
           "z": z, "w": w}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "faddeeva_kat.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/Faddeeva.cpp")
