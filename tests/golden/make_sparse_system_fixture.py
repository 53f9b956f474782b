"""Transcribe the reference's 36x36 block test system into a JSON fixture.

Reads the data (numbers only) held by g2o's unit-test helper
unit_test/solver/sparse_system_helper.cpp: the serialized 12 x 12 block
matrix with 3x3 blocks (sparseMatrixString, :52-148), its dense inverse
(denseInverseMatrixString), and the vectors b (createTestVectorB, :255-294)
and x (createTestVectorX, :296-337). Run where the reference is mounted;
the tests only read the JSON it writes.
"""
import json
import os
import re
import sys

SRC = "/root/reference/unit_test/solver/sparse_system_helper.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sparse_system_36.json")


def function_body(text, name):
    i = text.index(name + "(")
    j = text.index("{", i)
    depth, k = 0, j
    while True:
        if text[k] == "{":
            depth += 1
        elif text[k] == "}":
            depth -= 1
            if depth == 0:
                return text[j:k]
        k += 1


def stream_lines(body):
    return re.findall(r'aux\s*<<\s*"([^"]*)"', body)


def vector_values(body):
    return [float(v) for v in re.findall(r"result\(idx\+\+\)\s*=\s*([-+0-9.eE]+);", body)]


def main():
    text = open(SRC).read()
    lines = stream_lines(function_body(text, "std::string sparseMatrixString"))
    rows, cols, blocks = [], [], []
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith("BLOCK"):
            r, c = (int(v) for v in ln.split(":")[1].split())
            vals = [float(v) for k in range(3) for v in lines[i + 1 + k].split()]
            rows.append(r)
            cols.append(c)
            blocks.append(vals)
            i += 4
        else:
            i += 1
    inv_lines = stream_lines(function_body(text, "std::string denseInverseMatrixString"))
    inv = [[float(v) for v in ln.split()] for ln in inv_lines if not ln.startswith("#") and ln.strip()]
    b = vector_values(function_body(text, "g2o::VectorX createTestVectorB"))
    x = vector_values(function_body(text, "g2o::VectorX createTestVectorX"))
    assert len(b) == 36 and len(x) == 36 and len(inv) == 36 and all(len(r) == 36 for r in inv)
    data = {"source": "unit_test/solver/sparse_system_helper.cpp", "block_size": 3, "num_blocks": 12,
            "rows": rows, "cols": cols, "blocks": [v for blk in blocks for v in blk], "b": b, "x": x,
            "inverse": [v for r in inv for v in r]}
    json.dump(data, open(OUT, "w"), indent=0)
    print(f"wrote {OUT}: {len(rows)} blocks")


if __name__ == "__main__":
    sys.exit(main())
