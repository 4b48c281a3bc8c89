"""Writes tests/golden/mc_table_digest.txt: the SHA-256 of the reference's
marching-cubes case table (src/geometry/MARCHING_CUBES_TRIANGLES.include, the
addVertexPairTriangle(a,b,c,index) calls of each case in order), serialised as
one line per case: "case: a,b,c a,b,c ...". The digest lets the table pin run
where the reference sources are not mounted. Run from the repository root:
    python tests/golden/make_mc_table_digest.py
"""
import hashlib
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
REF_TABLE = "/root/reference/src/geometry/MARCHING_CUBES_TRIANGLES.include"


def serialise(table):
    """table: 256 lists of (a, b, c) triangles in reference edge ids (1..12)."""
    return "\n".join("%d: %s" % (c, " ".join("%d,%d,%d" % t for t in table[c])) for c in range(256)) + "\n"


def reference_table(path=REF_TABLE):
    parts = re.split(r"case (\d+):", open(path).read())
    ref = {}
    for i in range(1, len(parts), 2):
        ref[int(parts[i])] = [tuple(int(v) for v in t) for t in
                              re.findall(r"addVertexPairTriangle\((\d+),(\d+),(\d+),index\)", parts[i + 1])]
    assert sorted(ref) == list(range(256))
    return [ref[c] for c in range(256)]


def main():
    digest = hashlib.sha256(serialise(reference_table()).encode()).hexdigest()
    with open(os.path.join(HERE, "mc_table_digest.txt"), "w") as f:
        f.write(digest + "\n")
    print(digest)


if __name__ == "__main__":
    main()
