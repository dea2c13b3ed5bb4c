"""Writes topology_L4.json: the interaction lists the reference assigns.

Reads the assignment statements of FMM2DTree::assign_Child{0..3}_Interaction
(src/AFMM2DTree.cpp:65-943 of the reference) as text and replays them level
by level (assign_Tree_Interactions, 947-967) for a tree of 4 levels.  The
output is data only: for every level 0..4 and box, its neighborNumbers[8],
innerNumbers[16] and outerNumbers[24] (-1 = absent).

Run in the build container (the reference is mounted there):
    python tests/golden/make_topology_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
LEVELS = 4


def parse_rules(path):
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)      # drop the box diagrams
    src = re.sub(r"//[^\n]*", "", src)
    rules = {}
    for m in re.finditer(r"void FMM2DTree::assign_Child(\d)_Interaction\(int j, int k\)\s*\{", src):
        child = int(m.group(1))
        body = src[m.end():]
        depth, end = 1, 0
        for end, ch in enumerate(body):
            depth += ch == "{"
            depth -= ch == "}"
            if depth == 0:
                break
        body = body[:end]
        g = re.search(r"int nC\s*=\s*4\*k\s*([+-]\s*\d+)?\s*;", body).group(1)
        off = int(g.replace(" ", "")) if g else 0
        assert off == child
        out = []
        cur = None
        for st in re.finditer(
                r"nN\s*=\s*tree\[j\]\[k\]\.neighborNumbers\[(\d+)\]"
                r"|tree\[nL\]\[nC\]\.(neighbor|inner|outer)Numbers\[(\d+)\]\s*=\s*"
                r"(?:nC\s*([+-])\s*(\d+)|tree\[j\]\[nN\]\.childrenNumbers\[(\d)\])", body):
            if st.group(1) is not None:
                cur = int(st.group(1))
                continue
            lst, idx = st.group(2), int(st.group(3))
            if st.group(4):
                d = int(st.group(5)) * (1 if st.group(4) == "+" else -1)
                out.append((lst, idx, "sib", d, None))
            else:
                out.append((lst, idx, "nbr", cur, int(st.group(6))))
        rules[child] = out
    assert sorted(rules) == [0, 1, 2, 3]
    return rules


def replay(rules, levels):
    tree = [[{"neighbor": [-1] * 8, "inner": [-1] * 16, "outer": [-1] * 24}]]
    for j in range(levels):
        nxt = [{"neighbor": [-1] * 8, "inner": [-1] * 16, "outer": [-1] * 24}
               for _ in range(4 ** (j + 1))]
        for k in range(4 ** j):
            for c in range(4):
                nC = 4 * k + c
                for lst, idx, kind, a, b in rules[c]:
                    if kind == "sib":
                        nxt[nC][lst][idx] = nC + a
                    else:
                        nN = tree[j][k]["neighbor"][a]
                        if nN != -1:
                            nxt[nC][lst][idx] = 4 * nN + b
        tree.append(nxt)
    return tree


def main():
    rules = parse_rules(os.path.join(REF, "src", "AFMM2DTree.cpp"))
    tree = replay(rules, LEVELS)
    data = {"source": "src/AFMM2DTree.cpp:65-967 (assignment statements replayed)",
            "levels": LEVELS,
            "neighbor": [[b["neighbor"] for b in lv] for lv in tree],
            "inner": [[b["inner"] for b in lv] for lv in tree],
            "outer": [[b["outer"] for b in lv] for lv in tree]}
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "topology_L4.json"), "w") as f:
        json.dump(data, f, separators=(",", ":"))
    print("rules per child:", {c: len(r) for c, r in rules.items()})


if __name__ == "__main__":
    main()
