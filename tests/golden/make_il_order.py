"""Extract the interaction-list push order of the reference octree as a data table.

Reads ``src/HODLR3DTree.cpp`` of SAFRAN-LAB/HODLR3D *as text* (lines 110-1489,
``assign_Child0_Interaction`` .. ``assign_Child7_Interaction``) and records, for
each child position c = 0..7, the sequence of assignments the reference makes:

* ``["sib_il", c2]``          sibling 8k+c2 appended to the interaction list;
* ``["sib_nbr", slot, c2]``   sibling 8k+c2 stored in neighbour slot ``slot``;
* ``["par", P, ops]``         block guarded by ``parent.neighborNumbers[P] != -1``
  where ``ops`` holds ``["il", c2]`` (child c2 of that parent neighbour appended
  to the interaction list) and ``["nbr", slot, c2]`` (stored in neighbour slot).

Nothing is compiled or executed; the output ``il_order.json`` is a fixture the
CPU tests compare the oracle's closed-form rule against.  Run in a checkout that
has the reference mounted at /root/reference:

    python tests/golden/make_il_order.py
"""
import json
import os
import re
import sys

REF = os.environ.get("HODLR3D_REFERENCE", "/root/reference")


def strip_comments(text):
    return "\n".join(line.split("//", 1)[0] for line in text.splitlines())


def match_brace(text, open_pos):
    depth = 0
    for i in range(open_pos, len(text)):
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return i
    raise ValueError("unbalanced braces")


RE_SIB_IL = re.compile(r"interactionList\.push_back\(8\*k\+(\d)\)")
RE_SIB_NBR = re.compile(r"neighborNumbers\[(\d+)\]\s*=\s*8\*k\+(\d)")
RE_PAR = re.compile(r"nN\s*=\s*tree\[j\]\[k\]\.neighborNumbers\[(\d+|n)\]")
RE_IL = re.compile(r"interactionList\.push_back\(tree\[j\]\[nN\]\.childrenNumbers\[(\d+|i)\]\)")
RE_NBR = re.compile(r"neighborNumbers\[(\d+)\]\s*=\s*tree\[j\]\[nN\]\.childrenNumbers\[(\d+)\]")
RE_FOR_N = re.compile(r"for \(size_t n = (\d+); n <= (\d+); n\+\+\)\s*\{")
RE_EXCL = re.compile(r"if \(([^)]*)\)\s*\{\}")


def events(body):
    """Yield assignment events of one child function in source order."""
    out = []
    pos = 0
    cur = None
    pat = re.compile("|".join([RE_FOR_N.pattern, RE_SIB_IL.pattern, RE_SIB_NBR.pattern,
                               RE_PAR.pattern, RE_IL.pattern, RE_NBR.pattern]))
    while True:
        m = pat.search(body, pos)
        if not m:
            break
        s = m.group(0)
        if RE_FOR_N.match(s):
            lo, hi = int(m.group(1)), int(m.group(2))
            open_pos = m.end() - 1
            close = match_brace(body, open_pos)
            loop = body[open_pos:close + 1]
            ex = RE_EXCL.search(loop)
            excluded = set()
            if ex:
                excluded = {int(v) for v in re.findall(r"n\s*==\s*(\d+)", ex.group(1))}
            assert RE_PAR.search(loop).group(1) == "n"
            assert RE_IL.search(loop).group(1) == "i"
            assert not RE_NBR.search(loop)
            for n in range(lo, hi + 1):
                if n not in excluded:
                    out.append(["par", n, [["il", c] for c in range(8)]])
            cur = None
            pos = close + 1
            continue
        if RE_SIB_IL.fullmatch(s):
            out.append(["sib_il", int(RE_SIB_IL.fullmatch(s).group(1))])
        elif RE_SIB_NBR.fullmatch(s):
            g = RE_SIB_NBR.fullmatch(s)
            out.append(["sib_nbr", int(g.group(1)), int(g.group(2))])
        elif RE_PAR.fullmatch(s):
            cur = ["par", int(RE_PAR.fullmatch(s).group(1)), []]
            out.append(cur)
        elif RE_IL.fullmatch(s):
            v = RE_IL.fullmatch(s).group(1)
            cur[2].extend([["il", c] for c in range(8)] if v == "i" else [["il", int(v)]])
        elif RE_NBR.fullmatch(s):
            g = RE_NBR.fullmatch(s)
            cur[2].append(["nbr", int(g.group(1)), int(g.group(2))])
        pos = m.end()
    return out


def main():
    path = os.path.join(REF, "src", "HODLR3DTree.cpp")
    text = strip_comments(open(path).read())
    table = {}
    for c in range(8):
        head = re.search(r"void HODLR3DTree::assign_Child%d_Interaction\(int j, int k\)\s*\{" % c, text)
        close = match_brace(text, head.end() - 1)
        table[str(c)] = events(text[head.end():close])
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "il_order.json")
    with open(dst, "w") as f:
        json.dump({"source": "src/HODLR3DTree.cpp:110-1489", "children": table}, f, indent=0)
    print("wrote", dst, file=sys.stderr)


if __name__ == "__main__":
    main()
