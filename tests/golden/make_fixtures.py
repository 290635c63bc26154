"""Regenerate the golden fixtures of tests/golden from the reference tree.

Run in the build container (the reference is not present on GPU machines):
    python tests/golden/make_fixtures.py [/root/reference]

Outputs (committed):
  bubble_2D_fi001.dat.xz       tests/bubble_2D/reference_solution/fi001.dat,
                               xz-compressed (PETSc binary Vec, fi at step 100)
  bubble_2D_hots_fi001.dat.xz  tests/bubble_2D_hots/reference_solution/fi001.dat
  ff_stencil_d3.json           the psi-gradient stencil of
  ff_stencil_d2.json           src/lbm/lbm_forcing.F90 (LBMAddFluidFluidForcesD3/D2)
                               as data: per neighbour block, in source order, the
                               isotropy section, the offset, the ffw index, the
                               gradient terms (axis, sign, factor) and the
                               weight-sum terms, plus the wall-test condition as
                               cells joined by and/or.
Only the reference's text and data files are read; nothing in it is run.
"""
import json
import lzma
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def copy_fixture(rel, name):
    with open(os.path.join(REF, rel), "rb") as f:
        raw = f.read()
    with open(os.path.join(HERE, name), "wb") as f:
        f.write(lzma.compress(raw, preset=9 | lzma.PRESET_EXTREME))


def parse_cond(cond, names):
    """'walls(i+2,j,k).eq.0.and.(walls(i+1,j,k).eq.0.or.walls(i+1,j,k+1).eq.0)'
    -> nested ['and'/'or', ...] of offsets"""
    s = cond.replace(" ", "")
    s = re.sub(r"\.eq\.0\.(?=\.and\.|\.or\.|\)|$)", ".eq.0", s)  # real literal 0. -> 0
    toks = re.findall(r"walls\([^)]*\)\.eq\.0|\.and\.|\.or\.|\(|\)", s)

    def off(tok):
        inside = re.match(r"walls\(([^)]*)\)", tok).group(1).split(",")
        o = []
        for p, v in zip(inside, names):
            o.append(0 if p == v else int(p[len(v):]))
        return o

    pos = 0

    def expr():
        nonlocal pos
        items, op = [term()], None
        while pos < len(toks) and toks[pos] in (".and.", ".or."):
            o = toks[pos][1:-1]
            pos += 1
            if op is None:
                op = o
            elif op != o:
                raise ValueError("mixed and/or without parentheses: " + cond)
            items.append(term())
        return items[0] if op is None else [op] + items

    def term():
        nonlocal pos
        t = toks[pos]
        pos += 1
        if t == "(":
            e = expr()
            assert toks[pos] == ")"
            pos += 1
            return e
        return off(t)

    e = expr()
    assert pos == len(toks), cond
    return e


def stencil(routine, names):
    src = open(os.path.join(REF, "src/lbm/lbm_forcing.F90")).read()
    body = src[src.index("subroutine %s(" % routine):src.index("end subroutine %s" % routine)]
    body = re.sub(r"&\s*\n\s*", " ", body)
    out, cur, section = [], None, 4
    for line in body.split("\n"):
        l = line.strip()
        if l.startswith("if(dist%disc%isotropy_order > 4)"):
            section = 8
            continue
        if l.startswith("if(dist%disc%isotropy_order > 8)"):
            section = 10
            continue
        m = re.match(r"if \((walls.*)\) then$", l)
        if m:
            cur = {"section": section, "cond": parse_cond(m.group(1), names), "grad": [], "wsum": []}
            continue
        if cur is None:
            continue
        mg = re.match(r"gradrho\(:,(\w)_DIRECTION,[^)]*\) = gradrho\(:,\w_DIRECTION,[^)]*\)\s*([+-])\s*"
                      r"(\d)?\.?\*?dist%disc%ffw\((\d+)\)\*\(rho\(:,([^)]*)\)-rho\(:,[^)]*\)\)", l)
        if mg:
            cur["grad"].append({"axis": "XYZ".index(mg.group(1)), "sign": 1 if mg.group(2) == "+" else -1,
                                "factor": int(mg.group(3) or 1), "k": int(mg.group(4)),
                                "target": [0 if p.strip() == v else int(p.strip()[len(v):])
                                           for p, v in zip(mg.group(5).split(","), names)]})
            continue
        mw = re.match(r"weightsum\((\w)_DIRECTION\) = weightsum\(\w_DIRECTION\) \+ dist%disc%ffw\((\d+)\)"
                      r"(\*(\d)\.)?", l)
        if mw:
            cur["wsum"].append({"axis": "XYZ".index(mw.group(1)), "k": int(mw.group(2)),
                                "mult": int(mw.group(4) or 1)})
            continue
        if l == "end if":
            out.append(cur)
            cur = None
    return out


if __name__ == "__main__":
    copy_fixture("tests/bubble_2D/reference_solution/fi001.dat", "bubble_2D_fi001.dat.xz")
    copy_fixture("tests/bubble_2D_hots/reference_solution/fi001.dat", "bubble_2D_hots_fi001.dat.xz")
    for name, routine, names in (("ff_stencil_d3.json", "LBMAddFluidFluidForcesD3", ("i", "j", "k")),
                                 ("ff_stencil_d2.json", "LBMAddFluidFluidForcesD2", ("i", "j"))):
        data = stencil(routine, names)
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print(name, len(data), "blocks")
