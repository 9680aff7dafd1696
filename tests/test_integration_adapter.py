"""The C adapter of INTEGRATION.md against the two headers it is written
for.  mdoodz-private.h includes CHOLMOD and CSparse headers, which this image
does not have, so the adapter cannot be compiled here; instead every member
the adapter reads from the reference's structs (grid, params, mat_prop,
scale, markers, InterpBufPool, Nparams, SparseMat) must be declared in that
struct in the reference's headers, and every mdg_* call must name a
function of include/mdgpu.h with its number of arguments.  Skipped where the
reference sources are absent."""
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/MDLIB"
VARS = {"mesh": "grid", "model": "params", "materials": "mat_prop", "scaling": "scale",
        "particles": "markers", "pool": "InterpBufPool", "Nmodel": "Nparams", "StokesA": "SparseMat",
        "StokesC": "SparseMat", "StokesD": "SparseMat"}


def _strip_comments(src):
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
    return re.sub(r"//[^\n]*", " ", src)


def _struct_members(src):
    """{typedef name: set of member names} for `typedef struct ... { } name;`
    and for `typedef struct tag name;` with its `struct tag { };`."""
    bodies = [(m.group(1), m.group(2)) for m in
              re.finditer(r"typedef\s+struct\s*\w*\s*\{(.*?)\}\s*(\w+)\s*;", src, flags=re.S)]
    tags = {m.group(1): m.group(2) for m in re.finditer(r"typedef\s+struct\s+(\w+)\s+(\w+)\s*;", src)}
    for m in re.finditer(r"(?<!typedef )struct\s+(\w+)\s*\{(.*?)\}\s*;", src, flags=re.S):
        if m.group(1) in tags:
            bodies.append((m.group(2), tags[m.group(1)]))
    out = {}
    for body, name in bodies:
        names = set()
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            # drop the type: everything up to the first declarator
            parts = decl.split(",")
            first = re.match(r"^(?:const\s+)?(?:unsigned\s+|signed\s+|struct\s+)?[\w]+[\s\*]+(.*)$", parts[0])
            if not first:
                continue
            for d in [first.group(1)] + parts[1:]:
                d = re.sub(r"\[.*?\]", "", d).replace("*", "").strip()
                d = d.split("=")[0].strip()
                if re.fullmatch(r"\w+", d):
                    names.add(d)
        out[name] = names
    return out


def _c_blocks():
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    return "\n".join(re.findall(r"```c\n(.*?)```", text, flags=re.S))


def _prototypes():
    src = _strip_comments(open(os.path.join(REPO, "include", "mdgpu.h")).read())
    protos = {}
    for m in re.finditer(r"\b(mdg_\w+)\s*\(([^;{]*?)\)\s*;", src, flags=re.S):
        args = m.group(2).strip()
        protos[m.group(1)] = 0 if args in ("", "void") else _top_level_commas(args) + 1
    return protos


def _top_level_commas(s):
    depth, n = 0, 0
    for ch in s:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        elif ch == "," and depth == 0:
            n += 1
    return n


def _calls(code, name):
    """Argument lists of every call of name in code."""
    out = []
    for m in re.finditer(r"\b%s\s*\(" % re.escape(name), code):
        i, depth = m.end(), 1
        while depth:
            depth += {"(": 1, ")": -1}.get(code[i], 0)
            i += 1
        out.append(code[m.end():i - 1])
    return out


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
def test_adapter_reads_only_declared_members():
    src = _strip_comments(open(os.path.join(REF, "include", "mdoodz.h")).read() + "\n" +
                          open(os.path.join(REF, "mdoodz-private.h")).read())
    members = _struct_members(src)
    for t in set(VARS.values()):
        assert t in members and members[t], "struct %s not found in the reference headers" % t
    code = _c_blocks()
    used = set(re.findall(r"\b(%s)\s*(?:->|\.)\s*(\w+)" % "|".join(VARS), code))
    assert len(used) > 40
    missing = sorted("%s->%s (%s)" % (v, f, VARS[v]) for v, f in used if f not in members[VARS[v]])
    assert not missing, missing


def test_adapter_calls_match_the_header():
    protos = _prototypes()
    code = _c_blocks()
    called = set(re.findall(r"\b(mdg_\w+)\s*\(", code))
    assert len(called) > 15
    for name in sorted(called):
        assert name in protos, "%s is not declared in include/mdgpu.h" % name
        for args in _calls(code, name):
            n = 0 if not args.strip() else _top_level_commas(args) + 1
            assert n == protos[name], "%s(%s): %d arguments, the header declares %d" % (name, args, n, protos[name])
