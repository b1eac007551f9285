"""Every prototype in include/vsip.h against the reference declaration it cites.

Each function declared in include/vsip.h carries a trailing `/* vsip.h:NNN */`
comment naming the line of the reference header c_VSIP_src/vsip.h it stands
for.  This test reads that line and compares return type, name and parameter
types (parameter names and white space ignored), so a drop-in caller that
compiles against the reference header compiles against this one.  It reads the
reference's header as text and is skipped where the reference is absent.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_HEADER = "/root/reference/c_VSIP_src/vsip.h"

_TYPE_WORDS = {"const", "void", "int", "unsigned", "signed", "long", "char", "float", "double", "short"}


def _norm_param(p):
    p = p.strip()
    toks = re.findall(r"[A-Za-z_]\w*|\*", p)
    # drop a trailing parameter name (anything that is neither a keyword nor a vsip_ type)
    if len(toks) > 1 and toks[-1] != "*" and toks[-1] not in _TYPE_WORDS and not toks[-1].startswith("vsip_"):
        toks = toks[:-1]
    return " ".join(toks).replace(" *", "*")


def _norm_proto(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    m = re.match(r"\s*(.*?)\b(vsip_\w+)\s*\((.*)\)\s*;?\s*$", text, flags=re.S)
    assert m, text
    ret = " ".join(re.findall(r"[A-Za-z_]\w*|\*", m.group(1))).replace(" *", "*")
    params = [_norm_param(p) for p in m.group(3).split(",")]
    if params == ["void"]:
        params = []
    return ret, m.group(2), params


def _our_prototypes():
    text = open(os.path.join(ROOT, "include", "vsip.h")).read()
    body = text[text.index("/* ---- library lifecycle"):]
    out = []
    # a statement up to ';' and the citation comment that follows it on the same line
    for m in re.finditer(r"([^;{}]*?\bvsip_\w+\s*\([^;]*\)\s*;)[ \t]*/\*\s*vsip\.h:(\d+)\s*\*/", body, flags=re.S):
        stmt = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
        if "typedef" in stmt:
            continue
        out.append((stmt.strip(), int(m.group(2))))
    return out


@pytest.mark.skipif(not os.path.exists(REF_HEADER), reason="reference sources not present")
def test_every_prototype_matches_the_cited_reference_line():
    ref_lines = open(REF_HEADER, errors="replace").read().split("\n")
    ours = _our_prototypes()
    assert len(ours) > 190
    bad = []
    for stmt, line in ours:
        mine = _norm_proto(stmt)
        ref = ref_lines[line - 1]
        if mine[1] not in ref:
            bad.append("%s cites vsip.h:%d which declares %r" % (mine[1], line, ref.strip()))
            continue
        theirs = _norm_proto(ref)
        if mine != theirs:
            bad.append("%s: ours %r, reference vsip.h:%d %r" % (mine[1], mine, line, theirs))
    assert not bad, "\n".join(bad)


def test_every_declared_function_cites_a_reference_line():
    text = open(os.path.join(ROOT, "include", "vsip.h")).read()
    body = re.sub(r"/\*(?!\s*vsip\.h:).*?\*/", "", text[text.index("/* ---- library lifecycle"):], flags=re.S)
    names = set(re.findall(r"\b(vsip_\w+)\s*\(", body))
    cited = set(_norm_proto(s)[1] for s, _ in _our_prototypes())
    assert names == cited, sorted(names ^ cited)


def test_planted_mismatch_is_caught():
    # a pointer return in place of void must differ after normalisation
    a = _norm_proto("vsip_cmattr_f* vsip_cmgetattrib_f(const vsip_cmview_f*, vsip_cmattr_f*);")
    b = _norm_proto("void vsip_cmgetattrib_f ( const vsip_cmview_f*, vsip_cmattr_f* ) ;")
    assert a != b
    # parameter names do not matter, types do
    assert _norm_proto("int f_x(vsip_length M, vsip_scalar_f* const x);".replace("f_x", "vsip_x")) == \
        _norm_proto("int vsip_x ( vsip_length, vsip_scalar_f* const ) ;")
    assert _norm_proto("int vsip_x(vsip_length);") != _norm_proto("int vsip_x(vsip_offset);")
