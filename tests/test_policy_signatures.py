"""The drop-in policies' hooks against the reference's own declarations.

The reference's policy headers need kv (interval, dd, hwround, rdouble),
which is not in this image, so include/vcp/pidhip.hpp, pddhip.hpp and
pidhip_fma.hpp cannot be compiled against vcp/ here, and this project builds
no reference code on stand-ins for what the image lacks.  These tests read
the reference's headers as text instead and check what a compile would:

  * every `override` in our headers has a declaration of the same name,
    parameter types (names and default values aside) and const qualification
    in the base policy, marked virtual or override there
    (vcp/pidblas.hpp:90,142,198,239,296; vcp/pddblas.hpp:477,490,567);
  * pddhip's hiding methods name methods of pddblas that are not virtual
    (vcp/pddblas.hpp:318,379,413,571,611,615), with pddhip in place of pddblas
    in the product's operands;
  * pidhip_fma's base derives from pidblas, so the pidblas hooks stay hooks;
  * every member the headers call on their bases, and every error class they
    throw, is declared by the reference, the backend error with the
    constructor arguments used (vcp/error.hpp:82-95).

Skipped where the reference tree is absent (these run in the build
container, never on a GPU box)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/vcp"
OURS = os.path.join(ROOT, "include", "vcp")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree absent")

DECL = re.compile(r"(virtual\s+)?void\s+(\w+)\s*\(([^()]*)\)\s*(const)?\s*(override)?\s*\{")


def _split_top(plist):
    out, depth, cur = [], 0, ""
    for ch in plist:
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return out


def _param_types(plist, rename=None):
    types = []
    for p in _split_top(plist):
        p = p.split("=")[0].strip()
        m = re.match(r"^(.*[&*>\s])(\w+)$", p)
        t = m.group(1) if m else p
        t = re.sub(r"\s+", "", t).replace("vcp::", "")
        for a, b in (rename or {}).items():
            t = t.replace(a, b)
        types.append(t)
    return tuple(types)


def _decls(path, rename=None):
    with open(path) as f:
        text = f.read()
    out = []
    for m in DECL.finditer(text):
        out.append({"virtual": bool(m.group(1)), "name": m.group(2), "params": _param_types(m.group(3), rename),
                    "const": bool(m.group(4)), "override": bool(m.group(5))})
    return out


def _find(decls, name, params, const):
    return [d for d in decls if d["name"] == name and d["params"] == params and d["const"] == const]


def test_pidhip_overrides_match_pidblas_hooks():
    ours = [d for d in _decls(os.path.join(OURS, "pidhip.hpp")) if d["override"]]
    names = sorted(d["name"] for d in ours)
    assert names == sorted(["mul_im_m", "mul_m_im", "vmulmm", "mulmm", "mulltmm"])
    ref = _decls(os.path.join(REF, "pidblas.hpp"))
    for d in ours:
        hits = _find(ref, d["name"], d["params"], d["const"])
        assert hits and all(h["virtual"] or h["override"] for h in hits), d


def test_pidhip_fma_base_keeps_the_hooks():
    with open(os.path.join(REF, "pidblas_fma.hpp")) as f:
        text = f.read()
    assert re.search(r"class\s+pidblas_fma\s*:\s*public\s+(vcp::)?pidblas\b", text)
    ref = _decls(os.path.join(REF, "pidblas_fma.hpp"))
    ours = {d["name"]: d for d in _decls(os.path.join(OURS, "pidhip.hpp")) if d["override"]}
    mm = _find(ref, "mulmm", ours["mulmm"]["params"], True)
    assert mm and mm[0]["override"]
    with open(os.path.join(OURS, "pidhip_fma.hpp")) as f:
        assert "pidhip_t<vcp::pidblas_fma, VCPHIP_FORM_INFSUP>" in f.read()


def test_pddhip_overrides_and_hiding_methods_match_pddblas():
    ours = _decls(os.path.join(OURS, "pddhip.hpp"), rename={"pddhip": "pddblas"})
    ref = _decls(os.path.join(REF, "pddblas.hpp"))
    over = [d for d in ours if d["override"]]
    assert sorted(d["name"] for d in over) == ["Cholesky", "eigsym", "inv"]
    for d in over:
        hits = _find(ref, d["name"], d["params"], d["const"])
        assert hits and all(h["override"] or h["virtual"] for h in hits), d
    hiding = [d for d in ours if not d["override"] and d["name"] in
              ("mulmm", "mulltmm", "linearsolve", "eigsym", "eigsymge")]
    assert sorted({d["name"] for d in hiding}) == ["eigsym", "eigsymge", "linearsolve", "mulltmm", "mulmm"]
    for d in hiding:
        hits = _find(ref, d["name"], d["params"], d["const"])
        assert hits, d
        assert not any(h["virtual"] or h["override"] for h in hits), d


def test_members_and_errors_used_are_declared_by_the_reference():
    ref_text = ""
    for name in ("mats.hpp", "imats.hpp", "pdblas.hpp", "pddblas.hpp", "pidblas.hpp", "pidblas_fma.hpp"):
        with open(os.path.join(REF, name)) as f:
            ref_text += f.read()
    for name in ("pidhip.hpp", "pddhip.hpp"):
        with open(os.path.join(OURS, name)) as f:
            ours = f.read()
        used = set(re.findall(r"this->(\w+)\s*\(", ours)) | set(re.findall(r"(?:Base|vcp::pddblas)::(\w+)\s*\(", ours))
        assert used, name
        for member in sorted(used):
            assert re.search(r"\b%s\s*\(" % re.escape(member), ref_text), (name, member)
        # data members read through this-> (row, column, type, v, n)
        for field in set(re.findall(r"this->(\w+)\b(?!\s*\()", ours)):
            assert re.search(r"\b%s\b" % re.escape(field), ref_text), (name, field)
    with open(os.path.join(REF, "error.hpp")) as f:
        err = f.read()
    for name in ("pidhip.hpp", "pddhip.hpp"):
        with open(os.path.join(OURS, name)) as f:
            ours = f.read()
        for cls in set(re.findall(r"vcp::(?!throw_error)(\w+_error|invalid_argument)\b", ours)):
            assert re.search(r"class\s+%s\s*:" % cls, err), (name, cls)
        if "vcp::throw_error<" in ours:
            assert re.search(r"template\s*<class Exception, class\.\.\. Args>\s*\[\[noreturn\]\]\s*void\s+throw_error|"
                             r"void\s+throw_error\s*\(", err)
    assert re.search(r"backend_error\(const std::string& backend, const std::string& routine, int code, "
                     r"const std::string& message\)", err)
