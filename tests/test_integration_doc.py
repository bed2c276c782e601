"""The Java binding shown in INTEGRATION.md against the C ABI it binds (no GPU).

Every function of include/mvr_hip.h must be declared in the JNA interface (MvrHipLibrary) with the same number of
arguments and a return type that JNA maps to the C one; the interface must declare nothing else.  The Java
classes of the document may call only functions the header declares, each with the header's argument count.
"""
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "mvr_hip.h")
DOC = os.path.join(REPO, "INTEGRATION.md")

JAVA_RETURN = {"int": "int", "long long": "long", "void": "void", "const char*": "String"}


def split_args(s):
    """Top-level comma split (no nesting in these signatures beyond brackets)."""
    s = " ".join(s.split())
    if s in ("", "void"):
        return []
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([<":
            depth += 1
        elif ch in ")]>":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    out.append(cur.strip())
    return out


def header_signatures():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    sigs = {}
    for m in re.finditer(r"^\s*((?:const\s+)?[a-z_ ]+?\s*\**)\s*(mvr_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text, flags=re.M):
        ret = " ".join(m.group(1).split()).replace(" *", "*")
        sigs[m.group(2)] = (ret, len(split_args(m.group(3))))
    return sigs


def java_blocks():
    return re.findall(r"```java\n(.*?)```", open(DOC).read(), flags=re.S)


def interface_signatures():
    blocks = [b for b in java_blocks() if "interface MvrHipLibrary" in b]
    assert len(blocks) == 1, "INTEGRATION.md must show the JNA interface once"
    body = re.sub(r"//[^\n]*", "", blocks[0])
    sigs = {}
    for m in re.finditer(r"\b(int|long|void|String)\s+(mvr_[a-z0-9_]+)\s*\(([^;]*?)\)\s*;", body, flags=re.S):
        assert m.group(2) not in sigs, f"{m.group(2)} declared twice"
        sigs[m.group(2)] = (m.group(1), len(split_args(m.group(3))))
    return sigs


def test_header_parses():
    sigs = header_signatures()
    assert len(sigs) >= 60
    assert sigs["mvr_decon_run"] == ("int", 2)
    assert sigs["mvr_get_num_devices"] == ("int", 0)
    assert sigs["mvr_last_error"] == ("const char*", 0)


def test_jna_interface_declares_every_header_function():
    h = header_signatures()
    j = interface_signatures()
    missing = sorted(set(h) - set(j))
    extra = sorted(set(j) - set(h))
    assert not missing, f"not in the JNA interface: {missing}"
    assert not extra, f"not in the header: {extra}"
    for name, (ret, n) in h.items():
        jret, jn = j[name]
        assert jn == n, f"{name}: {jn} arguments in Java, {n} in C"
        assert JAVA_RETURN[ret] == jret, f"{name}: returns {jret} in Java, {ret} in C"


def test_java_classes_call_the_abi_with_its_argument_counts():
    h = header_signatures()
    calls = 0
    for block in java_blocks():
        if "interface MvrHipLibrary" in block:
            continue
        for m in re.finditer(r"\.\s*(mvr_[a-z0-9_]+)\s*\(", block):
            name = m.group(1)
            assert name in h, f"the Java code calls {name}, which the header does not declare"
            # the argument list up to the matching parenthesis
            i, depth = m.end(), 1
            while depth:
                depth += {"(": 1, ")": -1}.get(block[i], 0)
                i += 1
            n = len(split_args(block[m.end():i - 1]))
            assert n == h[name][1], f"{name} called with {n} arguments, the header has {h[name][1]}"
            calls += 1
    assert calls >= 20


@pytest.mark.parametrize("name", ["mvr_decon_create", "mvr_decon_set_view", "mvr_decon_set_view_kernels",
                                  "mvr_decon_init", "mvr_decon_set_max", "mvr_decon_run", "mvr_decon_get_psi",
                                  "mvr_decon_get_stats", "mvr_decon_destroy"])
def test_resident_driver_uses_the_resident_entry_points(name):
    # MultiViewDeconvolutionHIP: views uploaded once, runNextIteration -> mvr_decon_run, getPSI <- mvr_decon_get_psi
    block = next(b for b in java_blocks() if "class MultiViewDeconvolutionHIP" in b)
    assert re.search(r"\." + name + r"\s*\(", block), name
