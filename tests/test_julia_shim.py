"""CPU tier: every `ccall` in the Julia shim (julia/ITensorsROC.jl, the
binding a maintainer adds to ITensors.jl; Julia itself is not in the image)
names a symbol include/itg.h declares, with the same number of arguments,
argument types that agree with the C prototype, and as many values passed as
types declared."""
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "itg.h")
SHIM = os.path.join(ROOT, "itensorsgpu.jl_amd", "julia", "ITensorsROC.jl")

# Julia scalar types -> the C scalar type they must meet
SCALAR = {"Cint": "int", "Int64": "int64_t", "UInt64": "uint64_t", "Csize_t": "size_t",
          "Cdouble": "double", "Float64": "double", "Cfloat": "float", "Float32": "float",
          "Cchar": "char", "UInt8": "char"}
# element types of Ptr{...}/Ref{...} that pin the C pointee
POINTEE = {"Int64": "int64_t", "Int32": "int32_t", "Cint": "int", "Cdouble": "double",
           "Float64": "double", "Float32": "float", "UInt64": "uint64_t"}


def split_top(s):
    """Split at commas outside brackets."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def balanced(s, i):
    """Index just past the bracket group opening at s[i]."""
    depth = 0
    for j in range(i, len(s)):
        if s[j] in "({[":
            depth += 1
        elif s[j] in ")}]":
            depth -= 1
            if depth == 0:
                return j + 1
    raise ValueError("unbalanced at %d" % i)


def header_prototypes():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    text = "\n".join(ln for ln in text.splitlines() if not ln.lstrip().startswith("#"))
    protos = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(itg_\w+)\s*\(([^;]*?)\)\s*;", text):
        ret, name, params = m.group(1).strip(), m.group(2), m.group(3).strip()
        ps = []
        if params and params != "void":
            for p in split_top(params):
                p = p.replace("*", " * ").split()
                ps.append(" ".join(p[:-1]))   # drop the parameter name
        protos[name] = (" ".join(ret.replace("*", " * ").split()), ps)
    return protos


def shim_ccalls():
    text = open(SHIM).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(itg_\w+),\s*libitg\),\s*", text):
        name = m.group(1)
        i = m.end()
        # return type up to the next top-level comma
        j = i
        depth = 0
        while not (text[j] == "," and depth == 0):
            depth += text[j] in "({[" and 1 or 0
            depth -= text[j] in ")}]" and 1 or 0
            j += 1
        ret = text[i:j].strip()
        k = j + 1
        while text[k].isspace():
            k += 1
        assert text[k] == "(", name
        e = balanced(text, k)
        types = split_top(text[k + 1:e - 1])
        # passed values up to the ccall's closing parenthesis
        start = text.rfind("ccall(", 0, m.end())
        close = balanced(text, start + len("ccall"))
        rest = text[e:close - 1].strip()
        assert rest.startswith(","), name
        args = split_top(rest[1:])
        calls.append((name, ret, types, args))
    return calls


def c_kind(ctype):
    t = ctype.replace("const", "").replace("*", " * ").split()
    stars = t.count("*")
    base = " ".join(x for x in t if x != "*")
    return base, stars


def compatible(jl, ctype):
    base, stars = c_kind(ctype)
    m = re.fullmatch(r"(Ptr|Ref)\{(.+)\}", jl)
    if jl == "Cstring":
        return base == "char" and stars == 1
    if m:
        if stars == 0:
            return False
        inner = m.group(2)
        if re.fullmatch(r"(Ptr|Ref)\{.+\}", inner):     # pointer to pointer
            return stars == 2
        if inner in POINTEE and stars == 1:
            return base in (POINTEE[inner], "void")
        return stars == 1 or (inner == "Cvoid" and stars >= 1)
    if jl in SCALAR:
        return stars == 0 and base == SCALAR[jl]
    return False


def test_shim_has_ccalls_for_the_boundary():
    names = {c[0] for c in shim_ccalls()}
    for n in ["itg_init", "itg_contract", "itg_permute", "itg_add_permuted", "itg_nrm2",
              "itg_truncate", "itg_svd_trunc", "itg_eigh_trunc", "itg_qr", "itg_randn",
              "itg_heff2_apply", "itg_heff2_prepare", "itg_davidson2", "itg_memcpy_d2h"]:
        assert n in names, n


def test_every_ccall_matches_its_prototype():
    protos = header_prototypes()
    calls = shim_ccalls()
    assert len(calls) > 20
    for name, ret, types, args in calls:
        assert name in protos, name
        cret, cparams = protos[name]
        assert compatible(ret, cret), (name, ret, cret)
        assert len(types) == len(cparams), (name, types, cparams)
        for jt, ct in zip(types, cparams):
            assert compatible(jt, ct), (name, jt, ct)
        n_args = sum(1 for a in args if not a.endswith("..."))
        splat = [a for a in args if a.endswith("...")]
        if splat:
            # dims... stands for the six Heff dimensions
            assert all(a == "dims..." for a in splat), (name, splat)
            n_args += 6 * len(splat)
        assert n_args == len(types), (name, args, types)


def test_compatible_rejects_mismatches():
    assert compatible("Int64", "int64_t")
    assert not compatible("Cint", "int64_t")
    assert not compatible("Ptr{Int32}", "const int64_t*")
    assert compatible("Ptr{Int64}", "const int64_t*")
    assert compatible("Ref{Ptr{Cvoid}}", "itg_ctx**")
    assert not compatible("Ptr{Cvoid}", "int")
