"""The Julia binding (gridapdistributed.jl_amd/julia/GridapDistributedHIP.jl)
against the C ABI it calls (include/gdx.h).  No Julia toolchain exists here,
so the shim is read as text: its mirrored structs must have the C structs'
fields in the same order with matching types, and every ccall must name a
declared entry point with the declared number and kinds of arguments and an
integer status return.  CPU only."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gdx.h")
SHIM = os.path.join(ROOT, "gridapdistributed.jl_amd", "julia", "GridapDistributedHIP.jl")

C_SCALARS = {"int32_t": "Int32", "int64_t": "Int64", "double": "Float64", "uint64_t": "UInt64", "float": "Float32"}


def strip_c_comments(s):
    s = re.sub(r"/\*.*?\*/", " ", s, flags=re.S)
    return re.sub(r"//[^\n]*", " ", s)


def c_structs():
    src = strip_c_comments(open(HEADER).read())
    out = {}
    for m in re.finditer(r"typedef\s+struct\s+(\w+)\s*\{(.*?)\}\s*(\w+)\s*;", src, flags=re.S):
        fields = []
        for decl in m.group(2).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            # "const double *modes", "double flux_p, flux_alpha", "int64_t cells[3]", "gdx_nn *precond"
            mt = re.match(r"((?:const\s+)?\w+)\s*(\*?)\s*(.*)", decl, flags=re.S)
            base, star, names = mt.group(1).replace("const ", "").strip(), mt.group(2), mt.group(3)
            for nm in names.split(","):
                nm = nm.strip()
                ptr = star or nm.startswith("*")
                nm = nm.lstrip("*").strip()
                arr = re.match(r"(\w+)\s*\[(\d+)\]", nm)
                if ptr:
                    t = "Ptr{Float64}" if base == "double" else "Ptr{Cvoid}"
                else:
                    t = C_SCALARS[base]
                if arr:
                    fields.append((arr.group(1), f"NTuple{{{arr.group(2)},{t}}}"))
                else:
                    fields.append((nm, t))
        out[m.group(3)] = fields
    return out


def julia_structs():
    src = open(SHIM).read()
    out = {}
    for m in re.finditer(r"^struct\s+(Gdx\w+)\s*\n(.*?)^end", src, flags=re.S | re.M):
        fields = []
        for line in m.group(2).splitlines():
            line = line.split("#", 1)[0]
            for part in line.split(";"):
                part = part.strip()
                if part:
                    nm, t = part.split("::")
                    fields.append((nm.strip(), t.strip().replace(" ", "")))
        out[m.group(1)] = fields
    return out


PAIRS = [("GdxProblemDesc", "gdx_problem_desc"), ("GdxNewtonOpts", "gdx_newton_opts"),
         ("GdxNewtonReport", "gdx_newton_report")]


@pytest.mark.parametrize("jl,c", PAIRS)
def test_struct_layout_matches_header(jl, c):
    C, J = c_structs(), julia_structs()
    assert jl in J and c in C
    assert J[jl] == C[c], (J[jl], C[c])


def c_prototypes():
    src = strip_c_comments(open(HEADER).read())
    protos = {}
    for m in re.finditer(r"\b(int|const char\s*\*)\s*(gdx_\w+)\s*\(([^)]*)\)\s*;", src):
        args = [a.strip() for a in m.group(3).split(",") if a.strip() and a.strip() != "void"]
        protos[m.group(2)] = (m.group(1).replace(" ", ""), args)
    return protos


def c_arg_kind(a):
    if "*" in a:
        return "ptr"
    base = a.split()[0]
    return C_SCALARS.get(base, base)


def julia_arg_kind(t):
    t = t.strip()
    if t.startswith(("Ptr{", "Ref{", "Cstring")):
        return "ptr"
    return t


def split_top(s):
    """split a Julia tuple body at top-level commas"""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return [x.strip() for x in out if x.strip()]


def julia_ccalls():
    src = open(SHIM).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(gdx_\w+),\s*libgdx\),\s*(\w+),\s*\(", src):
        # the argument-type tuple starts at m.end(); find its closing parenthesis
        depth, i = 1, m.end()
        while depth:
            depth += {"(": 1, ")": -1}.get(src[i], 0)
            i += 1
        calls.append((m.group(1), m.group(2), split_top(src[m.end():i - 1])))
    return calls


def test_ccalls_match_prototypes():
    protos, calls = c_prototypes(), julia_ccalls()
    assert len(calls) >= 20
    for name, ret, args in calls:
        assert name in protos, name
        cret, cargs = protos[name]
        assert (ret == "Cint") == (cret == "int") and (ret == "Cstring") == (cret == "constchar*"), name
        assert len(args) == len(cargs), (name, args, cargs)
        for ja, ca in zip(args, cargs):
            assert julia_arg_kind(ja) == c_arg_kind(ca), (name, ja, ca)


def test_shim_defines_the_assembler_interface():
    src = open(SHIM).read()
    assert re.search(r"struct HIPSparseMatrixAssembler\{[^}]*\}\s*<:\s*SparseMatrixAssembler", src)
    for f in ("get_rows", "get_cols", "get_assembly_strategy", "get_matrix_builder", "get_vector_builder",
              "assemble_matrix_and_vector", "assemble_matrix_and_vector!", "allocate_matrix_and_vector",
              "assemble_matrix", "assemble_vector", "assemble_matrix!", "assemble_vector!", "allocate_matrix",
              "allocate_vector", "collect_cell_matrix_and_vector"):
        assert re.search(rf"FESpaces\.{re.escape(f)}\(", src), f
    # the lowering returns nothing outside the vocabulary, and both local matrix types are built
    assert "return nothing" in src and "local_matrix(::Type{<:SparseMatrixCSC}" in src
    assert "local_matrix(::Type{<:SparseMatrixCSR}" in src
