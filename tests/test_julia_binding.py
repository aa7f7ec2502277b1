"""The Julia binding (culamp.jl_amd/julia/CuLAMPAMD.jl) against the C header.

Julia is not installed here, so the module cannot be loaded; instead every
`ccall` in it is parsed and its argument-type tuple, argument count and return
type are checked against the prototype of the same symbol in include/culamp.h,
and every Julia struct passed by reference is checked field by field against
the C struct it mirrors.  A wrong argument order, a missing argument or a
drifted struct layout fails here."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "culamp.h")
JULIA = os.path.join(ROOT, "culamp.jl_amd", "julia", "CuLAMPAMD.jl")

HANDLES = {"culamp_ctx", "culamp_problem", "culamp_batch", "culamp_graph", "culamp_dense"}
POINTEE = {"culamp_args": "CArgs", "culamp_result": "CResult", "culamp_instance": "CInstance",
           "culamp_batch_tuning": "CTuning", "culamp_divergence_opts": "DivOpts"}
SCALAR = {"int": {"Cint"}, "int32_t": {"Int32", "Cint"}, "int64_t": {"Int64"},
          "double": {"Float64"}, "size_t": {"Csize_t"}, "culamp_log_fn": {"Ptr{Cvoid}"}}
STRUCT_PAIRS = {"CArgs": "culamp_args", "CResult": "culamp_result", "CInstance": "culamp_instance",
                "CTuning": "culamp_batch_tuning", "DivOpts": "culamp_divergence_opts"}
C_FIELD = {"double": "Float64", "int64_t": "Int64", "int32_t": "Int32",
           "const double *": "Ptr{Float64}"}


def _strip_c_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _header():
    return _strip_c_comments(open(HEADER).read())


def c_prototypes():
    """symbol -> (return type, [parameter declarations])"""
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w \*]*?)\s*\b(culamp_\w+)\s*\(([^;{}]*?)\)\s*;", _header()):
        ret, name, params = m.group(1).strip(), m.group(2), m.group(3)
        ret = re.sub(r"\s+", " ", ret)
        params = [re.sub(r"\s+", " ", p).strip() for p in params.split(",")]
        if params == ["void"]:
            params = []
        out[name] = (ret, params)
    return out


def c_structs():
    """typedef name -> [(C type, field name)]"""
    out = {}
    for m in re.finditer(r"typedef struct\s*\{(.*?)\}\s*(\w+)\s*;", _header(), flags=re.S):
        fields = []
        for decl in m.group(1).split(";"):
            decl = re.sub(r"\s+", " ", decl).strip()
            if not decl:
                continue
            mm = re.match(r"((?:const )?\w+)\s*(\*?)\s*(.*)", decl)
            base, star, names = mm.group(1), mm.group(2), mm.group(3)
            for nm in names.split(","):
                nm = nm.strip()
                if nm.startswith("*"):
                    star, nm = "*", nm[1:].strip()
                ctype = base + (" *" if star else "")
                fields.append((ctype, nm))
        out[m.group(2)] = fields
    return out


def _c_enum(name):
    m = re.search(r"\b%s\s*=\s*(\d+)" % name, _header())
    return int(m.group(1))


def _balanced(text, i, open_ch="(", close_ch=")"):
    """index one past the bracket matching text[i] == open_ch"""
    assert text[i] == open_ch
    depth = 0
    for j in range(i, len(text)):
        if text[j] in "([{":
            depth += 1
        elif text[j] in ")]}":
            depth -= 1
            if depth == 0:
                return j + 1
    raise AssertionError("unbalanced at %d" % i)


def _split_top(text):
    parts, depth, cur = [], 0, ""
    for ch in text:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur.strip())
    return parts


def _strip_julia_comments(text):
    return "\n".join(re.sub(r"#(?![=]).*$", "", line) for line in text.splitlines())


def julia_ccalls():
    """[(symbol, return type, [argument types], [arguments])] for every ccall"""
    src = _strip_julia_comments(open(JULIA).read())
    calls = []
    for m in re.finditer(r"ccall\(", src):
        start = m.end() - 1
        end = _balanced(src, start)
        body = src[start + 1:end - 1]
        items = _split_top(body)
        sym = re.match(r"\(:(\w+)\s*,\s*libculamp\)", items[0])
        assert sym, items[0]
        ret = items[1]
        tup = items[2]
        assert tup.startswith("(") and tup.endswith(")"), tup
        types = _split_top(tup[1:-1])
        calls.append((sym.group(1), ret, types, items[3:]))
    return calls


def julia_structs():
    src = _strip_julia_comments(open(JULIA).read())
    out = {}
    for m in re.finditer(r"^(?:Base\.@kwdef )?(?:mutable )?struct (\w+)\n(.*?)^end", src,
                         flags=re.S | re.M):
        fields = []
        for line in m.group(2).splitlines():
            f = re.match(r"\s*(\w+)::([\w\{\}]+)", line)
            if f:
                fields.append((f.group(2), f.group(1)))
        out[m.group(1)] = fields
    return out


def _accepted(param):
    """Julia argument types that pass the C parameter `param` correctly."""
    p = param.replace("const ", "")
    depth = p.count("*")
    base = p.replace("*", " ").split()[0]
    if depth == 2:
        return {"Ref{Ptr{Cvoid}}"}
    if depth == 1:
        if base in HANDLES:
            return {"Ptr{Cvoid}"}
        if base == "void":
            return {"Ptr{Cvoid}", "Ptr{Float64}"}
        if base == "double":
            ok = {"Ptr{Float64}"}
            if not param.startswith("const"):
                ok.add("Ref{Float64}")
            return ok
        if base == "int32_t":
            return {"Ref{Int32}", "Ptr{Int32}"}
        if base in POINTEE:
            return {"Ref{%s}" % POINTEE[base], "Ptr{%s}" % POINTEE[base]}
        raise AssertionError("unmapped pointer parameter " + param)
    return SCALAR[base]


RETURNS = {"int": "Cint", "void": "Cvoid", "const char *": "Cstring"}


def test_every_ccall_matches_its_prototype():
    protos = c_prototypes()
    calls = julia_ccalls()
    assert len(calls) >= 40
    for sym, ret, types, args in calls:
        assert sym in protos, "%s is not declared in include/culamp.h" % sym
        cret, params = protos[sym]
        assert RETURNS[cret] == ret, (sym, cret, ret)
        assert len(types) == len(params), (sym, types, params)
        assert len(args) == len(types), (sym, "argument count", args)
        for k, (jt, cp) in enumerate(zip(types, params)):
            assert jt in _accepted(cp), "%s arg %d: %s does not pass %r" % (sym, k, jt, cp)


def test_structs_match_their_c_layout():
    cs, js = c_structs(), julia_structs()
    for jname, cname in STRUCT_PAIRS.items():
        cfields = cs[cname]
        jfields = js[jname]
        assert [n for _, n in jfields] == [n for _, n in cfields], jname
        assert [t for t, _ in jfields] == [C_FIELD[t] for t, _ in cfields], jname


def test_constants_match_the_header():
    src = open(JULIA).read()
    for jl, c in (("CULAMP_FP64", "CULAMP_FP64"), ("CULAMP_MIXED", "CULAMP_MIXED"),
                  ("CULAMP_BF16", "CULAMP_BF16"), ("CULAMP_BF16X3", "CULAMP_BF16X3")):
        m = re.search(r"const %s = Int32\((\d+)\)" % jl, src)
        assert int(m.group(1)) == _c_enum(c), jl
    metric = dict(re.findall(r":(\w+) => Int32\((\d+)\)", src))
    assert int(metric["l1"]) == _c_enum("CULAMP_METRIC_L1")
    assert int(metric["l2"]) == _c_enum("CULAMP_METRIC_L2")
    assert int(metric["cosine"]) == _c_enum("CULAMP_METRIC_COSINE")
    assert int(metric["pearson"]) == _c_enum("CULAMP_METRIC_CORRELATION")
    stop = dict(re.findall(r'"(reference|infeas)" => Int32\((\d+)\)', src))
    assert int(stop["reference"]) == _c_enum("CULAMP_STOP_REFERENCE")
    assert int(stop["infeas"]) == _c_enum("CULAMP_STOP_INFEAS")


def test_every_solver_entry_point_is_bound():
    bound = {sym for sym, *_ in julia_ccalls()}
    for sym in ("culamp_lamp_solve", "culamp_sinkhorn_solve", "culamp_annealed_sinkhorn_solve",
                "culamp_lamp_cellsim_solve", "culamp_sinkhorn_cellsim_solve",
                "culamp_annealed_sinkhorn_cellsim_solve", "culamp_batch_lamp_solve",
                "culamp_batch_sinkhorn_solve", "culamp_sinkhorn_divergence",
                "culamp_dense_lamp_solve", "culamp_lse_rows", "culamp_col_marginal_cost",
                "culamp_sinkhorn_half_step", "culamp_row_residual_cost"):
        assert sym in bound, sym


def test_parser_catches_a_wrong_binding():
    """The checks above are not vacuous: a swapped pair of arguments, a
    dropped argument and a mistyped struct field are all rejected."""
    protos = c_prototypes()
    _, params = protos["culamp_lse_rows"]
    good = ["Ptr{Cvoid}", "Ptr{Float64}", "Int64", "Ptr{Float64}", "Int64", "Ptr{Float64}",
            "Float64", "Float64", "Float64", "Float64", "Cint", "Ptr{Float64}"]
    assert all(t in _accepted(p) for t, p in zip(good, params))
    swapped = list(good)
    swapped[1], swapped[2] = swapped[2], swapped[1]
    assert not all(t in _accepted(p) for t, p in zip(swapped, params))
    assert len(good[:-1]) != len(params)
    cfields = c_structs()["culamp_result"]
    assert C_FIELD[cfields[-1][0]] == "Int64"   # fp64_from is 64-bit, not Cint
