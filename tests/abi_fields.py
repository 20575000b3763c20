"""Ordered field names of C struct definitions, for comparing this library's public header with
the reference's function-table headers (the tables are the drop-in ABI: same fields, same order).

A field is named as declared: `ret(PNANOVDB_ABI* name)(...)` and `ret(*name)(...)` give `name`,
`PNANOVDB_REFLECT_INTERFACE()` gives `interface_pnanovdb_reflect_data_type` (the field the macro
declares), other declarations their last identifier. Fields inside `#if`/`#ifdef` blocks are
reported separately as conditional."""
import re


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _split_decls(body):
    decls, depth, cur = [], 0, []
    for ch in body:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == ";" and depth == 0:
            decls.append("".join(cur))
            cur = []
        else:
            cur.append(ch)
    return [d.strip() for d in decls if d.strip()]


def _field(decl):
    """(name, type): type is "fn" for a function pointer, else the declaration without the name."""
    decl = " ".join(decl.split())
    if decl.startswith("PNANOVDB_REFLECT_INTERFACE("):
        return "interface_pnanovdb_reflect_data_type", "const pnanovdb_reflect_data_type_t*"
    m = re.search(r"\(\s*(?:PNANOVDB_ABI\s*)?\*\s*(\w+)\s*\)", decl)
    if m:
        return m.group(1), "fn"
    decl = re.sub(r"\[[^\]]*\]", "", decl)
    name = re.findall(r"\w+", decl)[-1]
    typ = decl[:decl.rindex(name)].strip()
    return name, re.sub(r"\s*\*", "*", typ)


def struct_fields(text, names):
    """{struct name: {"fields": [...], "conditional": [...]}} for each name defined in text."""
    text = _strip_comments(text)
    out = {}
    for name in names:
        for m in re.finditer(r"\bstruct\s+" + re.escape(name) + r"\s*\{", text):
            i = m.end()
            depth = 1
            j = i
            while depth:
                if text[j] == "{":
                    depth += 1
                elif text[j] == "}":
                    depth -= 1
                j += 1
            body = text[i:j - 1]
            plain, cond = [], []
            level = 0
            for line in body.split("\n"):
                s = line.strip()
                if s.startswith("#if"):
                    level += 1
                    continue
                if s.startswith("#endif"):
                    level -= 1
                    continue
                if s.startswith("#"):
                    continue
                (cond if level else plain).append(line)
            fields = [_field(d) for d in _split_decls("\n".join(plain))]
            out[name] = {"fields": [f[0] for f in fields], "types": [f[1] for f in fields],
                         "conditional": [_field(d)[0] for d in _split_decls("\n".join(cond))]}
            break
    return out


# the reference headers (nanovdb_editor/putil/) and the tables each defines
REFERENCE_TABLES = {
    "Compute.h": ["pnanovdb_compute_array_t", "pnanovdb_compute_profiler_entry_t",
                  "pnanovdb_compute_shader_interface_t", "pnanovdb_compute_device_interface_t", "pnanovdb_compute_t"],
    "Raster.h": ["pnanovdb_raster_t"],
    "ParallelPrimitives.h": ["pnanovdb_parallel_primitives_t"],
    "GridBuild.h": ["pnanovdb_grid_build_state_t", "pnanovdb_grid_build_fanout_state_t", "pnanovdb_grid_build_t"],
    "FileFormat.h": ["pnanovdb_fileformat_t"],
}
