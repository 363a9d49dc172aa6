"""The Rust binding written out in the repository's integration guide
declares exactly the entry points of include/mole_gpu.h, and its repr(C)
structs list the header's fields in the header's order (the Rust cannot be
compiled in this image)."""
import os
import re

ROOT = os.path.join(os.path.dirname(__file__), "..")


def header_functions_and_structs():
    src = open(os.path.join(ROOT, "include", "mole_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    funcs = set(re.findall(r"^(?:mole_status|void|const char \*)\s*(mole_\w+)\s*\(", src, flags=re.M))
    structs = {}
    for body, name in re.findall(r"typedef struct \{(.*?)\}\s*(\w+);", src, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = re.sub(r"\[[^\]]*\]", "", decl).replace("*", " ").strip()
            if not decl:
                continue
            # "double sum_e, sum_e2": the first name ends the type, the rest follow commas
            parts = decl.split(",")
            fields.append(parts[0].split()[-1])
            fields.extend(p.strip() for p in parts[1:])
        structs[name] = fields
    return funcs, structs


def rust_block():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```rust\n(.*?)```", doc, flags=re.S)
    return next(b for b in blocks if "extern \"C\"" in b)


def test_rust_declarations_match_header_functions():
    funcs, _ = header_functions_and_structs()
    rust = set(re.findall(r"pub fn (mole_\w+)\(", rust_block()))
    assert rust == funcs, (sorted(funcs - rust), sorted(rust - funcs))


def test_rust_structs_match_header_fields():
    _, structs = header_functions_and_structs()
    rb = rust_block()
    for name in ("mole_system_desc", "mole_run_desc", "mole_stats", "mole_sr_sums", "mole_train_desc", "mole_optimizer", "mole_block_record", "mole_term", "mole_operator"):
        m = re.search(r"pub struct " + name + r" \{(.*?)\n\}", rb, flags=re.S)
        assert m, name
        rust_fields = re.findall(r"pub (\w+):", m.group(1))
        assert rust_fields == structs[name], (name, rust_fields, structs[name])
