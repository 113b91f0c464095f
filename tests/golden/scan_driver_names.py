"""Writes driver_names.json: for each module of the reference package that the
drop-in replaces, the names the rest of the reference package reads from it (by
attribute or `from ... import`), with the parameter list of each that the module
defines as a function.  Reads the reference's sources as text (ast), imports nothing
from it.  Run in a container that holds the reference checkout:

    python tests/golden/scan_driver_names.py /root/reference/sttopt
"""

import ast
import json
import pathlib
import sys

ALIASED = ["torch_fem", "torch_mg", "torch_solve", "compliance", "sensitivity",
           "virtual_heat", "filters", "torch_util", "mma"]


def module_aliases(tree):
    """local name -> aliased module, for `import sttopt.m as x` / `from sttopt import m`."""
    alias = {}
    for node in ast.walk(tree):
        if isinstance(node, ast.ImportFrom) and node.module in ("sttopt", None):
            for a in node.names:
                if a.name in ALIASED:
                    alias[a.asname or a.name] = a.name
        elif isinstance(node, ast.Import):
            for a in node.names:
                if a.name.startswith("sttopt.") and a.name[7:] in ALIASED:
                    alias[a.asname or a.name] = a.name[7:]
    return alias


def used_names(pkg: pathlib.Path):
    uses = {m: set() for m in ALIASED}
    for f in sorted(pkg.glob("*.py")):
        tree = ast.parse(f.read_text())
        alias = module_aliases(tree)
        for node in ast.walk(tree):
            if isinstance(node, ast.ImportFrom) and node.module:
                m = node.module.split(".")[-1]
                if m in ALIASED and m != f.stem:
                    uses[m].update(a.name for a in node.names)
            if isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name):
                m = alias.get(node.value.id)
                if m is not None and m != f.stem:
                    uses[m].add(node.attr)
    return uses


def signatures(path: pathlib.Path):
    out = {}
    for node in ast.parse(path.read_text()).body:
        if isinstance(node, ast.FunctionDef):
            a = node.args
            out[node.name] = {
                "positional": [p.arg for p in a.posonlyargs + a.args],
                "keyword_only": [p.arg for p in a.kwonlyargs],
                "n_defaults": len(a.defaults),
            }
    return out


def main(pkg: str):
    pkg = pathlib.Path(pkg)
    uses = used_names(pkg)
    table = {}
    for m in ALIASED:
        sigs = signatures(pkg / f"{m}.py")
        table[m] = {name: sigs.get(name) for name in sorted(uses[m])}
    out = pathlib.Path(__file__).with_name("driver_names.json")
    out.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    print(f"wrote {out}: {sum(len(v) for v in table.values())} names")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/sttopt")
