"""The module boundary: every name the reference package reads from the nine modules
this package replaces (tests/golden/driver_names.json, scanned from the reference's
sources) exists here, and each function keeps the reference's positional parameters
(names and order) and keyword-only parameters, so the drivers (stto, stto_heat,
stto_laplace, seqopt) and their helpers run on the aliased modules unchanged."""

import importlib
import inspect
import json
import pathlib

import pytest

TABLE = json.loads((pathlib.Path(__file__).parent / "golden" / "driver_names.json").read_text())
CASES = [(m, name, sig) for m, names in sorted(TABLE.items()) for name, sig in sorted(names.items())]


def test_the_table_covers_the_aliased_modules():
    assert set(TABLE) == {"torch_fem", "torch_mg", "torch_solve", "compliance", "sensitivity",
                          "virtual_heat", "filters", "torch_util", "mma"}
    assert len(CASES) >= 40


@pytest.mark.parametrize("module, name, sig", CASES, ids=[f"{m}.{n}" for m, n, _ in CASES])
def test_module_exports_name_with_reference_signature(module, name, sig):
    mod = importlib.import_module(f"sttopt_amd.{module}")
    assert hasattr(mod, name), f"sttopt_amd.{module} lacks {name}"
    if sig is None:
        return
    fn = getattr(mod, name)
    params = list(inspect.signature(fn).parameters.values())
    positional = [p.name for p in params
                  if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD)]
    keyword_only = {p.name for p in params if p.kind == p.KEYWORD_ONLY}
    var_kw = any(p.kind == p.VAR_KEYWORD for p in params)
    assert positional[:len(sig["positional"])] == sig["positional"], (positional, sig)
    # extra positional parameters here must be optional
    for p in params[len(sig["positional"]):]:
        if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD):
            assert p.default is not inspect.Parameter.empty, p.name
    # the reference's optional positional parameters are optional here too
    n_opt = sig["n_defaults"]
    for pname in sig["positional"][len(sig["positional"]) - n_opt:]:
        p = inspect.signature(fn).parameters[pname]
        assert p.default is not inspect.Parameter.empty, pname
    for kw in sig["keyword_only"]:
        assert kw in keyword_only or var_kw, f"{module}.{name} lacks keyword {kw}"
