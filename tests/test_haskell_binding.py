"""The Haskell binding (shapes_amd/haskell/Physics/Engine/GPU.hs) against
the C ABI it imports: every foreign import names a symbol that
include/shapes_gpu.h declares and the library exports, with the same
number and kinds of arguments and result, and the byte layouts the module
writes are those of the C structures.  (No Haskell compiler is available
here; this checks the module's interface to the library, not its Haskell.)"""
import os
import re

import pytest

from shapes_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HS = os.path.join(ROOT, "shapes_amd", "haskell", "Physics", "Engine", "GPU.hs")
HEADER = os.path.join(ROOT, "include", "shapes_gpu.h")

HS_KIND = {"CInt": "i32", "Int32": "i32", "Int64": "i64", "CLLong": "i64",
           "CUInt": "u32", "Word32": "u32", "CDouble": "f64", "CFloat": "f32",
           "CString": "ptr"}


def hs_kind(t):
    t = t.strip()
    if t.startswith(("Ptr", "FunPtr")) or t == "CString":
        return "ptr"
    return HS_KIND[t]


def c_kind(arg):
    arg = arg.strip()
    if "*" in arg:
        return "ptr"
    t = arg.rsplit(None, 1)[0] if " " in arg else arg
    t = t.replace("const ", "").strip()
    return {"int32_t": "i32", "shp_status": "i32", "int64_t": "i64", "uint32_t": "u32",
            "double": "f64", "float": "f32", "void": None}[t]


def header_prototypes():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for ret, name, args in re.findall(r"\b(shp_status|int32_t|const char\*)\s+(shp_\w+)\s*\((.*?)\);",
                                      text, flags=re.S):
        kinds = [c_kind(a) for a in args.split(",")]
        out[name] = ("ptr" if "*" in ret else "i32", [k for k in kinds if k is not None])
    return out


def hs_imports():
    text = open(HS).read()
    out = []
    for amp, sym, hname, sig in re.findall(
            r'foreign import ccall (?:safe|unsafe) "(&?)(\w+)"\s*\n\s*(\w+)\s*::\s*(.*?)(?=\nforeign|\n\n)',
            text, flags=re.S):
        parts = [p.strip() for p in " ".join(sig.split()).split("->")]
        out.append((amp == "&", sym, hname, parts))
    return out


def test_imports_match_the_header():
    protos = header_prototypes()
    imports = hs_imports()
    assert len(imports) >= 12
    lib = engine.load_library()
    for is_addr, sym, hname, parts in imports:
        assert sym in protos, sym
        assert hasattr(lib, sym), sym
        assert sym in engine.EXPORTS, sym
        if is_addr:  # a finalizer: the address of the function
            assert parts == ["FunPtr (Ptr ShpBatch", "IO ())"] or parts[0].startswith("FunPtr")
            continue
        ret, args = protos[sym]
        assert parts[-1].startswith("IO "), (sym, parts)
        assert hs_kind(parts[-1][3:]) == ret, (sym, parts[-1])
        got = [hs_kind(p) for p in parts[:-1]]
        assert got == args, (sym, got, args)


def _const(name):
    m = re.search(rf"^{name} = (\d+)$", open(HS).read(), flags=re.M)
    assert m, name
    return int(m.group(1))


def test_layouts_match_the_c_structures():
    import ctypes as C
    assert _const("worldDescBytes") == C.sizeof(engine.WorldDescC)
    assert _const("worldDescCapacityOffset") == engine.WorldDescC.capacity.offset
    assert _const("configBytes") == C.sizeof(engine.ConfigC)
    assert _const("bodyDescBytes") == C.sizeof(engine.BodyDescC)
    assert _const("bodyDescVerticesOffset") == engine.BodyDescC.n_vertices.offset
    assert engine.BodyDescC.hull_x.offset == _const("bodyDescVerticesOffset") + 8
    assert engine.BodyDescC.hull_y.offset == _const("bodyDescVerticesOffset") + 16
    # the pointer columns of shp_world_desc follow n_bodies at 8-byte steps
    names = ["px", "py", "theta", "vx", "vy", "omega", "mass", "inertia", "mu", "bounce",
             "hull_off", "hull_x", "hull_y", "radius"]
    for k, n in enumerate(names):
        assert getattr(engine.WorldDescC, n).offset == 8 + 8 * k
    assert engine.ConfigC.iterations.offset == 40


@pytest.mark.parametrize("name", ["initEngine", "changeScene", "updateWorld", "stepWorld",
                                  "runWorld", "readPhysObj", "append", "delete", "filled",
                                  "worldContacts", "worldAabbs"])
def test_reference_names_are_exported(name):
    """The module exports the reference's engine and world operations
    under their own names (Engine/Main.hs:43-98, World.hs:77-94,
    IOWorld.hs:74-90)."""
    text = open(HS).read()
    head = text[text.index("module Physics.Engine.GPU"):text.index(") where")]
    assert re.search(rf"\b{name}\b", head)
    assert re.search(rf"^{name} ::", text, flags=re.M)
