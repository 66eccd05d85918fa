"""The autotuner's choice of launch shapes from timing lines (no GPU): each
option replaces the default only when more than 2% faster, and the emitted
header carries every chosen field."""
import importlib.util
import os

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location(
    "jms_tune", os.path.join(HERE, "..", "jmultislicelib_amd", "csrc", "tune", "tune.py"))
T = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(T)


def _lines(n, row_ms=1.0, flat_ms=1.0, panel=(1.0, 1.0, 1.0, 1.0), x16=None, x8=None):
    out = [{"n": n, "pass": "row", "variant": "default", "lpb": 4, "gtw": 0, "ms": row_ms},
           {"n": n, "pass": "row", "variant": "tuned", "lpb": 4, "gtw": 0, "ms": 1.0},
           {"n": n, "pass": "row", "variant": "tuned_flat", "lpb": 4, "gtw": 0, "ms": flat_ms},
           {"n": n, "pass": "layout", "row_ms": panel[0], "row_panel_ms": panel[1], "col_ms": panel[2],
            "col_panel_ms": panel[3]},
           {"n": n, "pass": "col_nt", "cw": 8, "nt": 0, "ms": 1.0, "panel_ms": 0.8 if n == 480 else 1.0}]
    if x16 is not None:
        out.append({"n": n, "pass": "xcol", "ct": 16, "g": 0, "rot": 0, "nt": 1, "ms": x16})
        out.append({"n": n, "pass": "xcol", "ct": 8, "g": 2, "rot": 16, "nt": 0, "ms": x8})
        out.append({"n": n, "pass": "xcol", "ct": 8, "g": 1, "rot": 8, "nt": 0, "ms": 0.1 * x8})  # not selectable
        out.append({"n": n, "pass": "xcol", "ct": 8, "g": 2, "rot": 16, "nt": 1, "ms": 0.1 * x8})  # other policy
    return out


def test_choices_per_length():
    res = (_lines(480, flat_ms=0.9, panel=(1.0, 0.95, 1.0, 0.9)) + _lines(512, flat_ms=0.99) +
           _lines(1024, x16=0.706, x8=0.597) + _lines(2048, x16=0.706, x8=0.70))
    ch = T.choose(res)
    assert ch[480] == {"padr": 24, "panel": True, "ntc": False}
    assert ch[512] == {}  # 1% faster: kept out
    assert ch[1024] == {"cw": 8}
    assert ch[2048] == {}


def test_tile_width_policy_and_layout_chosen_together():
    """768-like timings: the pair of passes
    (row in the staged layout + column) decides width, cache policy and
    layout together."""
    res = _lines(768, panel=(0.386, 0.422, 0.674, 0.671)) + [
        {"n": 768, "pass": "col_cw_nt", "cw": 16, "nt": 1, "ms": 0.674, "panel_ms": 0.670},
        {"n": 768, "pass": "col_cw_nt", "cw": 16, "nt": 0, "ms": 0.666, "panel_ms": 0.625},
        {"n": 768, "pass": "col_cw_nt", "cw": 8, "nt": 1, "ms": 0.739, "panel_ms": 0.572},
        {"n": 768, "pass": "col_cw_nt", "cw": 8, "nt": 0, "ms": 0.691, "panel_ms": 0.596}]
    k = T.choose(res)[768]
    assert k["cw"] == 8 and k["panel"] is True and k["ntc"] is True


def test_emit_writes_chosen_fields(tmp_path):
    res = _lines(480, flat_ms=0.9, panel=(1.0, 0.95, 1.0, 0.9)) + _lines(1024, x16=0.706, x8=0.597)
    path = tmp_path / "tuned.h"
    T.emit(res, str(path))
    txt = path.read_text()
    i = txt.index("template <> struct Tuned<480>")
    blk = txt[i:txt.index("};", i)]
    assert "PADR = 24;" in blk and "PANEL = true;" in blk and "NTC = false;" in blk
    i = txt.index("template <> struct Tuned<1024>")
    assert "CW = 8;" in txt[i:txt.index("};", i)]
