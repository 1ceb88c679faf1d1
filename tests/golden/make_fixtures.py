"""Write the golden fixtures in tests/golden/ from the reference's own literal tables.

Run in the development container only (it reads the reference checkout as
text; nothing is imported or executed from it):

    python tests/golden/make_fixtures.py /path/to/DTC-reference

The reference ships no numerical test data for this path, but it does hold
its element stiffness matrices and load-case tables as literal expressions.
This script reads those expressions as data, evaluates them with a small
arithmetic evaluator (numbers, the table's named constants, + - * / **), and
stores the evaluated matrices:

  ke08_nu{0.3,0.2}  24x24 table of Hexe08 (f-src/mod_linfe.f90:161-766) before `factor`
  ke20_nu{0.3,0.2}  60x60 table of Hexe20 (f-src/mod_linfe.f90:890-4509) before `factor`
  uu08              init_displ_hexe8 uu(l,d,lc) at eps=1e-6 (f-src/mod_write_deck.f90:63-107)
  uu20              init_displ_hexe20 uu(l,d,lc) (f-src/mod_write_deck.f90:118-279)
  vv_lin_unit       init_loadcase_el_order_lin vv(dof,lc) with eps=1 (f-src/mod_struct_calcmat.f90:1684-1770)
  vv_quad_unit      init_loadcase_el_order_quad vv(dof,lc) with e=1, which rows were assigned,
                    and rows whose literal length is not 60 (stored as NaN)
  meta_tc01_*       the keyword/value pairs of regression/TC01-1_*.meta
"""
from __future__ import annotations

import ast
import operator
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.Div: operator.truediv, ast.Pow: operator.pow, ast.USub: operator.neg, ast.UAdd: operator.pos}


_REAL = re.compile(r"(?<![\w.])(\d+\.\d*|\.\d+)([eEdD][+-]?\d+)?(_rk|_8)?(?![\w.])")


def _eval(expr: str, names: dict) -> float:
    """Arithmetic-only evaluation of a Fortran real expression.  A real literal
    with a kind suffix (_rk, _8) is double precision; one without is of the
    default real kind, i.e. single precision (the reference is built without a
    default-real promotion flag), so `.9*eps` is float32(0.9) * eps."""
    def lit(m):
        txt = m[1] + ((m[2] or "").replace("d", "e").replace("D", "e"))
        if "." in txt and txt.startswith("."):
            txt = "0" + txt
        if txt.endswith("."):
            txt += "0"
        txt = txt.replace(".e", ".0e")
        return txt if m[3] else f"_f4({txt})"
    e = _REAL.sub(lit, expr)
    tree = ast.parse(e.strip(), mode="eval")

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id == "_f4" and len(n.args) == 1:
            return float(np.float32(ev(n.args[0])))
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return float(n.value)
        if isinstance(n, ast.Name):
            return float(names[n.id])
        if isinstance(n, ast.BinOp) and type(n.op) in _OPS:
            return _OPS[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.UnaryOp) and type(n.op) in _OPS:
            return _OPS[type(n.op)](ev(n.operand))
        raise ValueError(f"unsupported expression {expr!r}")
    return ev(tree)


def _logical_lines(lines):
    """Join Fortran free-form continuation lines and strip comments."""
    out, cur = [], ""
    for raw in lines:
        s = raw.split("!")[0].rstrip()
        st = s.lstrip()
        if st.startswith("&"):
            st = st[1:]
        cont = st.endswith("&")
        if cont:
            st = st[:-1]
        cur += " " + st
        if not cont:
            out.append(cur.strip())
            cur = ""
    if cur.strip():
        out.append(cur.strip())
    return out


def _split_list(body: str):
    return [t.strip() for t in body.split(",") if t.strip()]


def parse_ke(src_lines, start, end, n, nu):
    K = np.full((n, n), np.nan)
    for ln in src_lines[start - 1:end]:
        m = re.match(r"\s*C_FE\((\d+),(\d+)\)\s*=\s*(.*)$", ln)
        if m:
            K[int(m[1]) - 1, int(m[2]) - 1] = _eval(m[3], {"nu": nu})
    assert not np.isnan(K).any()
    return K


def parse_uu(src_lines, start, end, nen, names):
    uu = np.zeros((nen, 3, 3 * nen))
    for ln in _logical_lines(src_lines[start - 1:end]):
        m = re.match(r"uu\(:,\s*([^,]+),\s*([^)]+)\)\s*=\s*\[(.*)\]$", ln)
        if not m:
            continue
        d = int(round(_eval(m[1], names))) - 1
        lc = int(round(_eval(m[2], names))) - 1
        vals = [_eval(t, names) for t in _split_list(m[3])]
        assert len(vals) == nen, (ln, len(vals))
        uu[:, d, lc] = vals
    return uu


def parse_vv(src_lines, start, end, n, names):
    vv = np.zeros((n, n))
    assigned = np.zeros(n, bool)
    bad = []
    for ln in _logical_lines(src_lines[start - 1:end]):
        m = re.match(r"vv\(\s*(\d+),\s*:\)\s*=\s*\[(.*)\]$", ln)
        if not m:
            continue
        r = int(m[1]) - 1
        vals = [_eval(t, names) for t in _split_list(m[2])]
        if len(vals) != n:
            # the HEX20 table holds a row with 61 entries (row 54); it cannot be
            # placed, so the row is recorded as NaN
            vv[r, :] = np.nan
            bad.append((r + 1, len(vals)))
            continue
        vv[r, :] = vals
        assigned[r] = True
    return vv, assigned, bad


def find_line(src_lines, pattern, start=1):
    for i in range(start - 1, len(src_lines)):
        if re.search(pattern, src_lines[i]):
            return i + 1
    raise KeyError(pattern)


def parse_meta(path):
    kv = {}
    for ln in open(path, encoding="utf-8", errors="replace"):
        m = re.match(r"^([wri])\s+(\w+)\s+(.*)$", ln.rstrip("\n"))
        if m:
            val = m[3].split("!")[0].strip()
            kv[m[2]] = val
    return kv


def main(ref):
    linfe = open(os.path.join(ref, "f-src/mod_linfe.f90")).read().splitlines()
    deck = open(os.path.join(ref, "f-src/mod_write_deck.f90")).read().splitlines()
    calc = open(os.path.join(ref, "f-src/mod_struct_calcmat.f90")).read().splitlines()
    out = {}
    h08 = find_line(linfe, r"Function Hexe08\(mc\)")
    h08e = find_line(linfe, r"end Function Hexe08", h08)
    h20 = find_line(linfe, r"Function Hexe20\(mc\)")
    h20e = find_line(linfe, r"C_FE = C_FE\*factor", h20)
    for nu in (0.3, 0.2):
        out[f"ke08_nu{nu}"] = parse_ke(linfe, h08, h08e, 24, nu)
        out[f"ke20_nu{nu}"] = parse_ke(linfe, h20, h20e, 60, nu)
    u8 = find_line(deck, r"Function init_displ_hexe8")
    u8e = find_line(deck, r"End Function init_displ_hexe8", u8)
    out["uu08"] = parse_uu(deck, u8, u8e, 8, {"eps": 1e-6})
    u20 = find_line(deck, r"Function init_displ_hexe20")
    u20e = find_line(deck, r"End Function init_displ_hexe20", u20)
    out["uu20"] = parse_uu(deck, u20, u20e, 20, {"eps": 1e-6, "z": 0.0, "f": 0.99794, "oh": 1.5,
                                                 "h": 0.5, "ish": 1})
    vl = find_line(calc, r"subroutine init_loadcase_el_order_lin")
    vle = find_line(calc, r"end subroutine init_loadcase_el_order_lin", vl)
    out["vv_lin_unit"], _, _ = parse_vv(calc, vl, vle, 24, {"eps": 1.0, "eps2": 2.0, "eps3": 3.0, "eps4": 4.0})
    vq = find_line(calc, r"subroutine init_loadcase_el_order_quad")
    vqe = find_line(calc, r"end subroutine init_loadcase_el_order_quad", vq)
    out["vv_quad_unit"], out["vv_quad_assigned"], bad = parse_vv(
        calc, vq, vqe, 60, {"e": 1.0, "z": 0.0, "f": 0.99794, "oh": 1.5, "h": 0.5, "zz": 2.0})
    out["vv_quad_bad_rows"] = np.array(bad, np.int64).reshape(-1, 2)
    np.savez_compressed(os.path.join(HERE, "reference_tables.npz"), **out)

    metas = {}
    for name in ("isotropic-std-HEX08-ME1", "isotropic-std-HEX08-ME2", "orthotropic-std-HEX08-ME1"):
        p = os.path.join(ref, f"regression/TC01-1_test_dev_dtc_{name}.meta")
        if os.path.exists(p):
            metas[name] = parse_meta(p)
    import json
    with open(os.path.join(HERE, "regression_meta.json"), "w") as f:
        json.dump(metas, f, indent=1, sort_keys=True)
    print("wrote", sorted(out), sorted(metas))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
