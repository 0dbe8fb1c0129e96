"""Extract a chemmech-generated mechanism's tables into a JSON file.

The ROS3 path consumes the mechanism only through the arrays of the
generated Fortran module ``RXNS_DATA`` (for CB6r5:
CCTM/src/MECHS/cb6r5_ae7_aq/RXNS_DATA_MODULE.F90).  CCTM links those arrays
into the solver; this library receives them through ``cmaq_ros3_init``.
Outside CCTM (tests, benchmark, PyTorch harness) the same arrays are read
from the JSON written here, which holds the table values unchanged, with
1-based species/reaction ids exactly as in the Fortran DATA statements.

Usage (maintainers only, needs the CMAQ source tree):
    python -m cmaq_amd.mechs.extract_tables <RXNS_DATA_MODULE.F90> <out.json>
"""

import json
import re
import sys

# Arrays copied into the JSON (name -> kind).  Everything the ROS3 path and
# CALC_RCONST read from RXNS_DATA (rbsparse.F:48, rbfeval.F:50,
# rbjacob.F:49, rbsolver.F:61, RXNS_FUNC_MODULE.F90:221).
_ARRAYS = {
    "CHEMISTRY_SPC": "str", "SPECIES_MOLWT": "real", "CONVERT_CONC": "bool",
    "SPECIES_TYPE": "str", "CGRID_INDEX": "int",
    "IRR": "int", "SC": "real", "NREACT": "int", "NPRDCT": "int",
    "KTYPE": "int", "IRXBITS": "int", "IORDER": "int", "RTDAT": "real",
    "IRRFALL": "int", "RFDAT": "real", "IPH": "int", "IHETERO": "int",
    "PHOTAB": "str", "HETERO": "str", "RXLABEL": "str",
    "NRXWM": "int", "NRXWW": "int", "NRXWO2": "int", "NRXWN2": "int",
    "NRXWCH4": "int", "NRXWH2": "int",
}
_SCALARS = [
    "MECHNAME", "N_GAS_CHEM_SPC", "NUMB_MECH_SPC", "NRXNS", "KUNITS", "NMPHOT",
    "NFALLOFF", "MXPRD", "NWM", "NWW", "NWO2", "NWN2", "NWCH4", "NWH2",
    "ATM_AIR", "ATM_O2", "ATM_N2", "ATM_CH4", "ATM_H2", "MHETERO", "NPHOTAB",
    "NHETERO", "NSPECIAL_RXN", "NSPECIAL", "N_SS_SPC",
]


def _strip_comment(line):
    out, q = [], None
    for ch in line:
        if q:
            out.append(ch)
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            out.append(ch)
        elif ch == "!":
            break
        else:
            out.append(ch)
    return "".join(out).rstrip()


def _statements(text):
    """Free-form statements with '&' continuations joined."""
    buf = ""
    for raw in text.splitlines():
        line = _strip_comment(raw)
        s = line.strip()
        if buf and s.startswith("&"):
            s = s[1:]
        if s.endswith("&"):
            buf += s[:-1] + " "
            continue
        buf += s
        if buf.strip():
            yield buf.strip()
        buf = ""


def _split_values(s):
    vals, cur, q = [], "", None
    for ch in s:
        if q:
            cur += ch
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            cur += ch
        elif ch == ",":
            vals.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        vals.append(cur.strip())
    out = []
    for v in vals:
        m = re.fullmatch(r"(\d+)\*(.+)", v)
        if m and not v.startswith("'"):
            out += [m.group(2).strip()] * int(m.group(1))
        else:
            out.append(v)
    return out


def _conv(v, kind):
    if kind == "str":
        return v.strip("'\"").rstrip()
    if kind == "bool":
        return v.upper() in ("T", ".TRUE.")
    if kind == "int":
        return int(v)
    return float(v.replace("D", "E").replace("d", "e"))


def extract(path):
    text = open(path).read()
    params = {}
    data = {}  # name -> {index tuple: value}
    for st in _statements(text):
        m = re.match(r"(INTEGER|REAL\s*\(\s*8\s*\)|CHARACTER\s*\(\s*\d+\s*\))\s*,\s*PARAMETER\s*::\s*(\w+)\s*=\s*(.+)$", st, re.I)
        if m:
            name, val = m.group(2).upper(), m.group(3).strip()
            if val.startswith("'"):
                params[name] = val.strip("'").rstrip()
            else:
                try:
                    params[name] = int(val)
                except ValueError:
                    params[name] = float(val.replace("D", "E"))
            continue
        m = re.match(r"DATA\s+(.*?)\s*/(.*)/\s*$", st, re.I)
        if not m:
            continue
        targets, values = m.group(1), _split_values(m.group(2))
        mi = re.fullmatch(r"\(\s*(\w+)\s*\((.*)\)\s*,\s*(\w+)\s*=\s*(\w+)\s*,\s*(\w+)\s*\)", targets)
        if mi:
            name, idx, var = mi.group(1).upper(), mi.group(2), mi.group(3)
            lo = _resolve(mi.group(4), params)
            hi = _resolve(mi.group(5), params)
            parts = [p.strip() for p in idx.split(",")]
            if name not in _ARRAYS:
                continue
            if len(values) != hi - lo + 1:
                raise ValueError(f"{name}: {len(values)} values for {lo}..{hi}")
            for n, v in zip(range(lo, hi + 1), values):
                key = tuple(n if p == var else _resolve(p, params) for p in parts)
                data.setdefault(name, {})[key] = _conv(v, _ARRAYS[name])
            continue
        tlist = re.findall(r"(\w+)\s*\(([^)]*)\)", targets)
        if len(tlist) != len(values):
            continue
        for (name, idx), v in zip(tlist, values):
            name = name.upper()
            if name not in _ARRAYS:
                continue
            key = tuple(_resolve(p.strip(), params) for p in idx.split(","))
            data.setdefault(name, {})[key] = _conv(v, _ARRAYS[name])
    return params, data


def _resolve(tok, params):
    tok = tok.strip()
    if re.fullmatch(r"-?\d+", tok):
        return int(tok)
    return int(params[tok.upper()])


def _dense(entries, shape, default):
    """Fortran-indexed sparse dict -> nested list in row-major (first index outer)."""
    if len(shape) == 1:
        return [entries.get((i,), default) for i in range(1, shape[0] + 1)]
    return [[entries.get((i, j), default) for j in range(1, shape[1] + 1)]
            for i in range(1, shape[0] + 1)]


def build_json(path):
    p, d = extract(path)
    ns, nr, mxprd = p["NUMB_MECH_SPC"], p["NRXNS"], p["MXPRD"]
    nf, nmphot, mhet = p["NFALLOFF"], p["NMPHOT"], p["MHETERO"]
    if p.get("NSPECIAL_RXN", 0) != 0:
        raise ValueError("SPECIAL rate operators are not supported")
    out = {
        "mechname": p["MECHNAME"],
        "n_gas_chem_spc": p["N_GAS_CHEM_SPC"],
        "numb_mech_spc": ns, "nrxns": nr, "mxprd": mxprd, "kunits": p["KUNITS"],
        "nfalloff": nf, "nmphot": nmphot, "nphotab": p["NPHOTAB"],
        "mhetero": mhet, "nhetero": p["NHETERO"],
        "chemistry_spc": _dense(d["CHEMISTRY_SPC"], (ns,), ""),
        "species_molwt": _dense(d["SPECIES_MOLWT"], (ns,), 0.0),
        "species_type": _dense(d["SPECIES_TYPE"], (ns,), ""),
        "convert_conc": _dense(d["CONVERT_CONC"], (ns,), False),
        # IRR(NRXNS, MXPRD+3), SC(NRXNS, MXPRD): reaction-major lists
        "irr": _dense(d["IRR"], (nr, mxprd + 3), 0),
        "sc": _dense(d["SC"], (nr, mxprd), 0.0),
        "nreact": _dense(d["NREACT"], (nr,), 0),
        "nprdct": _dense(d["NPRDCT"], (nr,), 0),
        "ktype": _dense(d["KTYPE"], (nr,), 0),
        "irxbits": _dense(d["IRXBITS"], (nr,), 0),
        "iorder": _dense(d["IORDER"], (nr,), 0),
        # RTDAT(3, NRXNS) -> rtdat[r] = [RTDAT(1,r), RTDAT(2,r), RTDAT(3,r)]
        "rtdat": [list(x) for x in zip(*_dense(d["RTDAT"], (3, nr), 0.0))],
        "irrfall": _dense(d["IRRFALL"], (nf,), 0),
        "rfdat": [list(x) for x in zip(*_dense(d["RFDAT"], (5, nf), 0.0))],
        "iph": _dense(d["IPH"], (nmphot, 3), 0),
        "ihetero": _dense(d["IHETERO"], (mhet, 2), 0),
        "photab": _dense(d["PHOTAB"], (p["NPHOTAB"],), ""),
        "hetero": _dense(d["HETERO"], (p["NHETERO"],), ""),
        "rxlabel": _dense(d["RXLABEL"], (nr,), ""),
        "atm_air": p["ATM_AIR"], "atm_o2": p["ATM_O2"], "atm_n2": p["ATM_N2"],
        "atm_ch4": p["ATM_CH4"], "atm_h2": p["ATM_H2"],
    }
    for key, nkey in (("NRXWM", "NWM"), ("NRXWW", "NWW"), ("NRXWO2", "NWO2"),
                      ("NRXWN2", "NWN2"), ("NRXWCH4", "NWCH4"), ("NRXWH2", "NWH2")):
        n = p[nkey]
        out[key.lower()] = _dense(d.get(key, {}), (n,), 0) if n > 0 else []
    return out


if __name__ == "__main__":
    tables = build_json(sys.argv[1])
    with open(sys.argv[2], "w") as f:
        json.dump(tables, f, separators=(",", ":"))
    print(f"{tables['mechname']}: {tables['numb_mech_spc']} species, "
          f"{tables['nrxns']} reactions -> {sys.argv[2]}")
