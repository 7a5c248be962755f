"""On-disk formats of the reference's console and GUI front ends, for images computed here.

- ``.dat``: the image writer of ``Source/widimage.cpp:1479-1500``: a 19-byte text header
  ``"%4d %4d %4d %4d" % (minX, maxX, minY, maxY)`` followed by the FP64 values row by row
  (y = minY..maxY-1, x = minX..maxX-1), native byte order (little-endian on x86-64).
- ``.csv``: ``Source/widimage.cpp:1505-1523``: ``X:;minX;maxX``, ``Y:;minY;maxY``, then one line
  per row ``Y=<y>:;v;v;...`` with the values in Qt's default stream notation (6 significant
  digits, like C's ``%g``).
- ``.dat.nofit``: the four no-fit rectangles read at ``Source/widimage.cpp:248-262``: one line
  per rectangle ``left top right bottom``, missing numbers read as -1.
- ``.ini``: the parameter files of ``SC_CalcCons::loadParameter``
  (``Source/sc_calcCons.cpp:107-399``): QSettings INI groups, the calculation's keys in the
  method group (``FCC Spheres``, historically, or ``CalcData``), grid and beam inputs in
  ``[Inputs]``, then the per-lattice-type overrides.
- key=value lists (``sas_scatter2Cons --mcpval "radius=5;cbparticle=cube;..."``,
  ``Source/sc_mainCons.cpp:76-118, 858-912``): ';'-separated pairs whose keys are the
  parameter names with their GUI prefixes shortened and lower-cased.

The image arrays are in the engine's IDX order (x = ihex fastest), which is the order of the
reference's ``data()``, so ``SasCalc.doCalculation()`` output goes to these writers as is.
"""
import re
import struct
from urllib.parse import unquote

import numpy as np

__all__ = ["write_dat", "read_dat", "write_csv", "read_csv", "read_nofit", "write_nofit", "read_ini",
           "PARAMETERS", "keyvalue_names", "read_keyvalues"]


def _bounds(b):
    x0, x1, y0, y1 = (int(v) for v in b)
    if x1 <= x0 or y1 <= y0:
        raise ValueError("empty bounds")
    return x0, x1, y0, y1


def write_dat(path, image, bounds):
    """Binary image file: 19-byte header, then (maxY-minY) x (maxX-minX) doubles."""
    x0, x1, y0, y1 = _bounds(bounds)
    img = np.ascontiguousarray(image, dtype="<f8").reshape(-1)
    if img.size != (x1 - x0) * (y1 - y0):
        raise ValueError("image size does not match the bounds")
    head = "%4d %4d %4d %4d" % (x0, x1, y0, y1)
    with open(path, "wb") as f:
        f.write(head.encode("ascii"))
        f.write(img.tobytes())


def read_dat(path):
    """(bounds, image) of a .dat file; the header is the text before the first double."""
    raw = open(path, "rb").read()
    # the header has 4 integers, each printed right-aligned in a field of 4 characters; wider
    # numbers make it longer, so parse the leading text and keep the rest as doubles
    m = re.match(rb"\s*(-?\d+)\s+(-?\d+)\s+(-?\d+)\s+(-?\d+)", raw)
    if not m:
        raise ValueError("not a .dat image: no header")
    x0, x1, y0, y1 = (int(v) for v in m.groups())
    n = (x1 - x0) * (y1 - y0)
    hl = len(raw) - 8 * n
    if hl < m.end() or n <= 0:
        raise ValueError("size of %s does not match its header" % path)
    return (x0, x1, y0, y1), np.frombuffer(raw, dtype="<f8", offset=hl).copy()


def _g(v):
    return "%g" % v


def write_csv(path, image, bounds):
    """Text image file readable by spreadsheets (';' separated)."""
    x0, x1, y0, y1 = _bounds(bounds)
    img = np.asarray(image, dtype=np.float64).reshape(y1 - y0, x1 - x0)
    with open(path, "w", newline="\n") as f:
        f.write("X:;%d;%d\n" % (x0, x1))
        f.write("Y:;%d;%d\n" % (y0, y1))
        for r, y in enumerate(range(y0, y1)):
            f.write("Y=%d:" % y + "".join(";" + _g(v) for v in img[r]) + "\n")


def read_csv(path):
    """(bounds, image) of a .csv image file (values at the file's 6-digit precision)."""
    with open(path) as f:
        lines = f.read().splitlines()
    _, x0, x1 = lines[0].split(";")
    _, y0, y1 = lines[1].split(";")
    x0, x1, y0, y1 = int(x0), int(x1), int(y0), int(y1)
    rows = [np.array([float(v) for v in ln.split(";")[1:]]) for ln in lines[2:2 + (y1 - y0)]]
    return (x0, x1, y0, y1), np.concatenate(rows)


def read_nofit(path):
    """Four (left, top, right, bottom) rectangles for setNoFitRect; absent lines/numbers -> -1."""
    rects = []
    try:
        lines = open(path).read().splitlines()
    except OSError:
        lines = []
    for i in range(4):
        sl = lines[i].split() if i < len(lines) else []
        sl = [int(v) for v in sl[:4]] + [-1] * max(0, 4 - len(sl))
        rects.append(tuple(sl))
    return rects


def write_nofit(path, rects):
    with open(path, "w", newline="\n") as f:
        for r in list(rects)[:4] + [(-1, -1, -1, -1)] * max(0, 4 - len(rects)):
            f.write("%d %d %d %d\n" % tuple(int(v) for v in r))


def _ini_groups(path):
    groups, cur = {}, None
    for ln in open(path, encoding="utf-8", errors="replace").read().splitlines():
        s = ln.strip()
        if not s or s.startswith(";") or s.startswith("#"):
            continue
        if s.startswith("[") and s.endswith("]"):
            cur = unquote(s[1:-1])
            groups.setdefault(cur, {})
            continue
        if "=" in s and cur is not None:
            k, v = s.split("=", 1)
            groups[cur][k.strip()] = v.strip()
    return groups


def _num(v):
    """QVariant::toDouble / toBool of an INI string: true/false -> 1/0, unparsable -> 0."""
    t = v.strip().strip('"')
    if t.lower() == "true":
        return 1.0
    if t.lower() == "false":
        return 0.0
    try:
        return float(t)
    except ValueError:
        return 0.0


# Keys of the method group that loadParameter skips (radio-button groups, :156-161).
_SKIPPED = {"EditQmaxPreset", "CenterMidpoint"}


def read_ini(path, known_keys=None):
    """Parameters of a reference .ini file as engine key=value pairs.

    Returns (params, fitflags, unused): params maps engine keys to floats (booleans as
    0/1), fitflags maps keys with a FITFLAG_ entry to whether it marks them as fitted, and
    unused lists method-group keys the engine does not know (the reference prints these as
    'ini keys not used').  known_keys: the engine's key names (default: every key of the
    group is returned and the caller's engine decides)."""
    groups = _ini_groups(path)
    names = [g for g in groups if g not in ("Inputs", "FFT", "AI")]
    used = "FCC Spheres" if "FCC Spheres" in names else "CalcData"
    if len(names) > 1:
        m = groups.get("Inputs", {}).get("CurMethod", "")
        if m:
            mm = m.split(" ")[0] if " " in m else m
            for g in names:
                if g.startswith(mm):
                    used = g
                    break
    grp = groups.get(used, {})
    params, fitflags, unused = {}, {}, []
    for k, v in grp.items():
        if k.startswith("FITFLAG_") or k in _SKIPPED:
            continue
        if known_keys is not None and k not in known_keys:
            unused.append(k)
            continue
        x = _num(v)
        if k in ("EditPixelX", "EditPixelY") and x < 0.01:
            x *= 1000.0  # old files give the pixel size in metres (:161-166)
        params[k] = x
        ff = grp.get("FITFLAG_" + k)
        if ff is not None:
            fitflags[k] = ff[:1] != "0"
    inp = groups.get("Inputs", {})
    gp = inp.get("EditGridPoints")
    params["GridPoints"] = float(int(_num(gp))) if gp is not None and _num(gp) >= 0 else float(int(_num(inp.get("GridPoints", "100"))))
    hk = inp.get("Edithklmax")
    params["HKLmax"] = float(int(_num(hk))) if hk is not None and _num(hk) >= 0 else float(int(_num(inp.get("HKLmax", "3"))))
    params["RadioButtonQ1"] = _num(inp.get("RadioButtonQ1", "false"))
    params["RadioButtonQ2"] = _num(inp.get("RadioButtonQ2", "false"))
    if _num(inp.get("RadioButtonQ4", "false")):
        params["RadioButtonQ4"], params["ExpandImage"] = 1.0, 0.0
    else:
        params["RadioButtonQ4"], params["ExpandImage"] = 0.0, _num(inp.get("ExpandImage", "false"))
    params["BeamPosX"] = float(int(_num(inp.get("EditCenterX", "0"))))
    params["BeamPosY"] = float(int(_num(inp.get("EditCenterY", "0"))))
    # Per-lattice-type overrides (:265-398).  Only updates of parameter keys take effect:
    # the table's "ComboboxParticle" differs in case from the key "ComboBoxParticle" and
    # "HKLmax" is a global input, so those entries change nothing in the reference either.
    lt = int(params.get("LType", 0))
    twin_off = {1, 2, 3, 7, 8, 17, 9, 10, 11, 12, 13, 14, 15, 16, 18, 19}
    if lt == 0:
        params["CheckBoxTwinned"] = 0.0
        params["Length"] = 250.0
    elif lt in (1, 2, 3):
        params["CheckBoxTwinned"] = 0.0
        params["Length"] = 500.0
    elif lt in twin_off:
        params["CheckBoxTwinned"] = 0.0
    if lt == 10:  # OBDD: radius from the cell (:324-337); "EditLength"/"EditPhi" are not keys
        params["EditRadius"] = 0.5484 * 0.6200 * params.get("uca", 0.0)
    if lt in (13, 19):
        params["ComboBoxPeak"] = 7.0
    return params, fitflags, unused


# The calculation's parameters as the console builds them from the key table
# (Source/sc_calc_generic.cpp:36-138 via Source/sc_calcCons.cpp:24-101): name, kind
# ('C' selection, 'N' number, 'I' integer, 'T' toggle) and, for selections, the item texts
# in index order.  Output-only entries (EditRelDis, EditDist) are not parameters.
_LTYPES = ("Lamellae|Hexagonally packed cylinders (P6/mm)|Square packed cylinders (P4/mm)|"
           "Rectangular centered cylinders (cmm)|BCC (lm3m)|FCC (Fm3m)|HCP (P6/mmc)|SC (Pm3m)|"
           "BCT (l4/mm)|Gyroid (Ia3d)|OBDD (Pn3m)|Plumbers Nightmare (Im3m)|None|CP-Layers|"
           "2DHex (GISAXS)|2DSquare (GISAXS)|1DLam (GISAXS)|Fd3m|Orthorombic|LDQ12").split("|")
PARAMETERS = [
    ("LType", "C", _LTYPES),
    ("uca", "N", None), ("ucb", "N", None), ("ucc", "N", None), ("ucalpha", "N", None),
    ("ucbeta", "N", None), ("ucgamma", "N", None), ("EditCeff", "N", None), ("CheckBoxTwinned", "T", None),
    ("EditCeffcyl", "N", None), ("reff", "N", None), ("acpl", "N", None), ("bcpl", "N", None),
    ("ifluc", "N", None), ("rfluc", "N", None),
    ("ComboBoxParticle", "C", ("sphere|cylinder|disk|vesicle|cube|ellipsoid|triaxial ellipsoid|"
                               "super ellipsoid, barrel|superball|excluded volume chain|Kratky Porod chain").split("|")),
    ("Ordis", "C", ("Gaussian|Exponential|Onsager|Maier-Saupe|Cut-off|Laguerre|z-dir|isotropic|"
                    "mirrored Gaussian|mirrored Exponential|mirrored Onsager|mirrored Maier-Saupe|"
                    "mirrored Cut-off|fiber pattern").split("|")),
    ("ComboBoxInterior", "C", "homogeneous|core + homogeneous sh|core + inhomogeneous sh|multi-shell|myelin".split("|")),
    ("EditRadius", "N", None), ("EditRadiusi", "N", None), ("EditSigma", "N", None), ("EditDbeta", "N", None),
    ("Length", "N", None), ("SigmaL", "N", None), ("Alpha", "N", None), ("EditRho", "N", None),
    ("RotAlpha", "N", None),
    ("ComboBoxPeak", "C", ("Lorentzian|Gaussian|mod. 1 Lorentzian|mod. 2 Lorentzian|Pseudo-Voigt|"
                           "Pearson VII|Gamma|Anisotropic Gaussian").split("|")),
    ("EditPeakPar", "N", None), ("EditDebyeWaller", "N", None), ("EditAzi", "N", None),
    ("EditDomainSize", "N", None), ("RadButDebyeScherrer", "T", None), ("RadioButtonPara", "T", None),
    ("VAx1", "N", None), ("VAx2", "N", None), ("VAx3", "N", None), ("Ay1", "N", None), ("Ay2", "N", None),
    ("Ay3", "N", None), ("Az1", "N", None), ("Az2", "N", None), ("Az3", "N", None),
    ("SigX", "N", None), ("SigY", "N", None), ("SigZ", "N", None),
    ("ucpsi", "N", None), ("ucn1", "N", None), ("ucn2", "N", None), ("ucn3", "N", None),
    ("theta", "N", None), ("phi", "N", None), ("rotTheta", "N", None), ("rotPhi", "N", None),
    ("GridPoints", "I", None), ("HKLmax", "I", None), ("EditQmax", "N", None), ("CalcQmax", "N", None),
    ("EditQmaxData", "T", None), ("EditQmaxPreset", "T", None), ("EditQmin", "N", None),
    ("EditQsteps", "I", None),
    ("EditPixelNoX", "N", None), ("EditPixelNoY", "N", None), ("EditPixelX", "N", None),
    ("EditPixelY", "N", None), ("EditDet", "N", None), ("EditWavelength", "N", None),
    ("BeamPosX", "N", None), ("BeamPosY", "N", None), ("CenterBeam", "T", None), ("CenterMidpoint", "T", None),
    ("EditBFactor", "N", None), ("CheckBoxWAXS", "T", None), ("P1", "N", None),
    ("iso", "N", None), ("I0", "N", None), ("Base", "N", None),
]
_KIND = {k: (t, items) for k, t, items in PARAMETERS}

# Keys left out of the key=value list (Source/sc_mainCons.cpp:87-97): prefix and substring tests.
_KV_PREFIX = ("EditAxis", "Editdom", "ExpandImage", "RadioButtonQ", "P1", "EditQmaxData", "EditQmaxPreset")
_KV_SUBSTR = ("EditQsteps", "EditQmin", "EditBFactor", "CheckBoxWAXS", "CalcQmax", "CenterBeam",
              "CenterMidpoint", "EditRelDis", "EditDist")


def _short_name(key):
    """The printed name of a key (Source/sc_mainCons.cpp:104-112): GUI prefixes shortened, lower case."""
    if key.startswith("VAx"):
        k = key[1:]
    elif key.startswith("Edit"):
        k = key[4:]
    elif key.startswith("CheckBox"):
        k = key[8:]
    elif key.startswith("ComboBox"):
        k = "CB" + key[8:]
    elif key.startswith("RadBut"):
        k = key[6:]
    elif key.startswith("RadioButton"):
        k = "RB" + key[11:]
    else:
        k = key
    return k.lower()


def keyvalue_names(center_midpoint=False):
    """{printed name: parameter key} of the key=value interface.  With the beam centre at the
    image midpoint the reference also drops BeamPosX/BeamPosY (Source/sc_mainCons.cpp:81-86)."""
    keys = sorted(k for k, _, _ in PARAMETERS)
    if center_midpoint:
        keys = [k for k in keys if k not in ("BeamPosX", "BeamPosY")]
    out = {}
    for k in keys:
        if k.startswith(_KV_PREFIX) or any(s in k for s in _KV_SUBSTR):
            continue
        out.setdefault(_short_name(k), k)
    return out


def _leading_number(v):
    """QString::toDouble of the whole text: the number, or 0 when the text is not one."""
    try:
        return float(v)
    except ValueError:
        return 0.0


def read_keyvalues(text, center_midpoint=False):
    """Parameters of a key=value list as engine keys (calculateFromChatbotKeyValList,
    Source/sc_mainCons.cpp:858-893).  Returns (params, ignored).

    Pairs are separated by ';' (empty pairs skipped); the name is taken as written (not
    trimmed), the value trimmed.  A value starting with a digit, '-' or '+' is a number; one
    starting with 'T' or 'F' a toggle (so is an item text such as "FCC (Fm3m)", which sets 0,
    as in the reference); "NULL" is skipped; anything else is a selection's item text, whose
    index is stored (-1 when no item has that text, as the reference does).
    Names the list does not define, and item texts given for non-selections, are returned in
    `ignored`."""
    names = keyvalue_names(center_midpoint)
    params, ignored = {}, []
    for kv in text.split(";"):
        if not kv:
            continue
        p = kv.find("=")
        if p < 0:
            continue
        val = kv[p + 1:].strip()
        key = names.get(kv[:p])
        if key is None or not val:
            ignored.append(kv)
            continue
        kind, items = _KIND[key]
        if val[0].isdigit() or val[0] in "-+":
            params[key] = _leading_number(val)
        elif val[0] in "FT":
            params[key] = 1.0 if val[0] == "T" else 0.0
        elif val == "NULL":
            continue
        elif kind == "C":
            params[key] = float(items.index(val)) if val in items else -1.0
        else:
            ignored.append(kv)
    return params, ignored
