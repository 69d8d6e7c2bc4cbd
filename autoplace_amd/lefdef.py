"""Minimal LEF / DEF writer and reader for the placer's inputs and its gp.def output.

Reads what a global placer needs (the fields of synth.Lef / synth.Def, as the reference's utils/io_parser.py:22-261
reads them) and writes LEF 5.8 / DEF 5.8 text that readers of either implementation accept. The writer of the
placed DEF rewrites only the COMPONENTS statements of the moved cells (io_parser.py:264-301 behaviour).
"""
import re

import numpy as np

from .synth import Def, Lef, Macro


def write_lef(lef, path, dbu=1000):
    out = ["VERSION 5.8 ;", "BUSBITCHARS \"[]\" ;", "DIVIDERCHAR \"/\" ;", "UNITS", "  DATABASE MICRONS %d ;" % dbu,
           "END UNITS", ""]
    for name, (w, h) in lef.sites.items():
        out += ["SITE %s" % name, "  CLASS CORE ;", "  SIZE %.4f BY %.4f ;" % (w, h), "END %s" % name, ""]
    for name, m in lef.macros.items():
        out += ["MACRO %s" % name, "  CLASS %s ;" % m.cls, "  ORIGIN 0 0 ;", "  SIZE %.4f BY %.4f ;" % (m.width, m.height)]
        for pin, (x, y) in m.pins.items():
            # a 2 x 2 dbu square around the pin centre (the centre reads back as (lo + hi) / 2)
            xd, yd = int(round(x * dbu)), int(round(y * dbu))
            out += ["  PIN %s" % pin, "    DIRECTION INOUT ;", "    PORT", "      LAYER M1 ;",
                    "        RECT %.4f %.4f %.4f %.4f ;" % ((xd - 1) / dbu, (yd - 1) / dbu, (xd + 1) / dbu, (yd + 1) / dbu),
                    "    END", "  END %s" % pin]
        out += ["END %s" % name, ""]
    out.append("END LIBRARY")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_def(d, path, comp_xy=None, moved=None):
    """Write the design. comp_xy / moved: lower-left positions (dbu) of the components `moved`, written PLACED N."""
    placed = {}
    if moved is not None:
        placed = {int(i): (int(x), int(y)) for i, (x, y) in zip(moved, comp_xy)}
    lines = ["VERSION 5.8 ;", "DIVIDERCHAR \"/\" ;", "BUSBITCHARS \"[]\" ;", "DESIGN %s ;" % d.design,
             "UNITS DISTANCE MICRONS %d ;" % d.dbu, "",
             "DIEAREA ( %d %d ) ( %d %d ) ;" % d.die, ""]
    for k, (site, x, y, nx, ny, sx, sy) in enumerate(d.rows):
        lines.append("ROW ROW_%d %s %d %d N DO %d BY %d STEP %d %d ;" % (k, site, x, y, nx, ny, sx, sy))
    lines += ["", "COMPONENTS %d ;" % len(d.comp_name)]
    for i, name in enumerate(d.comp_name):
        m = d.macro_names[d.comp_macro_id[i]]
        if i in placed:
            lines.append("- %s %s + PLACED ( %d %d ) N ;" % ((name, m) + placed[i]))
        elif d.comp_fixed[i]:
            lines.append("- %s %s + FIXED ( %d %d ) %s ;" % (name, m, d.comp_xy[i, 0], d.comp_xy[i, 1], d.comp_orient[i]))
        elif d.comp_placed[i]:
            lines.append("- %s %s + PLACED ( %d %d ) %s ;" % (name, m, d.comp_xy[i, 0], d.comp_xy[i, 1], d.comp_orient[i]))
        else:
            lines.append("- %s %s + UNPLACED ;" % (name, m))
    lines += ["END COMPONENTS", ""]
    if len(d.io_name):
        lines.append("PINS %d ;" % len(d.io_name))
        for k, name in enumerate(d.io_name):
            x, y = d.io_xy[k]
            lines.append("- %s + NET %s + DIRECTION INOUT + LAYER M1 ( -1 -1 ) ( 1 1 ) + FIXED ( %d %d ) N ;"
                         % (name, name, int(x), int(y)))
        lines += ["END PINS", ""]
    lines.append("NETS %d ;" % len(d.net_name))
    for i, name in enumerate(d.net_name):
        conn = []
        for p in range(d.net_start[i], d.net_start[i + 1]):
            c = d.pin_comp[p]
            if c < 0:
                conn.append("( PIN %s )" % d.io_name[-1 - c])
            else:
                conn.append("( %s %s )" % (d.comp_name[c], d.pin_names[d.pin_name[p]]))
        lines.append("- %s %s ;" % (name, " ".join(conn)))
    lines += ["END NETS", "", "END DESIGN"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def parse_lef(paths):
    """-> Lef(sites, macros): site sizes; per macro size, class, pin centres (bounding box of the pin's shapes)."""
    sites, macros = {}, {}
    for path in paths:
        with open(path) as f:
            t = re.sub(r"#[^\n]*", "", f.read()).replace(";", " ; ").split()
        i, n = 0, len(t)
        while i < n:
            if t[i] == "SITE" and i + 2 < n and t[i + 2] != ";":
                name = t[i + 1]
                while not (t[i] == "END" and t[i + 1] == name):
                    if t[i] == "SIZE":
                        sites[name] = (float(t[i + 1]), float(t[i + 3]))
                    i += 1
                i += 2
                continue
            if t[i] == "MACRO":
                name, size, origin, cls, pins = t[i + 1], None, (0.0, 0.0), "CORE", {}
                i += 2
                while not (t[i] == "END" and t[i + 1] == name):
                    if t[i] == "SIZE":
                        size = (float(t[i + 1]), float(t[i + 3]))
                    elif t[i] == "ORIGIN":
                        origin = (float(t[i + 1]), float(t[i + 2]))
                    elif t[i] == "CLASS":
                        j = t.index(";", i)
                        cls = " ".join(t[i + 1:j]).upper()
                    elif t[i] == "OBS":
                        while t[i] != "END":
                            i += 1
                    elif t[i] == "PIN":
                        pin, xs, ys, depth = t[i + 1], [], [], 0
                        i += 2
                        while depth or not (t[i] == "END" and t[i + 1] == pin):
                            if t[i] == "PORT":
                                depth = 1
                            elif t[i] == "END":
                                depth = 0
                            elif t[i] in ("RECT", "POLYGON"):
                                j = t.index(";", i)
                                v = [float(s) for s in t[i + 1:j] if s[0] in "-.0123456789"]
                                xs += v[0::2]
                                ys += v[1::2]
                                i = j
                            i += 1
                        if xs:
                            pins[pin] = ((min(xs) + max(xs)) / 2 + origin[0], (min(ys) + max(ys)) / 2 + origin[1])
                        i += 1
                    i += 1
                if size:
                    macros[name] = Macro(size[0], size[1], cls, pins)
                i += 2
                continue
            i += 1
    return Lef(sites, macros)


def _section(text, name):
    head = text.find("\n%s " % name)
    if head < 0:
        return None
    start = text.index(";", head) + 1
    end = text.find("\nEND %s" % name, start)
    return start, (end if end >= 0 else len(text))


def parse_def(path):
    """-> Def with the fields of synth.Def."""
    d = Def()
    d.path = path
    with open(path) as f:
        text = f.read()
    span = _section(text, "COMPONENTS")
    head = text[:span[0]] if span else text
    m = re.search(r"^[ \t]*DESIGN[ \t]+(\S+)", head, re.M)
    d.design = m.group(1) if m else ""
    m = re.search(r"UNITS\s+DISTANCE\s+MICRONS\s+(\d+)", head)
    d.dbu = int(m.group(1)) if m else 1000
    m = re.search(r"^[ \t]*DIEAREA([^;]*);", head, re.M)
    if m:
        v = [int(s) for s in re.findall(r"-?\d+", m.group(1))]
        d.die = (min(v[0::2]), min(v[1::2]), max(v[0::2]), max(v[1::2]))
    for m in re.finditer(r"^[ \t]*ROW\s+\S+\s+(\S+)\s+(-?\d+)\s+(-?\d+)\s+\S+"
                         r"(?:\s+DO\s+(\d+)\s+BY\s+(\d+)(?:\s+STEP\s+(-?\d+)\s+(-?\d+))?)?", head, re.M):
        g = m.groups()
        d.rows.append((g[0], int(g[1]), int(g[2]), int(g[3] or 1), int(g[4] or 1), int(g[5] or 0), int(g[6] or 0)))
    fixed, placed, xy, macro, macro_id = [], [], [], [], {}
    if span:
        for stmt in text[span[0]:span[1]].split(";"):
            t = stmt.split()
            if len(t) < 3 or t[0] != "-":
                continue
            d.comp_name.append(t[1])
            macro.append(macro_id.setdefault(t[2], len(macro_id)))
            status, x, y, orient = "", 0, 0, "N"
            for i in range(3, len(t) - 6):
                if t[i] == "+" and t[i + 1] in ("PLACED", "FIXED", "COVER"):
                    status, x, y, orient = t[i + 1], int(t[i + 3]), int(t[i + 4]), t[i + 6]
                    break
            fixed.append(status in ("FIXED", "COVER"))
            placed.append(status != "")
            xy.append((x, y))
            d.comp_orient.append(orient)
    d.macro_names, d.comp_macro_id = list(macro_id), np.array(macro, dtype=np.int64)
    d.comp_fixed = np.array(fixed, dtype=bool)
    d.comp_placed = np.array(placed, dtype=bool)
    d.comp_xy = np.array(xy, dtype=np.int64).reshape(-1, 2)
    span = _section(text, "PINS")
    io_xy = []
    if span:
        for stmt in text[span[0]:span[1]].split(";"):
            t = stmt.split()
            if len(t) < 2 or t[0] != "-":
                continue
            rect, place = None, None
            for i in range(2, len(t)):
                if t[i] == "LAYER" and rect is None:
                    j = t.index("(", i)
                    rect = (int(t[j + 1]), int(t[j + 2]), int(t[j + 5]), int(t[j + 6]))
                elif t[i] in ("PLACED", "FIXED", "COVER") and t[i - 1] == "+" and place is None:
                    place = (int(t[i + 2]), int(t[i + 3]), t[i + 5])
            d.io_name.append(t[1])
            if place is None:
                io_xy.append((np.nan, np.nan))
            else:
                cx, cy = ((rect[0] + rect[2]) / 2.0, (rect[1] + rect[3]) / 2.0) if rect else (0.0, 0.0)
                if place[2] != "N":
                    raise ValueError("design pin orientation %s not supported by this reader" % place[2])
                io_xy.append((place[0] + cx, place[1] + cy))
    d.io_xy = np.array(io_xy, dtype=np.float64).reshape(-1, 2)
    span = _section(text, "NETS")
    net_start, pin_comp, pin_name, name_id = [0], [], [], {"": 0}
    pat = re.compile(r"\(\s*(\S+)\s+(\S+)[^)]*\)")
    if span:
        comp_id = {n: i for i, n in enumerate(d.comp_name)}
        io_id = {n: i for i, n in enumerate(d.io_name)}
        for stmt in text[span[0]:span[1]].split(";"):
            t = stmt.split(None, 2)
            if len(t) < 2 or t[0] != "-":
                continue
            for a, b in pat.findall(stmt):
                if a == "PIN":
                    pin_comp.append(-1 - io_id[b])
                    pin_name.append(0)
                else:
                    pin_comp.append(comp_id[a])
                    pin_name.append(name_id.setdefault(b, len(name_id)))
            d.net_name.append(t[1])
            net_start.append(len(pin_comp))
    d.pin_names = list(name_id)
    d.pin_name = np.array(pin_name, dtype=np.int64)
    d.net_start = np.array(net_start, dtype=np.int64)
    d.pin_comp = np.array(pin_comp, dtype=np.int64)
    return d
