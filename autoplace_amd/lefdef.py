"""LEF / DEF reading and writing for the placer's inputs and its gp.def output (the behaviour of the reference's
utils/io_parser.py:22-301: what a global placer reads from LEF / DEF, and the placed DEF it writes).

Read:
  LEF  SITE sizes; per MACRO its SIZE, CLASS, ORIGIN and the centre of every PIN (the bounding box of the pin's PORT
       shapes, RECT and POLYGON, MASK numbers skipped). PROPERTYDEFINITIONS and OBS are skipped.
  DEF  DESIGN, UNITS, DIEAREA, ROWs; COMPONENTS (name, master, PLACED / FIXED / COVER location and orientation);
       PINS (location of the first LAYER shape's centre, turned by the pin's orientation); NETS (the connections
       before the first '+'); hard PLACEMENT BLOCKAGES (SOFT / PARTIAL ones are counted, not used). Every other
       section is skipped.
Written:
  write_lef / write_def   a library and a design (synth.Lef / synth.Def) as LEF 5.8 / DEF 5.8 text;
  write_placed_def        the input DEF's text with the moved components' placement statements replaced: every other
                          statement, and every other attribute of a moved component, is copied.

Both readers work on tokens: the DEF statement by statement (text between ';'), carrying which section it is in;
the LEF as nested blocks (SITE / MACRO / PIN / PORT ... END).
"""
import re

import numpy as np

from .synth import Def, Lef, Macro

_COMMENT = re.compile(r"#[^\n]*")
_TOKEN = re.compile(r"[;()]|[^\s;()]+")
_PLACEMENT = ("PLACED", "FIXED", "COVER")

# a design pin's shape point (x, y), relative to its placement point, in each orientation: (x, y) -> M (x, y)
_PIN_TURN = {"N": ((1, 0), (0, 1)), "S": ((-1, 0), (0, -1)), "W": ((0, -1), (1, 0)), "E": ((0, 1), (-1, 0)),
             "FN": ((-1, 0), (0, 1)), "FS": ((1, 0), (0, -1)), "FW": ((0, 1), (1, 0)), "FE": ((0, -1), (-1, 0))}


# ------------------------------------------------------------------------------------------------------ writing
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
    if len(d.blockages):
        lines.append("BLOCKAGES %d ;" % len(d.blockages))
        for x0, y0, x1, y1 in d.blockages:
            lines.append("- PLACEMENT RECT ( %d %d ) ( %d %d ) ;" % (x0, y0, x1, y1))
        lines += ["END BLOCKAGES", ""]
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


def write_placed_def(d, path, comp_index, xy, orient="N"):
    """gp.def: the components `comp_index` PLACED at `xy` (lower-left corners, DBU) with `orient`. A design read
    from a file keeps its text: only the moved components' placement attributes change. A design built in memory
    (synth) is written whole."""
    if d.text is None:
        write_def(d, path, xy, comp_index)
        return
    moved = {int(i): (int(x), int(y)) for i, (x, y) in zip(comp_index, np.asarray(xy))}
    start, end = d.comp_span
    body, k = [], 0
    for stmt in d.text[start:end].split(";"):
        toks = _TOKEN.findall(stmt)
        if len(toks) < 3 or toks[0] != "-":
            if stmt.strip():
                body.append(stmt + ";")
            continue
        if k in moved:
            rest = _without_placement(toks[3:])
            body.append("\n- %s %s%s + PLACED ( %d %d ) %s ;"
                        % (toks[1], toks[2], "".join(" " + t for t in rest), moved[k][0], moved[k][1], orient))
        else:
            body.append(stmt + ";")
        k += 1
    with open(path, "w") as f:
        f.write(d.text[:start])
        f.write("".join(body))
        f.write("\n")
        f.write(d.text[end:])


def _without_placement(attrs):
    """A component's attributes without its '+ UNPLACED' / '+ PLACED|FIXED|COVER ( x y ) orient' group."""
    out, i = [], 0
    while i < len(attrs):
        if attrs[i] == "+" and i + 1 < len(attrs) and attrs[i + 1] == "UNPLACED":
            i += 2
        elif attrs[i] == "+" and i + 1 < len(attrs) and attrs[i + 1] in _PLACEMENT:
            i += 7                                   # + STATUS ( x y ) orient
        else:
            out.append(attrs[i])
            i += 1
    return out


# -------------------------------------------------------------------------------------------------------- LEF
def _numbers(toks):
    """The coordinates of a shape statement's tokens (a MASK number skipped)."""
    out, skip = [], False
    for t in toks:
        if skip:
            skip = False
        elif t == "MASK":
            skip = True
        elif t != "ITERATE":
            out.append(float(t))
    return out


class _LefBlocks(object):
    """Nested-block reader over one LEF file's tokens."""

    def __init__(self, toks):
        self.t, self.i = toks, 0

    def statement(self):
        """Tokens up to the next ';' (consumed), the keyword first."""
        j = self.t.index(";", self.i)
        out = self.t[self.i:j]
        self.i = j + 1
        return out

    def at_end_of(self, name):
        return self.t[self.i] == "END" and self.i + 1 < len(self.t) and self.t[self.i + 1] == name

    def skip_to_end(self, name):
        while not self.at_end_of(name):
            self.i += 1
        self.i += 2

    def site(self, name):
        size = None
        while not self.at_end_of(name):
            st = self.statement()
            if st[0] == "SIZE":
                size = (float(st[1]), float(st[3]))
        self.i += 2
        return size

    def shapes(self):
        """A PORT block: the coordinates of its shapes, up to its bare END."""
        xs, ys = [], []
        while self.t[self.i] != "END":
            st = self.statement()
            if st[0] in ("RECT", "POLYGON"):
                v = _numbers(st[1:])
                xs += v[0::2]
                ys += v[1::2]
        self.i += 1
        return xs, ys

    def pin(self, name):
        xs, ys = [], []
        while not self.at_end_of(name):
            if self.t[self.i] == "PORT":
                self.i += 1
                px, py = self.shapes()
                xs += px
                ys += py
            else:
                self.statement()
        self.i += 2
        return xs, ys

    def macro(self, name):
        size, origin, cls, pins = None, (0.0, 0.0), "CORE", {}
        while not self.at_end_of(name):
            kw = self.t[self.i]
            if kw == "PIN":
                pin = self.t[self.i + 1]
                self.i += 2
                xs, ys = self.pin(pin)
                if xs:
                    pins[pin] = ((min(xs) + max(xs)) / 2 + origin[0], (min(ys) + max(ys)) / 2 + origin[1])
            elif kw == "OBS":
                self.i += 1
                self.shapes()
            else:
                st = self.statement()
                if kw == "SIZE":
                    size = (float(st[1]), float(st[3]))
                elif kw == "ORIGIN":
                    origin = (float(st[1]), float(st[2]))
                elif kw == "CLASS":
                    cls = " ".join(st[1:]).upper()
        self.i += 2
        return None if size is None else Macro(size[0], size[1], cls, pins)


def parse_lef(paths):
    """Read LEF files (technology and cell libraries, any order). -> Lef(sites, macros)."""
    sites, macros = {}, {}
    for path in paths:
        with open(path) as f:
            toks = _TOKEN.findall(_COMMENT.sub("", f.read()))
        r = _LefBlocks(toks)
        while r.i < len(toks):
            kw = toks[r.i]
            name = toks[r.i + 1] if r.i + 1 < len(toks) else ""
            if kw == "PROPERTYDEFINITIONS":
                r.i += 1
                r.skip_to_end("PROPERTYDEFINITIONS")
            elif kw == "SITE" and r.i + 2 < len(toks) and toks[r.i + 2] != ";":
                r.i += 2
                size = r.site(name)
                if size is not None:
                    sites[name] = size
            elif kw == "MACRO":
                r.i += 2
                m = r.macro(name)
                if m is not None:
                    macros[name] = m
            else:
                r.i += 1
    return Lef(sites, macros)


# -------------------------------------------------------------------------------------------------------- DEF
class _DefBuilder(object):
    """The fields of a Def, filled statement by statement."""

    def __init__(self, d):
        self.d = d
        self.macro_id = {}
        self.macro, self.fixed, self.placed, self.xy = [], [], [], []
        self.io_xy = []
        self.net_start, self.pin_comp, self.pin_name, self.name_id = [0], [], [], {"": 0}
        self.comp_id = self.io_id = None

    def header(self, toks):
        d, kw = self.d, toks[0]
        if kw == "DESIGN" and len(toks) > 1:
            d.design = toks[1]
        elif kw == "UNITS" and len(toks) >= 4:
            d.dbu = int(toks[3])
        elif kw == "DIEAREA":
            v = [int(t) for t in toks[1:] if t not in "()"]
            d.die = (min(v[0::2]), min(v[1::2]), max(v[0::2]), max(v[1::2]))
        elif kw == "ROW" and len(toks) >= 6:
            site, x, y = toks[2], int(toks[3]), int(toks[4])
            nx = ny = 1
            sx = sy = 0
            if "DO" in toks:
                k = toks.index("DO")
                nx, ny = int(toks[k + 1]), int(toks[k + 3])
                if "STEP" in toks:
                    k = toks.index("STEP")
                    sx, sy = int(toks[k + 1]), int(toks[k + 2])
            d.rows.append((site, x, y, nx, ny, sx, sy))

    def component(self, toks):
        self.d.comp_name.append(toks[1])
        self.macro.append(self.macro_id.setdefault(toks[2], len(self.macro_id)))
        status, x, y, orient = None, 0, 0, "N"
        for k in range(3, len(toks) - 6):
            if toks[k] == "+" and toks[k + 1] in _PLACEMENT:
                status, x, y, orient = toks[k + 1], int(toks[k + 3]), int(toks[k + 4]), toks[k + 6]
                break
        self.fixed.append(status in ("FIXED", "COVER"))
        self.placed.append(status is not None)
        self.xy.append((x, y))
        self.d.comp_orient.append(orient)

    def design_pin(self, toks):
        shape = where = None
        for k in range(2, len(toks)):
            if toks[k] == "LAYER" and shape is None:
                p = toks.index("(", k)
                shape = (int(toks[p + 1]), int(toks[p + 2]), int(toks[p + 5]), int(toks[p + 6]))
            elif toks[k] in _PLACEMENT and toks[k - 1] == "+" and where is None:
                where = (int(toks[k + 2]), int(toks[k + 3]), toks[k + 5])
        self.d.io_name.append(toks[1])
        if where is None:
            self.io_xy.append((np.nan, np.nan))
            return
        cx, cy = ((shape[0] + shape[2]) / 2.0, (shape[1] + shape[3]) / 2.0) if shape else (0.0, 0.0)
        (a, b), (c, e) = _PIN_TURN[where[2]]
        self.io_xy.append((where[0] + (a * cx + b * cy), where[1] + (c * cx + e * cy)))

    def net(self, toks):
        if self.comp_id is None:
            self.comp_id = {n: i for i, n in enumerate(self.d.comp_name)}
            self.io_id = {n: i for i, n in enumerate(self.d.io_name)}
        k = 2
        while k < len(toks) and toks[k] == "(":
            close = toks.index(")", k)
            who, what = toks[k + 1], toks[k + 2]
            if who == "PIN":
                self.pin_comp.append(-1 - self.io_id[what])
                self.pin_name.append(0)
            else:
                self.pin_comp.append(self.comp_id[who])
                self.pin_name.append(self.name_id.setdefault(what, len(self.name_id)))
            k = close + 1
        self.d.net_name.append(toks[1])
        self.net_start.append(len(self.pin_comp))

    def blockage(self, toks):
        if toks[1] != "PLACEMENT":
            return
        if "SOFT" in toks or "PARTIAL" in toks:
            self.d.soft_blockages += 1
            return
        for k in range(2, len(toks) - 8):
            if toks[k] == "RECT":
                self.d.blockages.append((int(toks[k + 2]), int(toks[k + 3]), int(toks[k + 6]), int(toks[k + 7])))

    def finish(self):
        d = self.d
        d.macro_names = list(self.macro_id)
        d.comp_macro_id = np.array(self.macro, dtype=np.int64)
        d.comp_fixed = np.array(self.fixed, dtype=bool)
        d.comp_placed = np.array(self.placed, dtype=bool)
        d.comp_xy = np.array(self.xy, dtype=np.int64).reshape(-1, 2)
        d.io_xy = np.array(self.io_xy, dtype=np.float64).reshape(-1, 2)
        d.pin_names = list(self.name_id)
        d.pin_name = np.array(self.pin_name, dtype=np.int64)
        d.net_start = np.array(self.net_start, dtype=np.int64)
        d.pin_comp = np.array(self.pin_comp, dtype=np.int64)
        return d


_SECTION_HANDLER = {"COMPONENTS": "component", "PINS": "design_pin", "NETS": "net", "BLOCKAGES": "blockage"}


def parse_def(path):
    """Read a DEF. -> Def (the fields of synth.Def, with the text kept for write_placed_def)."""
    d = Def()
    d.path = path
    with open(path) as f:
        d.text = f.read()
    b = _DefBuilder(d)
    section, pos, comp_start = None, 0, None
    for stmt in d.text.split(";"):
        toks = _TOKEN.findall(_COMMENT.sub("", stmt))
        while len(toks) >= 2 and toks[0] == "END":      # 'END SECTION' has no ';': it opens the next statement
            if toks[1] == "COMPONENTS" and comp_start is not None:
                d.comp_span = (comp_start, pos + stmt.index("END"))
            section = None
            toks = toks[2:]
        if toks:
            if section is None:
                if toks[0] == "PROPERTYDEFINITIONS" or (len(toks) == 2 and toks[1].isdigit()):
                    section = toks[0]                   # COMPONENTS 12, NETS 40, VIAS 3, ... or PROPERTYDEFINITIONS
                    if section == "COMPONENTS":
                        comp_start = pos + len(stmt) + 1
                else:
                    b.header(toks)
            elif toks[0] == "-" and section in _SECTION_HANDLER and len(toks) >= 2:
                getattr(b, _SECTION_HANDLER[section])(toks)
        pos += len(stmt) + 1
    if comp_start is not None and d.comp_span is None:
        d.comp_span = (comp_start, len(d.text))
    return b.finish()
