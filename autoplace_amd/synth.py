"""Synthetic LEF + DEF designs for the placement workloads.

A design is built at the level of parsed LEF / DEF data (`Lef`, `Def`), the same data a LEF / DEF reader produces,
so `PlaceData` builds the placement problem from it exactly as from files; `lefdef.write_lef` / `write_def` turn it
into text that any LEF 5.8 / DEF 5.8 reader accepts.

Generator (seeded numpy PCG64, deterministic across machines):
  * standard cells one row tall (row = 10 sites, site 0.2 um x 2.0 um), widths uniform integers 1..8 sites;
    16 masters per width, each with 16 pins at uniform random points inside the cell box;
  * a square die at `util` = movable area / free area, rows spanning the die, side a whole number of rows;
  * `num_macros` fixed rectangular blocks covering `macro_frac` of the die, aligned to sites and rows, disjoint;
  * locality-driven netlist: every cell gets a hidden uniform point of the unit square; a net is a seed cell and its
    k - 1 nearest hidden neighbours, k = 2 (60 %), 3 (20 %), U{4..10} (17 %), U{11..50} (3 %);
  * `io_pins` optional design pins on the die boundary, each added to a random net.
"""
from collections import namedtuple

import numpy as np

Macro = namedtuple("Macro", "width height cls pins")      # microns; pins: {name: (x, y)} from the lower left
Lef = namedtuple("Lef", "sites macros")                   # sites: {name: (width, height)} microns

SITE_UM, ROW_UM, DBU = 0.2, 2.0, 1000
SITE_DBU, ROW_DBU = 200, 2000
ROW_SITES = 10
PINS_PER_MASTER = 16
VARIANTS = 16


class Def(object):
    """Parsed DEF (database-unit integers), the fields a global placer reads."""

    def __init__(self):
        self.path = ""
        self.design = ""
        self.dbu = DBU
        self.die = (0, 0, 0, 0)
        self.rows = []                   # (site, x, y, num_x, num_y, step_x, step_y)
        self.comp_name = []
        self.macro_names = []
        self.comp_macro_id = None        # int64 [n]
        self.comp_fixed = None           # bool [n]: FIXED / COVER
        self.comp_placed = None          # bool [n]: has a location
        self.comp_xy = None              # int64 [n, 2] lower left
        self.comp_orient = []
        self.io_name = []
        self.io_xy = None                # float64 [n, 2]
        self.net_name = []
        self.net_start = None            # int64 [nets + 1]
        self.pin_comp = None             # int64 [pins]: component, or -1 - io index
        self.pin_names = [""]
        self.pin_name = None             # int64 [pins]
        self.blockages = []
        self.soft_blockages = 0
        self.text = None                 # the DEF's text, when read from a file (the placed DEF rewrites it)
        self.comp_span = None            # (start, end) of the COMPONENTS statements in `text`


CONFIGS = {
    # name: (cells, nets, macros, macro_frac)
    "A": (10_000, 10_500, 0, 0.0),
    "B": (200_000, 210_000, 8, 0.20),
    "C": (2_000_000, 2_100_000, 16, 0.15),
}


def _degrees(rng, num_nets):
    u = rng.random(num_nets)
    deg = np.where(u < 0.6, 2, 3).astype(np.int64)
    mid = (u >= 0.8) & (u < 0.97)
    big = u >= 0.97
    deg[mid] = rng.integers(4, 11, size=int(mid.sum()))
    deg[big] = rng.integers(11, 51, size=int(big.sum()))
    return deg


def _cell_masters(rng):
    """-> {name: Macro}, names ordered by (width, variant). Pin shapes are 2 x 2 dbu squares."""
    macros = {}
    for w in range(1, 9):
        wd = w * SITE_DBU
        for v in range(VARIANTS):
            x = rng.integers(1, wd, size=PINS_PER_MASTER)
            y = rng.integers(1, ROW_DBU, size=PINS_PER_MASTER)
            pins = {}
            for j in range(PINS_PER_MASTER):
                # centre of the shape as a reader computes it from the written rectangle
                pins["P%d" % j] = (((x[j] - 1) / DBU + (x[j] + 1) / DBU) / 2, ((y[j] - 1) / DBU + (y[j] + 1) / DBU) / 2)
            macros["C%d_%d" % (w, v)] = Macro(wd / DBU, ROW_DBU / DBU, "CORE", pins)
    return macros


def _place_macros(rng, side, num_macros, macro_frac, max_tries=20000):
    """Disjoint fixed blocks, total area ~ macro_frac of the die. -> list of (x0, y0, w, h) in sites."""
    if num_macros == 0 or macro_frac <= 0:
        return []
    rows = side // ROW_SITES
    area_each = macro_frac * side * side / num_macros
    out = []
    for _ in range(num_macros):
        for _try in range(max_tries):
            r = rng.uniform(0.5, 2.0)
            w = int(round(np.sqrt(area_each * r)))
            h_rows = max(1, int(round(area_each / max(w, 1) / ROW_SITES)))
            w = max(1, int(round(area_each / (h_rows * ROW_SITES))))
            if w >= side or h_rows >= rows:
                continue
            x0 = int(rng.integers(0, side - w + 1))
            y0 = int(rng.integers(0, rows - h_rows + 1)) * ROW_SITES
            h = h_rows * ROW_SITES
            if all(x0 + w <= a or a + c <= x0 or y0 + h <= b or b + d <= y0 for a, b, c, d in out):
                out.append((x0, y0, w, h))
                break
        else:
            raise RuntimeError("could not place %d disjoint macros" % num_macros)
    return out


def generate(num_cells, num_nets=None, num_macros=0, macro_frac=0.0, util=0.70, seed=0, io_pins=0, name="synth"):
    """-> (Lef, Def) of a synthetic design (see the module docstring)."""
    from scipy.spatial import cKDTree

    if num_nets is None:
        num_nets = int(round(1.05 * num_cells))
    rng = np.random.Generator(np.random.PCG64(seed))
    masters = _cell_masters(rng)
    widths = rng.integers(1, 9, size=num_cells)
    variant = rng.integers(0, VARIANTS, size=num_cells)
    cell_area = float(widths.sum() * ROW_SITES)
    die_area = cell_area / util / (1.0 - macro_frac)
    side = int(np.ceil(np.sqrt(die_area) / ROW_SITES)) * ROW_SITES
    rows = side // ROW_SITES
    blocks = _place_macros(rng, side, num_macros, macro_frac)

    lef_macros = dict(masters)
    for k, (x0, y0, w, h) in enumerate(blocks):
        lef_macros["BLK%d" % k] = Macro(w * SITE_DBU / DBU, h * SITE_DBU / DBU, "BLOCK", {})
    lef = Lef({"core": (SITE_UM, ROW_UM)}, lef_macros)

    d = Def()
    d.design = name
    d.dbu = DBU
    d.die = (0, 0, side * SITE_DBU, side * SITE_DBU)
    d.rows = [("core", 0, r * ROW_DBU, side, 1, SITE_DBU, 0) for r in range(rows)]
    master_names = list(masters)                                  # index = (w - 1) * VARIANTS + v
    d.macro_names = master_names + ["BLK%d" % k for k in range(len(blocks))]
    n_comp = num_cells + len(blocks)
    d.comp_name = ["c%d" % i for i in range(num_cells)] + ["m%d" % k for k in range(len(blocks))]
    d.comp_macro_id = np.concatenate([(widths - 1) * VARIANTS + variant,
                                      len(master_names) + np.arange(len(blocks))]).astype(np.int64)
    d.comp_fixed = np.zeros(n_comp, dtype=bool)
    d.comp_fixed[num_cells:] = True
    d.comp_placed = d.comp_fixed.copy()
    xy = np.zeros((n_comp, 2), dtype=np.int64)
    for k, (x0, y0, w, h) in enumerate(blocks):
        xy[num_cells + k] = (x0 * SITE_DBU, y0 * SITE_DBU)
    d.comp_xy = xy
    d.comp_orient = ["N"] * n_comp

    # netlist: seed + k - 1 nearest hidden neighbours
    hidden = rng.random((num_cells, 2))
    deg = np.minimum(_degrees(rng, num_nets), num_cells)
    seeds = rng.integers(0, num_cells, size=num_nets)
    tree = cKDTree(hidden)
    net_start = np.zeros(num_nets + 1, dtype=np.int64)
    net_start[1:] = np.cumsum(deg)
    pin_comp = np.empty(int(net_start[-1]), dtype=np.int64)
    for k in np.unique(deg):
        nets = np.nonzero(deg == k)[0]
        _, nn = tree.query(hidden[seeds[nets]], k=int(k))
        nn = np.asarray(nn).reshape(len(nets), int(k))
        nn[:, 0] = seeds[nets]                                    # the seed itself (ties at distance 0)
        pin_comp[(net_start[nets][:, None] + np.arange(k)[None, :]).reshape(-1)] = nn.reshape(-1)
    # pin of each connection: the cell's connections take its pins in turn from a random first one
    first = rng.integers(0, PINS_PER_MASTER, size=num_cells)
    order = np.argsort(pin_comp, kind="stable")
    rank = np.empty(len(pin_comp), dtype=np.int64)
    counts = np.bincount(pin_comp, minlength=num_cells)
    starts = np.cumsum(counts) - counts
    rank[order] = np.arange(len(pin_comp)) - np.repeat(starts, counts)
    pin_idx = (first[pin_comp] + rank) % PINS_PER_MASTER
    d.pin_names = [""] + ["P%d" % j for j in range(PINS_PER_MASTER)]
    d.pin_name = pin_idx + 1

    # optional design pins on the boundary, each appended to a random net
    d.io_name = ["io%d" % k for k in range(io_pins)]
    if io_pins:
        t = rng.random(io_pins) * 4
        s = (t % 1) * side * SITE_DBU
        edge = t.astype(np.int64)
        W = side * SITE_DBU
        io = np.stack([np.where(edge == 0, s, np.where(edge == 1, W, np.where(edge == 2, W - s, 0))),
                       np.where(edge == 0, 0, np.where(edge == 1, s, np.where(edge == 2, W, W - s)))], 1)
        d.io_xy = np.rint(io).astype(np.float64)
        target = np.sort(rng.choice(num_nets, size=io_pins, replace=False))
        ins = net_start[target + 1]                               # insert at the end of each chosen net
        pin_comp = np.insert(pin_comp, ins, -1 - np.arange(io_pins))
        d.pin_name = np.insert(d.pin_name, ins, 0)
        add = np.zeros(num_nets + 1, dtype=np.int64)
        add[target + 1] = 1
        net_start = net_start + np.cumsum(add)
    else:
        d.io_xy = np.zeros((0, 2), dtype=np.float64)
    d.pin_comp = pin_comp
    d.net_start = net_start
    d.net_name = ["n%d" % i for i in range(num_nets)]
    return lef, d


def generate_config(name, seed=0, **kw):
    """Designs of the standard workloads: 'A' (10k cells), 'B' (200k cells, 8 macros), 'C' (2M cells, 16 macros)."""
    cells, nets, macros, frac = CONFIGS[name]
    return generate(cells, nets, macros, frac, seed=seed, name="synth_%s" % name, **kw)
