"""The placement problem as arrays, built from parsed LEF + DEF data.

Follows src/database.py:11-154 of the reference (units, region, movable / fixed split, blocked rectangles, pin
indexing, free-area fractions, DEF coordinates). Length unit: one site width of the narrowest ROW site; origin at
the lower left of the placement region (bounding box of the rows); positions are cell centres.

Pin indexing (database.py:65-100): a movable pin is (cell index, offset from the cell centre); a pin of a fixed
component or a design pin is (num_mov, absolute position). Pins of a net are contiguous, nets in DEF order.
"""
import numpy as np

_SWAP = ("W", "E", "FW", "FE")


def orient_size(w, h, orient):
    return (h, w) if orient in _SWAP else (w, h)


def orient_point(x, y, w, h, orient):
    """Point (x, y) of a w x h cell (from its lower left, orientation N) after placing it with `orient`."""
    return {"N": (x, y), "S": (w - x, h - y), "FN": (w - x, y), "FS": (x, h - y),
            "W": (h - y, x), "E": (y, w - x), "FW": (y, x), "FE": (h - y, w - x)}[orient]


class PlaceData(object):
    def __init__(self, lef, d):
        self.d = d
        self.name = d.design
        if not d.rows:
            raise ValueError("the DEF has no ROW: no placement region")
        sites = sorted({r[0] for r in d.rows})
        missing = [s for s in sites if s not in lef.sites]
        if missing:
            raise ValueError("ROW site(s) not in the LEF: %s" % ", ".join(missing))
        dbu = d.dbu
        self.unit_dbu = int(round(min(lef.sites[s][0] for s in sites) * dbu))
        self.row_height_dbu = int(round(min(lef.sites[s][1] for s in sites) * dbu))

        rects = []
        for site, x, y, nx, ny, sx, sy in d.rows:
            sw, sh = (int(round(v * dbu)) for v in lef.sites[site])
            rects.append((x, y, x + (nx - 1) * sx + sw, y + (ny - 1) * sy + sh))
        self.row_rects = np.array(rects, dtype=np.int64)
        self.region = (int(self.row_rects[:, 0].min()), int(self.row_rects[:, 1].min()),
                       int(self.row_rects[:, 2].max()), int(self.row_rects[:, 3].max()))
        u = float(self.unit_dbu)
        self.width = (self.region[2] - self.region[0]) / u
        self.height = (self.region[3] - self.region[1]) / u
        self.row_height = self.row_height_dbu / u

        names, comp_macro = d.macro_names, d.comp_macro_id
        unknown = [m for m in names if m not in lef.macros]
        if unknown:
            raise ValueError("%d macro(s) of the DEF are not in the LEF: %s" % (len(unknown), ", ".join(unknown[:5])))
        master_size = np.array([[lef.macros[m].width, lef.macros[m].height] for m in names]).reshape(-1, 2) * dbu
        size_dbu = master_size[comp_macro]
        core = np.array([lef.macros[m].cls.startswith("CORE") for m in names], dtype=bool)
        # fixed: FIXED / COVER, and placed components that are not core cells (they stay where they are)
        self.fixed = d.comp_fixed | (d.comp_placed & ~core[comp_macro])
        self.kept_in_place = int((self.fixed & ~d.comp_fixed).sum())
        self.mov_index = np.nonzero(~self.fixed)[0]
        self.num_mov = len(self.mov_index)
        mov_id = np.full(len(comp_macro), self.num_mov, dtype=np.int64)
        mov_id[self.mov_index] = np.arange(self.num_mov)
        self.size = size_dbu[self.mov_index] / u
        self.mov_area = float(self.size.prod(1).sum())

        blocked = [tuple(b) for b in d.blockages]
        for i in np.nonzero(self.fixed)[0]:
            w, h = orient_size(size_dbu[i, 0], size_dbu[i, 1], d.comp_orient[i])
            x, y = d.comp_xy[i]
            blocked.append((x, y, x + w, y + h))
        self.block_rects = np.array(blocked, dtype=np.int64).reshape(-1, 4)

        origin = np.array(self.region[:2], dtype=np.float64)
        io = d.pin_comp < 0
        comp = np.where(io, 0, d.pin_comp)
        npn = len(d.pin_names)
        key = comp_macro[comp] * npn + d.pin_name                  # (master, pin name) of every connection
        pairs, slot = np.unique(key, return_inverse=True)
        pair_xy = np.full((len(pairs), 2), np.nan)
        for k, pk in enumerate(pairs):
            xy_m = lef.macros[names[pk // npn]].pins.get(d.pin_names[pk % npn])
            if xy_m is not None:
                pair_xy[k] = xy_m
        xy = pair_xy[slot.reshape(-1)] * dbu
        bad = np.nonzero(np.isnan(xy[:, 0]) & ~io)[0]
        if len(bad):
            raise ValueError("pin %s of macro %s is not in the LEF"
                             % (d.pin_names[d.pin_name[bad[0]]], names[comp_macro[comp[bad[0]]]]))
        pin_cell = np.where(io, self.num_mov, mov_id[comp])
        pin_off = (xy - size_dbu[comp] / 2) / u
        if io.any():
            pin_off[io] = (d.io_xy[-1 - d.pin_comp[io]] - origin) / u
        for k in np.nonzero(~io & (pin_cell == self.num_mov))[0]:
            c = comp[k]
            ox, oy = orient_point(xy[k, 0], xy[k, 1], size_dbu[c, 0], size_dbu[c, 1], d.comp_orient[c])
            pin_off[k] = (d.comp_xy[c] + (ox, oy) - origin) / u
        keep = ~np.isnan(pin_off[:, 0])
        pin_net = np.repeat(np.arange(len(d.net_name)), np.diff(d.net_start))
        self.unplaced_io_pins = int((~keep).sum())
        self.pin_cell, self.pin_off, self.pin_net = pin_cell[keep], pin_off[keep], pin_net[keep]
        self.num_nets = len(d.net_name)
        if not self.num_mov or not self.num_nets:
            raise ValueError("nothing to place: %d movable cells, %d nets" % (self.num_mov, self.num_nets))
        self.num_pins = len(self.pin_cell)
        self.cell_pins = np.bincount(self.pin_cell, minlength=self.num_mov + 1)[:self.num_mov].astype(np.float64)

    def fixed_fraction(self, nbx, nby):
        """[nbx, nby] share of every bin that is not placeable (outside the rows or under a fixed object).

        Exact (database.py:103-134): union of rows minus union of blockers on the grid of all rectangle edges, then
        the overlap of every edge interval with every bin."""
        x0, y0, x1, y1 = self.region
        rects = np.concatenate([self.row_rects, self.block_rects])
        xs = np.unique(np.clip(np.concatenate([rects[:, 0], rects[:, 2], [x0, x1]]), x0, x1))
        ys = np.unique(np.clip(np.concatenate([rects[:, 1], rects[:, 3], [y0, y1]]), y0, y1))

        def covered(r):
            cnt = np.zeros((len(xs), len(ys)), dtype=np.int32)
            if len(r):
                a = np.searchsorted(xs, np.clip(r[:, 0], x0, x1))
                b = np.searchsorted(xs, np.clip(r[:, 2], x0, x1))
                c = np.searchsorted(ys, np.clip(r[:, 1], y0, y1))
                e = np.searchsorted(ys, np.clip(r[:, 3], y0, y1))
                np.add.at(cnt, (a, c), 1)
                np.add.at(cnt, (b, e), 1)
                np.add.at(cnt, (a, e), -1)
                np.add.at(cnt, (b, c), -1)
            return (cnt.cumsum(0).cumsum(1) > 0)[:-1, :-1]

        free = (covered(self.row_rects) & ~covered(self.block_rects)).astype(np.float64)

        def overlap(edges, lo, hi, nb):
            b = lo + (hi - lo) * np.arange(nb + 1) / float(nb)
            return np.clip(np.minimum(edges[1:, None], b[None, 1:]) - np.maximum(edges[:-1, None], b[None, :-1]), 0, None)

        cap = overlap(xs, x0, x1, nbx).T.dot(free).dot(overlap(ys, y0, y1, nby))
        bin_area = (x1 - x0) / float(nbx) * (y1 - y0) / float(nby)
        return np.clip(1.0 - cap / bin_area, 0.0, 1.0)

    def to_def_xy(self, pos):
        """Cell centres (placer units) -> lower-left corners in database units (rounded)."""
        ll = (np.asarray(pos, dtype=np.float64) - self.size / 2) * self.unit_dbu
        return np.rint(ll + np.array(self.region[:2], dtype=np.float64)).astype(np.int64)

    def placed_pos(self, placed=None):
        """Cell centres of the movable cells as placed in a DEF of the same netlist."""
        d, i = placed or self.d, self.mov_index
        if list(d.comp_name) != list(self.d.comp_name):
            raise ValueError("the placed DEF has other components than the design")
        if not d.comp_placed[i].all():
            raise ValueError("%d movable cells of the DEF have no location" % int((~d.comp_placed[i]).sum()))
        swap = np.array([d.comp_orient[k] in _SWAP for k in i], dtype=bool)
        size = np.where(swap[:, None], self.size[:, ::-1], self.size)
        return (d.comp_xy[i] - np.array(self.region[:2])) / float(self.unit_dbu) + size / 2
