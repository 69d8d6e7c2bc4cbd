"""The placement problem as arrays, from a parsed LEF library and DEF design (the behaviour of src/database.py:11-154
of the reference: units, region, movable / fixed split, blocked area, pin indexing, free-area fractions, DEF
coordinates).

Units. One placer length unit is the width of the narrowest site a ROW uses; the origin is the lower-left corner of
the placement region (the bounding box of the rows); a cell's position is its centre.

Pins (database.py:65-100). Every connection of a net becomes one pin row: a pin of a movable cell is (its cell
index, offset from the cell centre); a pin of a fixed component or a design pin is (num_mov, absolute position).
The pins of a net are contiguous, nets in DEF order. The arithmetic of the offsets is the reference's, operation for
operation, so the offsets agree bit for bit.
"""
import numpy as np

# orientation -> (swap the axes, mirror the first coordinate, mirror the second), applied to a point of an N-oriented
# w x h master: after the optional swap the point is (a, b) in an A x B box, and a mirror maps a to A - a
_ORIENT = {"N": (False, False, False), "S": (False, True, True), "FN": (False, True, False),
           "FS": (False, False, True), "W": (True, True, False), "E": (True, False, True),
           "FW": (True, False, False), "FE": (True, True, True)}


def orient_size(w, h, orient):
    """Footprint (width, height) of a w x h master placed with `orient`."""
    return (h, w) if _ORIENT[orient][0] else (w, h)


def orient_point(x, y, w, h, orient):
    """Where the point (x, y) of a w x h master (from its lower left, orientation N) lands, from the lower left of
    the placed footprint."""
    swap, mirror_a, mirror_b = _ORIENT[orient]
    a, b, A, B = (y, x, h, w) if swap else (x, y, w, h)
    return (A - a if mirror_a else a), (B - b if mirror_b else b)


def _site_units(lef, d):
    """-> (site width, row height) in DBU: the narrowest site width and the lowest site height of the rows."""
    if not d.rows:
        raise ValueError("cannot place %s: its DEF defines no ROW" % d.design)
    used = sorted({row[0] for row in d.rows})
    unknown = [s for s in used if s not in lef.sites]
    if unknown:
        raise ValueError("the LEF does not define the site(s) the rows use: %s" % ", ".join(unknown))
    widths = [lef.sites[s][0] * d.dbu for s in used]
    heights = [lef.sites[s][1] * d.dbu for s in used]
    return int(round(min(widths))), int(round(min(heights)))


def _row_boxes(lef, d):
    """[rows, 4] int64 (x0, y0, x1, y1) DBU: each row's sites, the last one ending its box."""
    out = np.empty((len(d.rows), 4), dtype=np.int64)
    for k, (site, x, y, nx, ny, sx, sy) in enumerate(d.rows):
        site_w, site_h = (int(round(v * d.dbu)) for v in lef.sites[site])
        out[k] = (x, y, x + (nx - 1) * sx + site_w, y + (ny - 1) * sy + site_h)
    return out


def _mark_union(boxes, xs, ys, x0, y0, x1, y1):
    """Cells of the grid of edges xs x ys (intervals [xs[i], xs[i+1]) x [ys[j], ys[j+1])) covered by the union of
    the boxes, clipped to the region. -> bool [len(xs) - 1, len(ys) - 1]."""
    cover = np.zeros((len(xs) - 1, len(ys) - 1), dtype=bool)
    if len(boxes):
        lo_x = np.searchsorted(xs, np.clip(boxes[:, 0], x0, x1))
        hi_x = np.searchsorted(xs, np.clip(boxes[:, 2], x0, x1))
        lo_y = np.searchsorted(ys, np.clip(boxes[:, 1], y0, y1))
        hi_y = np.searchsorted(ys, np.clip(boxes[:, 3], y0, y1))
        for a, b, c, e in zip(lo_x, hi_x, lo_y, hi_y):
            cover[a:b, c:e] = True
    return cover


def _interval_bin_overlap(edges, lo, hi, nb):
    """[len(edges) - 1, nb] length of each edge interval inside each of nb equal bins of [lo, hi]."""
    bins = lo + (hi - lo) * np.arange(nb + 1) / float(nb)
    left = np.maximum(edges[:-1, None], bins[None, :-1])
    right = np.minimum(edges[1:, None], bins[None, 1:])
    return np.clip(right - left, 0, None)


class PlaceData(object):
    """The arrays of one design (see the module docstring for units and pin indexing)."""

    def __init__(self, lef, d):
        self.d, self.name = d, d.design
        self.unit_dbu, self.row_height_dbu = _site_units(lef, d)
        self.row_rects = _row_boxes(lef, d)
        lo, hi = self.row_rects[:, :2].min(0), self.row_rects[:, 2:].max(0)
        self.region = (int(lo[0]), int(lo[1]), int(hi[0]), int(hi[1]))
        unit = float(self.unit_dbu)
        self.width = (self.region[2] - self.region[0]) / unit
        self.height = (self.region[3] - self.region[1]) / unit
        self.row_height = self.row_height_dbu / unit
        self._split_components(lef, d, unit)
        self._index_pins(lef, d, unit)

    def _split_components(self, lef, d, unit):
        """Movable cells and blocked area. A component stays where the DEF puts it when it is FIXED / COVER, or
        placed without being a core cell (pads, endcaps, macro blocks)."""
        masters = d.macro_names
        absent = [m for m in masters if m not in lef.macros]
        if absent:
            raise ValueError("the LEF lacks %d master(s) the DEF instantiates, e.g. %s"
                             % (len(absent), ", ".join(absent[:5])))
        dims = np.array([[lef.macros[m].width, lef.macros[m].height] for m in masters]).reshape(-1, 2) * d.dbu
        self._comp_dbu = dims[d.comp_macro_id]
        is_core = np.array([lef.macros[m].cls.startswith("CORE") for m in masters], dtype=bool)[d.comp_macro_id]
        self.fixed = d.comp_fixed | (d.comp_placed & ~is_core)
        self.kept_in_place = int(np.count_nonzero(self.fixed & ~d.comp_fixed))
        self.mov_index = np.flatnonzero(~self.fixed)
        self.num_mov = len(self.mov_index)
        self._cell_of = np.full(len(self.fixed), self.num_mov, dtype=np.int64)
        self._cell_of[self.mov_index] = np.arange(self.num_mov)
        self.size = self._comp_dbu[self.mov_index] / unit
        self.mov_area = float(self.size.prod(1).sum())
        boxes = [tuple(b) for b in d.blockages]
        for i in np.flatnonzero(self.fixed):
            fw, fh = orient_size(self._comp_dbu[i, 0], self._comp_dbu[i, 1], d.comp_orient[i])
            boxes.append((d.comp_xy[i, 0], d.comp_xy[i, 1], d.comp_xy[i, 0] + fw, d.comp_xy[i, 1] + fh))
        self.block_rects = np.array(boxes, dtype=np.int64).reshape(-1, 4)

    def _index_pins(self, lef, d, unit):
        """pin_cell / pin_off / pin_net of every connection (design pins without a location dropped)."""
        masters, names = d.macro_names, d.pin_names
        design_pin = d.pin_comp < 0
        comp = np.where(design_pin, 0, d.pin_comp)
        # each (master, pin name) the netlist uses is looked up in the LEF once
        pair = d.comp_macro_id[comp].astype(np.int64) * len(names) + d.pin_name
        uniq, which = np.unique(pair, return_inverse=True)
        where = {}
        for p in uniq.tolist():
            where[p] = lef.macros[masters[p // len(names)]].pins.get(names[p % len(names)])
        centre = np.array([where[p] if where[p] is not None else (np.nan, np.nan) for p in uniq.tolist()],
                          dtype=np.float64).reshape(-1, 2)
        xy = centre[which.reshape(-1)] * d.dbu
        lost = np.flatnonzero(np.isnan(xy[:, 0]) & ~design_pin)
        if len(lost):
            k = lost[0]
            raise ValueError("the LEF master %s has no pin %s (a connection of the DEF's nets)"
                             % (masters[d.comp_macro_id[comp[k]]], names[d.pin_name[k]]))
        origin = np.array(self.region[:2], dtype=np.float64)
        cell = np.where(design_pin, self.num_mov, self._cell_of[comp])
        off = (xy - self._comp_dbu[comp] / 2) / unit                   # from the cell centre
        if design_pin.any():
            off[design_pin] = (d.io_xy[-1 - d.pin_comp[design_pin]] - origin) / unit
        for k in np.flatnonzero(~design_pin & (cell == self.num_mov)):   # pins of fixed components: where they are
            c = comp[k]
            px, py = orient_point(xy[k, 0], xy[k, 1], self._comp_dbu[c, 0], self._comp_dbu[c, 1], d.comp_orient[c])
            off[k] = (d.comp_xy[c] + (px, py) - origin) / unit
        located = ~np.isnan(off[:, 0])
        net = np.repeat(np.arange(len(d.net_name)), np.diff(d.net_start))
        self.unplaced_io_pins = int(np.count_nonzero(~located))
        self.pin_cell, self.pin_off, self.pin_net = cell[located], off[located], net[located]
        self.num_nets, self.num_pins = len(d.net_name), int(np.count_nonzero(located))
        if self.num_mov == 0 or self.num_nets == 0:
            raise ValueError("%s has %d movable cells and %d nets: nothing to place"
                             % (d.design, self.num_mov, self.num_nets))
        self.cell_pins = np.bincount(self.pin_cell, minlength=self.num_mov + 1)[:self.num_mov].astype(np.float64)

    def fixed_fraction(self, nbx, nby):
        """[nbx, nby] share of every bin of the region that no cell may use: outside every row, or under a fixed
        component or a blockage (database.py:103-134). Exact: the union of the rows minus the union of the blockers on
        the grid of all box edges, each grid cell's area split over the bins by the separable interval overlaps."""
        x0, y0, x1, y1 = self.region
        xs, ys, usable = self.edge_grid()
        ox = _interval_bin_overlap(xs, x0, x1, nbx)
        oy = _interval_bin_overlap(ys, y0, y1, nby)
        free = ox.T.dot(usable.astype(np.float64)).dot(oy)
        bin_area = (x1 - x0) / float(nbx) * (y1 - y0) / float(nby)
        return np.clip(1.0 - free / bin_area, 0.0, 1.0)

    def edge_grid(self):
        """(xs, ys, usable): the edges of every row / blocker box clipped to the region, and which cells of that edge
        grid lie in a row and under no blocker. Depends on the design only: computed once."""
        if getattr(self, "_edge_grid", None) is None:
            x0, y0, x1, y1 = self.region
            boxes = np.concatenate([self.row_rects, self.block_rects])
            xs = np.unique(np.clip(np.concatenate([boxes[:, 0], boxes[:, 2], [x0, x1]]), x0, x1))
            ys = np.unique(np.clip(np.concatenate([boxes[:, 1], boxes[:, 3], [y0, y1]]), y0, y1))
            usable = _mark_union(self.row_rects, xs, ys, x0, y0, x1, y1) & \
                ~_mark_union(self.block_rects, xs, ys, x0, y0, x1, y1)
            self._edge_grid = (xs, ys, usable)
        return self._edge_grid

    def fixed_fraction_on(self, device, nbx, nby):
        """fixed_fraction on the device (ap_free_fraction): float64 tensor [nbx, nby], equal to the host form to
        rounding (its sums run in another order)."""
        import ctypes

        import torch

        from . import _lib
        cache = getattr(self, "_edge_grid_dev", None)
        if cache is None or cache[0] != str(device):
            xs, ys, usable = self.edge_grid()
            cache = (str(device), torch.as_tensor(xs, dtype=torch.int64, device=device),
                     torch.as_tensor(ys, dtype=torch.int64, device=device),
                     torch.as_tensor(np.ascontiguousarray(usable, dtype=np.uint8), device=device))
            self._edge_grid_dev = cache
        _, xs, ys, usable = cache
        out = torch.empty((nbx, nby), dtype=torch.float64, device=device)
        x0, y0, x1, y1 = (float(v) for v in self.region)
        p = lambda t: ctypes.c_void_p(t.data_ptr())   # noqa: E731
        stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
        _lib.check(_lib.lib().ap_free_fraction(p(xs), len(xs), p(ys), len(ys), p(usable), x0, y0, x1, y1, nbx, nby,
                                               p(out), stream), "free_fraction")
        return out

    def to_def_xy(self, pos):
        """Cell centres (placer units) -> lower-left corners in DBU, rounded to the nearest integer."""
        corner = (np.asarray(pos, dtype=np.float64) - self.size / 2) * self.unit_dbu
        return np.rint(corner + np.array(self.region[:2], dtype=np.float64)).astype(np.int64)

    def placed_pos(self, placed=None):
        """Centres (placer units) of this design's movable cells as another DEF of the same netlist places them (the
        design's own DEF by default); the movable set stays this design's."""
        d = placed or self.d
        if list(d.comp_name) != list(self.d.comp_name):
            raise ValueError("the placed DEF's components differ from the design's")
        i = self.mov_index
        missing = int(np.count_nonzero(~d.comp_placed[i]))
        if missing:
            raise ValueError("%d movable cells have no location in the placed DEF" % missing)
        swapped = np.array([_ORIENT[d.comp_orient[k]][0] for k in i], dtype=bool)
        footprint = np.where(swapped[:, None], self.size[:, ::-1], self.size)
        return (d.comp_xy[i] - np.array(self.region[:2])) / float(self.unit_dbu) + footprint / 2
