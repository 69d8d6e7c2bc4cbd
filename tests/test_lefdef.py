"""The LEF / DEF reader and the placed-DEF writer on a handwritten design with the constructs the synthetic files do
not use (comments, PROPERTYDEFINITIONS, MASK shapes, POLYGON pins, OBS, ORIGIN, turned design pins, soft and hard
placement blockages, a SPECIALNETS section, UNPLACED components, routing after a net's connections, a ROW without
DO), each expected value worked out by hand; and gp.def written from a read design keeps every other statement."""
import os
import tempfile

import numpy as np

from autoplace_amd import lefdef

LEF = """VERSION 5.8 ;
# a comment ; with a semicolon
PROPERTYDEFINITIONS
  MACRO SITE STRING ;
END PROPERTYDEFINITIONS
UNITS
  DATABASE MICRONS 1000 ;
END UNITS
SITE core
  CLASS CORE ;
  SYMMETRY Y ;
  SIZE 0.2 BY 2.0 ;
END core
MACRO INV
  CLASS CORE ;
  ORIGIN 0.1 0 ;
  SIZE 0.6 BY 2.0 ;
  SITE core ;
  PIN A
    DIRECTION INPUT ;
    PORT
      LAYER M1 ;
        RECT MASK 1 0.0 0.5 0.2 0.9 ;
    END
  END A
  PIN Z
    DIRECTION OUTPUT ;
    PORT
      LAYER M1 ;
        POLYGON 0.3 0.2 0.5 0.2 0.5 1.8 0.3 1.8 ;
    END
  END Z
  OBS
    LAYER M1 ;
      RECT 0 0 0.6 0.1 ;
  END
END INV
MACRO RAM
  CLASS BLOCK ;
  SIZE 10 BY 8 ;
  PIN D
    PORT
      LAYER M2 ;
        RECT 1 1 2 3 ;
    END
  END D
END RAM
END LIBRARY
"""

DEF = """VERSION 5.8 ;
DIVIDERCHAR "/" ;
DESIGN tiny ;
UNITS DISTANCE MICRONS 1000 ;
PROPERTYDEFINITIONS
  COMPONENTPIN text STRING ;
END PROPERTYDEFINITIONS
DIEAREA ( 0 0 ) ( 40000 20000 ) ;
ROW r0 core 0 0 N DO 200 BY 1 STEP 200 0 ;
ROW r1 core 0 2000 FS DO 200 BY 1 STEP 200 0 ;
ROW r2 core 0 4000 N ;
COMPONENTS 4 ;
- u1 INV + PLACED ( 1000 2000 ) N ;
- u2 INV + SOURCE DIST + UNPLACED ;
- mem RAM + FIXED ( 20000 4000 ) W ;
- u3 INV + PLACED ( 3000 0 ) FS + WEIGHT 2 ;
END COMPONENTS
BLOCKAGES 3 ;
- PLACEMENT RECT ( 0 10000 ) ( 4000 12000 ) ;
- PLACEMENT + SOFT RECT ( 5000 10000 ) ( 6000 12000 ) ;
- LAYER M1 RECT ( 0 0 ) ( 10 10 ) ;
END BLOCKAGES
PINS 2 ;
- in1 + NET n1 + DIRECTION INPUT + LAYER M3 ( -100 0 ) ( 100 400 ) + FIXED ( 0 9000 ) E ;
- out1 + NET n2 + DIRECTION OUTPUT + LAYER M3 ( -50 -50 ) ( 50 50 ) + PLACED ( 40000 5000 ) N ;
END PINS
SPECIALNETS 1 ;
- VDD ( * VDD ) + USE POWER ;
END SPECIALNETS
NETS 2 ;
- n1 ( PIN in1 ) ( u1 A ) ( u2 A ) ( mem D ) + USE SIGNAL
  + ROUTED M1 ( 10 10 ) ( 20 * ) ;
- n2 ( u1 Z ) ( u3 A ) ( PIN out1 ) ;
END NETS
END DESIGN
"""


def _read(tmp):
    lp, dp = os.path.join(tmp, "t.lef"), os.path.join(tmp, "t.def")
    open(lp, "w").write(LEF)
    open(dp, "w").write(DEF)
    return lefdef.parse_lef([lp]), lefdef.parse_def(dp)


def test_lef_reader_by_hand():
    with tempfile.TemporaryDirectory() as tmp:
        lef, _ = _read(tmp)
    assert lef.sites == {"core": (0.2, 2.0)}
    inv, ram = lef.macros["INV"], lef.macros["RAM"]
    assert (inv.width, inv.height, inv.cls) == (0.6, 2.0, "CORE") and ram.cls == "BLOCK"
    # A: the MASK rect's box centre (0.1, 0.7) + ORIGIN (0.1, 0); Z: the polygon's box centre (0.4, 1.0) + ORIGIN
    assert inv.pins["A"] == ((0.0 + 0.2) / 2 + 0.1, (0.5 + 0.9) / 2 + 0.0)
    assert inv.pins["Z"] == ((0.3 + 0.5) / 2 + 0.1, (0.2 + 1.8) / 2 + 0.0)
    assert ram.pins["D"] == (1.5, 2.0)
    assert set(inv.pins) == {"A", "Z"}                 # nothing from OBS


def test_def_reader_by_hand():
    with tempfile.TemporaryDirectory() as tmp:
        _, d = _read(tmp)
    assert d.design == "tiny" and d.dbu == 1000 and d.die == (0, 0, 40000, 20000)
    assert d.rows == [("core", 0, 0, 200, 1, 200, 0), ("core", 0, 2000, 200, 1, 200, 0), ("core", 0, 4000, 1, 1, 0, 0)]
    assert d.comp_name == ["u1", "u2", "mem", "u3"] and d.macro_names == ["INV", "RAM"]
    assert d.comp_macro_id.tolist() == [0, 0, 1, 0]
    assert d.comp_placed.tolist() == [True, False, True, True]
    assert d.comp_fixed.tolist() == [False, False, True, False]
    assert d.comp_xy.tolist() == [[1000, 2000], [0, 0], [20000, 4000], [3000, 0]]
    assert d.comp_orient == ["N", "N", "W", "FS"]
    assert d.blockages == [(0, 10000, 4000, 12000)] and d.soft_blockages == 1
    assert d.io_name == ["in1", "out1"]
    # in1: shape centre (0, 200) turned E: (x, y) -> (y, -x) = (200, 0), at (0, 9000); out1: centre (0, 0), N
    assert d.io_xy.tolist() == [[200.0, 9000.0], [40000.0, 5000.0]]
    assert d.net_name == ["n1", "n2"] and d.net_start.tolist() == [0, 4, 7]
    assert d.pin_comp.tolist() == [-1, 0, 1, 2, 0, 3, -2]
    assert [d.pin_names[k] for k in d.pin_name] == ["", "A", "A", "D", "Z", "A", ""]


def test_placed_def_keeps_the_rest_of_the_design():
    with tempfile.TemporaryDirectory() as tmp:
        lef, d = _read(tmp)
        out = os.path.join(tmp, "gp.def")
        lefdef.write_placed_def(d, out, np.array([1, 3]), np.array([[5000, 2000], [7000, 4000]]))
        text = open(out).read()
        d2 = lefdef.parse_def(out)
    assert d2.comp_placed.tolist() == [True, True, True, True]
    assert d2.comp_xy.tolist() == [[1000, 2000], [5000, 2000], [20000, 4000], [7000, 4000]]
    assert d2.comp_orient == ["N", "N", "W", "N"]
    assert "- u3 INV + WEIGHT 2 + PLACED ( 7000 4000 ) N ;" in text
    assert "- u2 INV + SOURCE DIST + PLACED ( 5000 2000 ) N ;" in text
    # everything outside the moved components' statements is the input's text
    for part in ("PROPERTYDEFINITIONS", "- PLACEMENT + SOFT RECT ( 5000 10000 ) ( 6000 12000 ) ;",
                 "+ ROUTED M1 ( 10 10 ) ( 20 * ) ;", "- mem RAM + FIXED ( 20000 4000 ) W ;", "END DESIGN"):
        assert part in text
    for k in ("net_start", "pin_comp", "pin_name", "io_xy"):
        assert np.array_equal(getattr(d2, k), getattr(d, k))
    assert d2.blockages == d.blockages and d2.rows == d.rows


def test_placedata_from_the_handwritten_design():
    """The problem arrays of the handwritten design: the W-turned RAM's pin lands where the orientation puts it."""
    from autoplace_amd.database import PlaceData
    with tempfile.TemporaryDirectory() as tmp:
        lef, d = _read(tmp)
    data = PlaceData(lef, d)
    assert data.num_mov == 3 and data.unit_dbu == 200 and data.row_height_dbu == 2000
    assert data.region == (0, 0, 40000, 6000)
    # mem: 10 x 8 um master turned W, at (20000, 4000); its pin D at (1500, 2000) from the master's lower left turns
    # to (h - y, x) = (8000 - 2000, 1500) in the footprint
    k = 3                                              # the fourth connection of n1
    assert data.pin_cell[k] == data.num_mov
    assert np.allclose(data.pin_off[k], [(20000 + 6000) / 200.0, (4000 + 1500) / 200.0])
    # the design pin in1 (no cell): its absolute position in sites
    assert data.pin_cell[0] == data.num_mov and np.allclose(data.pin_off[0], [1.0, 45.0])


def test_no_cpu_path():
    """gpu = -1 asks for the reference's CPU path, which this placer does not have: a clear error, not a fallback."""
    import pytest

    from autoplace_amd.problem import default_args
    from autoplace_amd.run_placement import run_placement_main
    with pytest.raises(RuntimeError, match="HIP device"):
        run_placement_main(default_args(gpu=-1, lef=["x.lef"], def_file="x.def"))
