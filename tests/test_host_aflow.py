"""Irreducible zones of the 582 lattices of the reference's AFLOW fixture (CPU only).

The fixture is the reference's own data file (wrap/tests/aflow_lattices.json,
copied unchanged): [Hall number, (a, b, c), (alpha, beta, gamma), symbol],
where the symbol is a Hall symbol or a Hermann-Mauguin symbol (full, spaced or
short).  As in wrap/tests/test_3_aflow_spacegroups.py:27-79, every lattice must
give a zone whose irreducible part holds 1/|G| of the first zone's volume.

Checks that do not come from the code under test:
* the first zone's volume is that of the primitive reciprocal cell;
* |G| agrees with the crystal system that the Hall number's range names
  (Hall numbers 1-2 triclinic, 3-107 monoclinic, 108-348 orthorhombic,
  349-429 tetragonal, 430-461 trigonal, 462-488 hexagonal, 489-530 cubic)
  and is the same for every entry of one Hall number;
* where the symbol names a default setting, its ITA number rises with the
  Hall number (the Hall table is ordered by ITA number) and |G| is the point
  group order of that ITA number.

Two entries give Hall number 460 (R-3c) the symbol '-R 3 2c'.  Read as a Hall
symbol, as the reference reads any string its table does not hold
(src/lattice_dual.hpp:510-526), its second generator is a 2-fold along a-b
(the implicit axis of a 2 after a 3, src/hall_symbol.cpp:220-223), which is
not a symmetry of the rhombohedral lattice: no irreducible zone has volume
1/12 of the first, and the construction raises, as the reference's check
(src/bz.cpp:69-80) reports a wrong volume.  The Hall symbol of R-3c,
'-R 3 2"c', builds the zone on the same two lattices.
"""
import json
import os

import numpy as np
import pytest

import brille_amd as b
from test_host import _pg_order

HERE = os.path.dirname(os.path.abspath(__file__))
AFLOW = json.load(open(os.path.join(HERE, "golden", "aflow_lattices.json")))
SYSTEM_ORDERS = [((1, 2), {1, 2}), ((3, 107), {2, 4}), ((108, 348), {4, 8}), ((349, 429), {4, 8, 16}),
                 ((430, 461), {3, 6, 12}), ((462, 488), {6, 12, 24}), ((489, 530), {12, 24, 48})]
MISQUOTED = "-R 3 2c"


def _orders(hall):
    for (lo, hi), orders in SYSTEM_ORDERS:
        if lo <= hall <= hi:
            return orders
    raise AssertionError(hall)


def _point_order(symbol):
    return len({np.asarray(w).tobytes() for w in b.Symmetry(symbol).W})


def _zone(afl, symbol=None):
    lat = b.Direct(np.array(afl[1], float), np.array(afl[2], float), symbol or afl[3])
    return lat, b.BrillouinZone(lat)


def test_fixture_is_the_reference_file():
    assert len(AFLOW) == 582
    assert len({a[0] for a in AFLOW}) == 230


def test_point_group_orders_follow_the_hall_numbers():
    seen, last = {}, (0, 0)
    for hall, _, _, sym in sorted(AFLOW, key=lambda a: a[0]):
        order = _point_order(sym)
        assert order in _orders(hall), (hall, sym, order)
        assert seen.setdefault(hall, order) == order, (hall, sym)
        ita = b.spacegroup_number(sym)
        if ita:
            assert order == _pg_order(ita), (hall, sym, ita)
            assert ita >= last[1] or hall == last[0], (hall, sym, ita, last)
            last = (hall, ita)


@pytest.mark.parametrize("k", range(0, len(AFLOW), 97))
def test_chunk_of_lattices_has_the_right_irreducible_volume(k):
    for afl in AFLOW[k:k + 97]:
        if afl[3] == MISQUOTED:
            continue
        lat, bz = _zone(afl)
        order = _point_order(afl[3])
        conv = abs(np.linalg.det(np.asarray(lat.reciprocal_vectors)))
        cent = {1: 1, 2: 2, 3: 2, 4: 2, 5: 2, 6: 4, 7: 3}[lat.bravais]
        assert np.isclose(bz.polyhedron.volume, conv * cent, rtol=1e-10), afl
        assert np.isclose(bz.ir_polyhedron.volume * order, bz.polyhedron.volume, rtol=1e-8), afl


def test_misquoted_r3c_symbol():
    bad = [a for a in AFLOW if a[3] == MISQUOTED]
    assert len(bad) == 2 and all(a[0] == 460 for a in bad)
    for afl in bad:
        with pytest.raises(RuntimeError, match="irreducible"):
            _zone(afl)
        lat, bz = _zone(afl, '-R 3 2"c')
        assert _point_order('-R 3 2"c') == 12
        assert np.isclose(bz.ir_polyhedron.volume * 12, bz.polyhedron.volume, rtol=1e-8)


@pytest.mark.parametrize("symbol,ita", [
    ("P 1 2_1/c 1", 14), ("P2_1/c", 14), ("P21/c", 14), ("P 2_1/n 2_1/m 2_1/a", 62), ("P n m a", 62),
    ("C 2/m 2/c 2_1/e", 64), ("P 4/m 2_1/b m", 127), ("P 4_2/m 2_1/n 2/m", 136), ("I 4_1/a 2/m 2/d", 141),
    ("P -3 2/m 1", 164), ("R -3 2/c", 167), ("P 6_3/m 2/m 2/c", 194), ("P63/mmc", 194), ("P 2/m -3", 200),
    ("F 4_1/d -3 2/m", 227), ("I 4_1/a -3 2/d", 230), ("I-4", 82), ("P 3_1 2 1", 152), ("P 3 1 2", 149),
    ("P 4_2", 77), ("-P 4 2", 123), ("P 4 2", 89), ("194", 194)])
def test_hermann_mauguin_names(symbol, ita):
    assert b.spacegroup_number(symbol) == ita
    assert len(b.Symmetry(symbol)) == len(b.Symmetry(str(ita)))
    assert _point_order(symbol) == _pg_order(ita)


def test_other_settings_are_hall_symbols_or_errors():
    # "P 2" is the c unique-axis Hall symbol, as in the reference's table
    W = b.Symmetry("P 2").W
    assert any(np.array_equal(w, np.diag([-1, -1, 1])) for w in W)
    assert b.spacegroup_number("P 2") == 3
    with pytest.raises(RuntimeError):
        b.Symmetry("P b n m")  # an orthorhombic setting the default table does not hold
