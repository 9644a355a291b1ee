"""Read ranges of the boundary kernel's stencils (EdgeEnv in
nr-amssncku_amd/csrc/bssn_env.h), enumerated on the CPU for every point of
the 3-deep shell of every array shape with 5 to 9 points per axis, without
a symmetry plane.

The reference serves reads below index 0 from symmetry_bd's mirrored halo
(fmisc.f90:877-901) for every Symmetry, and its stencils do read there: the
upwind stencil of a positive shift at the first points (lopsidediff.f90:
324-374) and the second-order fallbacks of short axes.  So the kernel keeps
its mirror test on such reads.  What this pins: which shells read below 0 or
past n - 1, so that a kernel may drop the test only where no read leaves
the array -- the points whose centred stencils are of fourth or second
order (ORD 2 and 1); the outermost layer (ORD 0) keeps it.
"""
import itertools

import pytest


def _room_order(p, n):
    """The joint stencil order at point p of an array of shape n
    (diff_new.f90:1084-1096): 2 with +-2 room on every axis, 1 with +-1, else 0."""
    r = min(min(p[a], n[a] - 1 - p[a]) for a in range(3))
    return 2 if r >= 2 else (1 if r >= 1 else 0)


def _reads(p, n, order):
    """Every (axis, offset) the kernel reads at p, for shifts of either sign
    (the upwind branch depends on the sign of beta at run time)."""
    out = set()
    last = any(p[a] == n[a] - 1 for a in range(3))
    for a in range(3):
        # grad: fourth order -2..2, second order -1..1
        if order == 2:
            out |= {(a, m) for m in (-2, -1, 1, 2)}
        elif order == 1:
            out |= {(a, m) for m in (-1, 1)}
    if order == 2:  # Hessian: diagonal -2..2, mixed -2..2 on both axes
        for a in range(3):
            out |= {(a, m) for m in (-2, -1, 0, 1, 2)}
        for a, b in ((0, 1), (0, 2), (1, 2)):
            out |= {((a, ma), (b, mb)) for ma in (-2, -1, 1, 2) for mb in (-2, -1, 1, 2)}
    elif order == 1:
        for a in range(3):
            out |= {(a, m) for m in (-1, 0, 1)}
        for a, b in ((0, 1), (0, 2), (1, 2)):
            out |= {((a, ma), (b, mb)) for ma in (-1, 1) for mb in (-1, 1)}
    if not last:  # lopsidediff.f90:324-374, both signs of beta
        for a in range(3):
            pos, nn = p[a], n[a]
            # beta > 0
            if pos + 3 <= nn - 1:
                out |= {(a, m) for m in (-1, 0, 1, 2, 3)}
            elif pos + 2 <= nn - 1:
                out |= {(a, m) for m in (-2, -1, 1, 2)}
            elif pos + 1 <= nn - 1:
                out |= {(a, m) for m in (1, 0, -1, -2, -3)}
            # beta < 0 (lowest readable index 0 without a symmetry plane)
            if pos - 3 >= 0:
                out |= {(a, m) for m in (1, 0, -1, -2, -3)}
            elif pos - 2 >= 0:
                out |= {(a, m) for m in (-2, -1, 1, 2)}
            elif pos - 1 >= 0:
                out |= {(a, m) for m in (-1, 0, 1, 2, 3)}
    return out


def _out_of_range(p, n, reads):
    bad = []
    for r in reads:
        parts = r if isinstance(r[0], tuple) else (r,)
        q = list(p)
        for a, m in parts:
            q[a] += m
        if any(q[a] < 0 or q[a] > n[a] - 1 for a in range(3)):
            bad.append(tuple(q))
    return bad


SHAPES = list(itertools.product(range(5, 10), repeat=3))


def _shell(n):
    for p in itertools.product(*(range(x) for x in n)):
        if min(min(p[a], n[a] - 1 - p[a]) for a in range(3)) < 3:
            yield p


@pytest.mark.parametrize("order", [2, 1])
def test_fourth_and_second_order_shell_reads_stay_inside_every_array(order):
    checked = 0
    for n in SHAPES:
        for p in _shell(n):
            if _room_order(p, n) == order:
                assert not _out_of_range(p, n, _reads(p, n, order)), (n, p)
                checked += 1
    assert checked > 3000


def test_zero_order_shell_reads_the_mirrored_layer():
    """The outermost layer reads below index 0 (the upwind stencil of a
    positive shift at the first point): its mirror test must stay."""
    below = 0
    for n in SHAPES[:30]:
        for p in _shell(n):
            if _room_order(p, n) == 0:
                below += any(min(q) < 0 for q in _out_of_range(p, n, _reads(p, n, 0)))
    assert below > 0


@pytest.mark.parametrize("n", [(5, 5, 5), (6, 7, 8), (9, 5, 6)])
def test_every_shell_point_has_an_order(n):
    orders = {_room_order(p, n) for p in _shell(n)}
    assert orders <= {0, 1, 2} and 0 in orders
