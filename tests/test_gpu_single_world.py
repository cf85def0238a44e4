"""The single-world entry points (arc_gpu_world.h) beyond the reference's
known-answer tests (those run in test_kat.py): hand-built worlds stepped on
the GPU in lockstep with the FP32 oracle, bit for bit after every step, where
the 128-record contact cap binds (the kernel's serial collision path, with the
spheres of two bodies interleaved), for many steps in one call (several
launches), on generated terrain; and the builders' and the step's refusals.
"""

import math

import numpy as np
import pytest

import pyoracle as P
from helpers import TwinWorld, assert_same

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from arc import native
    return native


def _plates(w, L, z, cell=0.02, difficulty=0.3):
    """two plates of 24 spheres each, the spheres' order alternating between
    the plates, over rough terrain of 2 cm columns: every sphere's cell window
    holds 144 columns, far beyond the cap, so the kernel records contacts in
    the world's sphere order"""
    w.terrain_generate(64, cell, difficulty, 3)
    he = (0.5, 0.3, 0.05)
    a = w.add_body(6.0, L.inertia(2, 6.0, *he), (0.01, -0.37, z))
    b = w.add_body(6.0, L.inertia(2, 6.0, *he), (0.02, 0.33, z))
    for ix in range(8):
        for iy in range(3):
            for body in (a, b):
                w.add_sphere(body, (-0.42 + 0.12 * ix, -0.2 + 0.2 * iy, -0.05), 0.1)


def _compare(t: TwinWorld):
    """every body and sphere, the flag and the clock (the twin asserts each)"""
    for i in range(t.g.counts()[0]):
        t.body(i)
    for k in range(t.g.counts()[2]):
        t.sphere(k)
    assert t.unstable in (0, 1) and t.time > 0.0


@pytest.mark.parametrize("z", [0.1, 0.2])
def test_contact_cap_serial_order(z):
    """At z = 0.1 the first substep records 128 contacts (the cap binds, as
    the oracle's count confirms); at 0.2 the plates land and rest.  The state
    equals the oracle's after every step either way."""
    L = P.Lib("o32")
    t = TwinWorld()
    _plates(t, L, z)
    caps = []
    for _ in range(3):
        t.step(1.0 / 600.0, 1, 1)
        caps.append(int(t.o.counts()[3]))
        _compare(t)
    for _ in range(30):
        t.step(1.0 / 30.0, 20, 1)
        _compare(t)
    print(f"\nz {z}: contacts of the first substeps {caps}, unstable {t.unstable}")
    if z == 0.1:
        assert caps[0] == 128
    else:
        assert not t.unstable


def test_many_steps_per_call():
    """600 steps in one call (three launches of at most 256 steps) give the
    bits of 600 calls of one step, and the oracle's."""
    L = P.Lib("o32")
    ws = [TwinWorld(), TwinWorld()]
    for w in ws:
        w.terrain_flat(1, 1, 1.0, -1e6)
        n, seg = 6, 0.3
        half = math.sqrt(0.5)
        for i in range(n):
            m = 1.0 - 0.9 * i / (n - 1.0)
            w.add_body(m, L.inertia(1, m, 0.03, 0.5 * seg), ((i + 0.5) * seg, 0, 0), (half, 0, half, 0))
            w.add_hinge(i - 1, i, (i * seg, 0, 0), (0, 1, 0))
            w.joint_set(i, -0.5, 0.5, 20.0, 0.5, 5.0, 0.1 * i)
    ws[0].step(1.0 / 30.0, 20, 600)
    for _ in range(600):
        ws[1].step(1.0 / 30.0, 20, 1)
    for i in range(6):
        assert_same(ws[0].body(i), ws[1].body(i), f"body {i}")
        assert_same(ws[0].joint_out(i), ws[1].joint_out(i), f"joint {i}")
    assert ws[0].time == ws[1].time


@pytest.mark.parametrize("n,cell,difficulty,seed", [(32, 0.4, 0.5, 9), (64, 0.25, 1.0, 0), (70, 0.1, 1.7, 2 ** 32 - 1),
                                                    (16, 0.5, 0.0, 4), (8, 0.3, -1.0, 5)])
def test_generated_terrain(n, cell, difficulty, seed):
    """arc_terrain_generate's columns at every cell centre and beyond the
    border, and the grid's description, against the oracle's."""
    t = TwinWorld()
    t.terrain_generate(n, cell, difficulty, seed)
    info = t.terrain_info()
    nx = int(info[0])
    assert nx == min(n, 64)
    xs = info[3] + (np.arange(-1, nx + 1) + 0.5) * info[2]
    for x in xs:
        for y in xs:
            t.height(float(x), float(y))


def test_refusals(native):
    """Invalid arguments raise ValueError with the library's message; a world
    beyond the device's bounds is refused before any launch."""
    w = native.World()
    with pytest.raises(ValueError, match="cell"):
        w.terrain_flat(4, 4, 0.0, 0.0)
    with pytest.raises(ValueError, match="cell"):
        w.terrain_generate(4, 1e-5, 1.0, 0)
    with pytest.raises(ValueError, match="no such body"):
        w.add_sphere(0, (0, 0, 0), 0.1)
    b = w.add_body(1.0, (0.1, 0.1, 0.1), (0, 0, 1))
    assert b == 0
    with pytest.raises(ValueError, match="no such body"):
        w.add_hinge(-1, 1, (0, 0, 0), (0, 1, 0))
    with pytest.raises(ValueError):
        w.add_hinge(0, 0, (0, 0, 0), (0, 1, 0))
    with pytest.raises(ValueError, match="no such joint"):
        w.joint_set(0, -1, 1)
    with pytest.raises(ValueError, match="column index"):
        w.terrain_set(1, 0.5)
    with pytest.raises(ValueError):
        w.step(0.0, 20, 1)
    with pytest.raises(ValueError):
        w.step(1.0 / 30.0, 0, 1)
    w.add_sphere(0, (0, 0, 0), 100.0)
    with pytest.raises(ValueError, match="window"):
        w.step(1.0 / 30.0, 20, 1)
    for _ in range(15):
        w.add_body(1.0, (0.1, 0.1, 0.1), (0, 0, 1))
    with pytest.raises(ValueError, match="16 bodies"):
        w.add_body(1.0, (0.1, 0.1, 0.1), (0, 0, 1))
    assert list(w.counts()) == [16, 0, 1]
    assert w.time == 0.0
