"""arc_sim_world, the host snapshot the reference's renderer and viewer read
(csrc/arc_render.c:1041, csrc/tools/arcview.c:46), field by field against
the reference engine built from its own sources (oracle/_ref):

- the state a step produces (bodies, their cached rotations and world
  inverse inertias, joints, spheres with touch counts and friction anchors,
  time, instability, terrain) and the contacts of the last substep, in the
  reference's record order, against the reference compiled in FP32, on the
  same seeds and targets: equal bit for bit after the rounding of the FP32
  build's floats to the snapshot's doubles, in every form;
- the robot's 74 visuals against the FP64 reference build: equal exactly;
- a reset world reports no contacts.
The cached rotation and world inverse inertia of each body (rot, inv_iw) are
compared within 16 binary32 ulp of each matrix's largest entry (an ulp or
two in the rotation, scaled by the largest inverse inertia in inv_iw): the
reference's cache after a step holds those of the orientation before the
last substep's normalization (its
velocity update normalizes q without refreshing them, csrc/arc_phys.c:674-687),
the snapshot derives them from the normalized orientation it stores, as the
reference's next step does first (csrc/arc_phys.c:712).  Not compared:
x_prev / q_prev (the pose at the start of the last substep) and torque_acc
(the torque sum of the last step), which the device does not keep after the
step; the renderer reads neither.
"""

import ctypes

import numpy as np
import pytest

import pyoracle as P
from helpers import action_stream, struct_values, world_type

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (P.available("r32") and P.available("r64s")),
                                 reason="reference builds (oracle/_ref) absent")]

W64 = world_type(ctypes.c_double)
W32 = world_type(ctypes.c_float)
SKIP = ("x_prev", "q_prev", "torque_acc", "visuals", "contacts", "h")


@pytest.fixture(scope="module")
def native():
    from arc import native
    return native


def _snapshot(native, g, i):
    p = native.sim_lib().arc_sim_world(g.ptr, i)
    assert p
    return W64.from_address(p)


def _reference_world(lib, sim, i, W):
    return W.from_address(lib.lib.orc_sim_world(sim.p, i))


@pytest.mark.parametrize("form", ["leg", "lane", "pair", "shared", "wave"])
def test_snapshot_matches_reference(native, form):
    n, steps = 24, 4
    ids = np.arange(n, dtype=np.int32)
    diff = P.difficulties(7, n)
    g = native.Sim(n)
    g.set_form(form)
    ref = P.Lib("r32")
    r = ref.sim(n)
    for s in (g, r):
        s.reset(ids, ids.astype(np.uint32), diff)
    ts = action_stream(n, steps, seed=11)
    for k in range(steps):
        g.step(ts[k])
        r.step(ts[k])
    contacts, worst = 0, 0.0
    for i in range(n):
        a = _snapshot(native, g, i)
        b = _reference_world(ref, r, i, W32)
        va, vb = struct_values(a, SKIP), struct_values(b, SKIP)
        assert va.keys() == vb.keys()
        cache = [k for k in va if ".rot." in k or ".inv_iw." in k]
        for k in range(0, len(cache), 9):  # one 3 x 3 matrix at a time
            x = np.array([va[c] for c in cache[k:k + 9]])
            y = np.array([vb[c] for c in cache[k:k + 9]], np.float64)
            ulps = np.max(np.abs(x - y)) / np.spacing(np.float32(np.max(np.abs(y))))
            worst = max(worst, float(ulps))
            assert ulps <= 16, f"world {i}: {cache[k]}: {ulps:.1f} ulp"
        bad = [k for k in va if k not in cache and not (va[k] == vb[k] or (va[k] != va[k] and vb[k] != vb[k]))]
        assert not bad, f"world {i}: {len(bad)} fields differ, first {bad[0]}: {va[bad[0]]!r} vs {vb[bad[0]]!r}"
        cells = a.terrain.nx * a.terrain.ny
        assert np.array_equal(np.array(a.terrain.h[:cells]), np.array(b.terrain.h[:cells], np.float64))
        assert a.n_contacts == b.n_contacts, f"world {i}"
        for k in range(a.n_contacts):
            assert struct_values(a.contacts[k]) == struct_values(b.contacts[k]), f"world {i} contact {k}"
        contacts += a.n_contacts
    print(f"\n{form}: {n} worlds after {steps} steps, {contacts} contacts, snapshots equal the reference's; "
          f"cached rotations within {worst:.1f} ulp of the largest entry")
    assert contacts > 0


def test_visuals_match_reference(native):
    g = native.Sim(2)
    ref = P.Lib("r64s")
    r = ref.sim(2)
    a = _snapshot(native, g, 1)
    b = _reference_world(ref, r, 1, W64)
    assert a.n_visuals == b.n_visuals == 74
    for k in range(74):
        assert struct_values(a.visuals[k]) == struct_values(b.visuals[k]), f"visual {k}"


def test_reset_world_has_no_contacts(native):
    n = 8
    ids = np.arange(n, dtype=np.int32)
    g = native.Sim(n)
    g.reset(ids, ids.astype(np.uint32), np.full(n, 0.5))
    for t in action_stream(n, 3, seed=2):
        g.step(t)
    assert sum(_snapshot(native, g, i).n_contacts for i in range(n)) > 0
    g.reset(ids[:4], ids[:4].astype(np.uint32), np.full(4, 0.5))
    assert [_snapshot(native, g, i).n_contacts for i in range(4)] == [0] * 4
