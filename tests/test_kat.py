"""The reference's own known-answer tests and invariants: csrc/tests/test_phys.c
(analytic physics cases) and csrc/tests/test_robot.c (standing, random
actions, determinism, void), each run on

- "o64": the FP64 oracle, with the reference's tolerances;
- "o32": the FP32 oracle (the precision the HIP engine computes in), which
  passes the same checks except where the reference's tolerance is below
  binary32 resolution (1e-9 / 1e-12 on quantities of order 1, and vz after
  600 substeps of position-derived velocity); those use the tolerance stated
  next to them.  The reference's own FP32 build shows the same six failures
  (oracle/Makefile `make ref` builds it);
- "gpu" (marked gpu): the HIP engine -- the single-world entry points of
  arc_gpu_world.h for the physics cases, the arc_sim ABI for the robot -- with
  the FP32 tolerances, driven in lockstep with the FP32 oracle: every value a
  check reads is also asserted equal, bit for bit, to the oracle's.
"""

import math

import numpy as np
import pytest

import pyoracle as P
from helpers import TwinSim, TwinWorld

DT, SUB = 1.0 / 30.0, 20
PREC = ["o64", "o32", pytest.param("gpu", marks=pytest.mark.gpu)]


def tol(kind, t64, t32):
    return t64 if kind == "o64" else t32


class GpuLib:
    """The "gpu" kind: worlds and sims twinned with the FP32 oracle; the
    inertia helpers are the FP32 oracle's (host-side arithmetic)."""

    def world(self):
        return TwinWorld()

    def sim(self, n, grid=64, cell=0.4):
        return TwinSim(n, grid=grid, cell=cell)

    def inertia(self, *args):
        return P.Lib("o32").inertia(*args)


def lib(kind):
    return GpuLib() if kind == "gpu" else P.Lib(kind)


def empty_world(L):
    w = L.world()
    w.terrain_flat(1, 1, 1.0, -1e6)
    return w


def run(w, seconds):
    w.step(DT, SUB, int(seconds / DT + 0.5))


@pytest.mark.parametrize("kind", PREC)
def test_free_fall(kind):
    """test_phys.c:23-35: z(1 s) = -g/2 within the symplectic-Euler error."""
    L = lib(kind)
    w = empty_world(L)
    w.add_body(2.0, L.inertia(0, 2.0, 0.1), (0, 0, 0))
    st = w.body(0)
    st[7] = 1.0
    w.set_body(0, st)
    run(w, 1.0)
    b = w.body(0)
    assert abs(b[2] - (-0.5 * 9.81)) <= 1e-2
    assert abs(b[0] - 1.0) <= tol(kind, 1e-9, 2e-6)
    assert abs(b[9] - (-9.81)) <= tol(kind, 1e-9, 1e-2)


@pytest.mark.parametrize("kind", PREC)
def test_free_rotation(kind):
    """test_phys.c:37-59: angular momentum drift <= 2 %, no energy gain."""
    L = lib(kind)
    w = empty_world(L)
    w.gravity(0, 0, 0)
    w.add_body(1.0, (1.0, 2.0, 3.0), (0, 0, 0))
    st = w.body(0)
    st[10:13] = (0.2, 0.3, 4.0)
    w.set_body(0, st)

    def L_world():
        b = w.body(0)
        q, om = b[3:7], b[10:13]
        R = _rot(q)
        return R @ (np.array([1.0, 2.0, 3.0]) * (R.T @ om))

    L0, e0 = L_world(), w.energy()
    run(w, 5.0)
    L1, e1 = L_world(), w.energy()
    assert np.linalg.norm(L1 - L0) / np.linalg.norm(L0) <= 0.02
    assert (e1 - e0) / e0 <= tol(kind, 1e-6, 1e-6)
    assert (e0 - e1) / e0 <= 0.05


def _rot(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("kind", PREC)
def test_pendulum(kind):
    """test_phys.c:61-101: small-angle period 2 pi sqrt(I / (m g l)) within 0.2 %."""
    L = lib(kind)
    w = empty_world(L)
    ln, m, r, th0 = 1.0, 1.0, 0.05, 0.1
    w.add_body(m, L.inertia(0, m, r), (0, 0, -ln))
    w.add_hinge(-1, 0, (0, 0, 0), (0, 1, 0))
    w.set_pose((0, 0, 0), (1, 0, 0, 0), [th0])
    assert abs(w.body(0)[0] - (-ln * math.sin(th0))) <= tol(kind, 1e-12, 1e-7)
    i_piv = 0.4 * m * r * r + m * ln * ln
    period = 2 * math.pi * math.sqrt(i_piv / (m * 9.81 * ln))
    e0 = w.energy()
    a_prev = w.joint_angle(0)
    assert abs(a_prev - th0) <= tol(kind, 1e-9, 1e-7)
    first, last, cross = -1.0, 0.0, 0
    for _ in range(6000):
        w.step(1.0 / 600.0, 1, 1)
        a = w.joint_angle(0)
        if a_prev > 0.0 and a <= 0.0:
            t = w.time - (1.0 / 600.0) * a / (a - a_prev)
            if first < 0:
                first = t
            last = t
            cross += 1
        a_prev = a
    assert abs((last - first) / (cross - 1) - period) <= 2e-3 * period
    swing = m * 9.81 * ln * (1 - math.cos(th0))
    assert (w.energy() - e0) / swing <= tol(kind, 1e-6, 1e-3)
    assert (e0 - w.energy()) / swing <= 0.2


@pytest.mark.parametrize("kind", PREC)
def test_chain(kind):
    """test_phys.c:103-131: 6-link chain, 10:1 masses, joints stay within 2 mm."""
    L = lib(kind)
    w = empty_world(L)
    n, seg = 6, 0.3
    half = math.sqrt(0.5)
    for i in range(n):
        m = 1.0 - 0.9 * i / (n - 1.0)
        w.add_body(m, L.inertia(1, m, 0.03, 0.5 * seg), ((i + 0.5) * seg, 0, 0), (half, 0, half, 0))
        w.add_hinge(i - 1, i, (i * seg, 0, 0), (0, 1, 0))
    maxsep = 0.0
    for _ in range(90):
        w.step(DT, SUB, 1)
        maxsep = max(maxsep, _chain_sep([w.body(i) for i in range(n)], seg))
    assert maxsep <= 2e-3
    assert not w.unstable
    assert w.body(n - 1)[2] <= 1e-6


def _chain_sep(bodies, seg):
    """Largest gap at the hinges: link k's start (0, 0, -seg/2 in its frame)
    against link k-1's end (0, 0, +seg/2), or the world origin for k = 0 --
    the anchors arc_add_hinge derives for this chain (links lie along body z)."""
    half = np.array([0.0, 0.0, 0.5 * seg])
    worst = 0.0
    for k, b in enumerate(bodies):
        start = b[:3] - _rot(b[3:7]) @ half
        if k == 0:
            end = np.zeros(3)
        else:
            a = bodies[k - 1]
            end = a[:3] + _rot(a[3:7]) @ half
        worst = max(worst, float(np.linalg.norm(end - start)))
    return worst


@pytest.mark.parametrize("kind", PREC)
def test_limits(kind):
    """test_phys.c:133-154: overshoot <= 5e-3 rad, rests on the upper limit."""
    L = lib(kind)
    w = empty_world(L)
    half = math.sqrt(0.5)
    w.add_body(1.0, L.inertia(1, 1.0, 0.03, 0.25), (0.25, 0, 0), (half, 0, half, 0))
    w.add_hinge(-1, 0, (0, 0, 0), (0, 1, 0))
    lo, hi = -0.3, 0.4
    w.joint_set(0, lo, hi)
    amin, amax = 1e9, -1e9
    for _ in range(120):
        w.step(DT, SUB, 1)
        a = w.joint_out(0)[0]
        amin, amax = min(amin, a), max(amax, a)
    assert amax - hi <= 5e-3 and lo - amin <= 5e-3
    assert abs(w.joint_out(0)[0] - hi) <= 5e-3


@pytest.mark.parametrize("kind", PREC)
def test_drive(kind):
    """test_phys.c:156-186: PD drive reaches its target, static sag
    m g (L/2) / kp, holding torque, saturation of a weak motor."""
    L = lib(kind)
    w = empty_world(L)
    w.gravity(0, 0, 0)
    m, ln, target = 1.0, 0.5, 0.5
    half = math.sqrt(0.5)
    w.add_body(m, L.inertia(1, m, 0.03, 0.5 * ln), (0.5 * ln, 0, 0), (half, 0, half, 0))
    w.add_hinge(-1, 0, (0, 0, 0), (0, 1, 0))
    w.joint_set(0, 1.0, -1.0, 50.0, 2.0, 100.0, target)
    run(w, 2.0)
    a, rate, _ = w.joint_out(0)
    assert abs(a - target) <= 1e-3 and abs(rate) <= 1e-3
    w.gravity(0, 0, -9.81)
    w.joint_set(0, 1.0, -1.0, 50.0, 2.0, 100.0, 0.0)
    run(w, 3.0)
    a, _, torque = w.joint_out(0)
    assert abs(a - m * 9.81 * 0.5 * ln / 50.0) <= 2e-3
    assert abs(torque - (-m * 9.81 * 0.5 * ln)) <= 0.05
    w.joint_set(0, 1.0, -1.0, 50.0, 2.0, 0.2, 0.0)
    run(w, 6.0)
    a, _, torque = w.joint_out(0)
    assert abs(abs(torque) - 0.2) <= 0.02
    assert a > 1.2


@pytest.mark.parametrize("kind", PREC)
def test_resting_sphere(kind):
    """test_phys.c:188-206: rests at height r, still, touching every substep."""
    L = lib(kind)
    w = L.world()
    w.terrain_flat(8, 8, 0.5, 0.0)
    r = 0.1
    w.add_body(1.0, L.inertia(0, 1.0, r), (0.01, 0.02, 0.5))
    w.add_sphere(0, (0, 0, 0), r)
    run(w, 1.0)
    maxv = 0.0
    for _ in range(60):
        w.step(DT, SUB, 1)
        maxv = max(maxv, float(np.linalg.norm(w.body(0)[7:10])))
    assert abs(w.body(0)[2] - r) <= 1e-4
    assert maxv <= 1e-3
    assert w.sphere(0)[0] == SUB


@pytest.mark.parametrize("kind", PREC)
def test_sliding(kind):
    """test_phys.c:217-249: stopping distance v^2 / (2 mu g); static hold on
    a 20 degree tilt; slides once mu < tan(20 deg)."""
    L = lib(kind)
    w = L.world()
    v0, mu = 3.0, 0.5
    w.mu(mu, mu)
    w.terrain_flat(32, 32, 0.5, 0.0)
    he, r = (0.2, 0.2, 0.05), 0.03
    b = w.add_body(2.0, L.inertia(2, 2.0, *he), (-4.0, 0, 0.05))
    for sx in (-1, 1):
        for sy in (-1, 1):
            w.add_sphere(b, (sx * (he[0] - r), sy * (he[1] - r), -he[2] + r), r)
    run(w, 0.2)
    x0 = w.body(0)[0]
    st = w.body(0)
    st[7:10] = (v0, 0, 0)
    w.set_body(0, st)
    run(w, 2.0)
    assert abs(w.body(0)[0] - x0 - v0 * v0 / (2 * mu * 9.81)) <= 0.05
    assert np.linalg.norm(w.body(0)[7:10]) <= 5e-3
    x1, ang = w.body(0)[0], math.radians(20.0)
    w.gravity(9.81 * math.sin(ang), 0, -9.81 * math.cos(ang))
    run(w, 2.0)
    assert abs(w.body(0)[0] - x1) <= 5e-3
    w.mu(0.3, 0.3)
    run(w, 1.0)
    assert w.body(0)[0] - x1 > 0.25


@pytest.mark.parametrize("kind", PREC)
def test_blocks(kind):
    """test_phys.c:251-283: lookups, seam, step edge and edge-resting spheres."""
    L = lib(kind)
    w = L.world()
    w.terrain_flat(8, 8, 0.5, 0.0)
    for i in range(64):
        if i % 8 >= 4:
            w.terrain_set(i, 0.3)
    assert abs(w.height(0.1, 0.1) - 0.3) <= tol(kind, 1e-12, 1e-7)
    assert abs(w.height(-0.1, 0.1)) <= 1e-12
    r = 0.1
    for pos in ((0.5, 0.0, 0.6), (-0.12, 0.25, 0.6), (0.02, -0.75, 0.6)):
        b = w.add_body(1.0, L.inertia(0, 1.0, r), pos)
        w.add_sphere(b, (0, 0, 0), r)
    run(w, 3.0)
    b0, b1, b2 = w.body(0), w.body(1), w.body(2)
    assert abs(b0[2] - (0.3 + r)) <= 1e-4 and abs(b0[0] - 0.5) <= 1e-4
    assert abs(b1[2] - r) <= 1e-4 and b1[0] <= -r + 1e-4
    assert abs(b2[2] - (0.3 + r)) <= 1e-3
    assert np.linalg.norm(b2[7:10]) <= 1e-3


@pytest.mark.parametrize("kind", PREC)
def test_set_pose(kind):
    """test_phys.c:285-302: forward kinematics reads back its angles."""
    L = lib(kind)
    w = empty_world(L)
    w.add_body(5.0, (1, 1, 1), (0, 0, 0))
    w.add_body(1.0, (.1, .1, .1), (0.2, 0, -0.2))
    w.add_body(1.0, (.1, .1, .1), (0.2, 0, -0.6))
    w.add_body(1.0, (.1, .1, .1), (0.2, 0.1, -1.0))
    w.add_hinge(0, 1, (0.2, 0, 0), (1, 0, 0))
    w.add_hinge(1, 2, (0.2, 0, -0.4), (0, 1, 0))
    w.add_hinge(2, 3, (0.2, 0, -0.8), (0, 0.6, 0.8))
    angles = [0.3, -0.7, 1.1]
    w.set_pose((1, 2, 3), (math.cos(0.4), 0, 0, math.sin(0.4)), angles)
    for i, a in enumerate(angles):
        assert abs(w.joint_angle(i) - a) <= tol(kind, 1e-9, 1e-6)


@pytest.mark.parametrize("kind", PREC)
@pytest.mark.parametrize("bad", [1e12, -1e300, float("nan")])
def test_blow_up(kind, bad):
    """test_phys.c:304-324: far or NaN states are flagged, lookups stay safe."""
    L = lib(kind)
    w = L.world()
    w.terrain_flat(8, 8, 0.5, 0.0)
    w.add_body(1.0, L.inertia(0, 1.0, 0.1), (bad, 0.0, 0.05))
    w.add_sphere(0, (0, 0, 0), 0.1)
    w.step(DT, SUB, 1)
    assert w.unstable
    assert w.height(bad, bad) == 0.0


@pytest.mark.parametrize("kind", PREC)
def test_void(kind):
    """test_robot.c:168-188: a ball rolled off the generated terrain falls."""
    L = lib(kind)
    w = L.world()
    w.terrain_generate(32, 0.4, 0.5, 9)
    info = w.terrain_info()
    edge = info[3] + info[0] * info[2]
    r = 0.1
    w.add_body(1.0, L.inertia(0, 1.0, r), (edge - 1.0, 0.0, 1.0))
    w.add_sphere(0, (0, 0, 0), r)
    st = w.body(0)
    st[7] = 4.0
    w.set_body(0, st)
    w.step(DT, SUB, 90)
    assert info[5] < -10.0
    assert w.body(0)[0] > edge
    assert w.body(0)[2] <= -2.0


@pytest.mark.parametrize("kind", PREC)
def test_stand(kind):
    """test_robot.c:70-99: the quadruped settles on flat ground, all feet down."""
    L = lib(kind)
    s = L.sim(1, grid=32)
    d, lo, hi, misc = s.robot_info()
    s.reset([0], [1], [0.0])
    maxv = 0.0
    for i in range(240):
        s.step(d[None, :])
        root, rot, joints, contacts, flags = s.observe()
        if i >= 180:
            maxv = max(maxv, float(np.linalg.norm(root[0, 7:10])))
    assert not flags[0]
    assert root[0, 2] > 0.8 * misc[0]
    assert maxv <= 2e-3
    assert (contacts[0, :4] > 0.99).all()
    assert contacts[0, 4] <= 0.0
    assert 1.0 - 2.0 * (root[0, 4] ** 2 + root[0, 5] ** 2) > 0.99


@pytest.mark.parametrize("kind", PREC)
def test_random_actions_stable(kind):
    """test_robot.c:101-146: random full-range targets every 3 steps on
    difficulty-1 terrain: nothing blows up, torso speed < 6 m/s.  The
    reference's 16 worlds x 600 steps on the GPU; 8 x 300 on the oracles."""
    L = lib(kind)
    n, steps = (16, 600) if kind == "gpu" else (8, 300)
    s = L.sim(n, grid=48)
    d, lo, hi, _ = s.robot_info()
    s.reset(np.arange(n), np.arange(n) + 100, np.ones(n))
    u = P.Lib("o64").rng_uniform(7, n * 12 * (steps // 3))
    maxspeed, unstable = 0.0, 0
    t = np.tile(d, (n, 1))
    for k in range(steps):
        if k % 3 == 0:
            uu = u[(k // 3) * n * 12:(k // 3 + 1) * n * 12].reshape(n, 12)
            t = lo + (hi - lo) * uu
        s.step(t)
        root, _, _, _, flags = s.observe()
        maxspeed = max(maxspeed, float(np.linalg.norm(root[:, 7:10], axis=1).max()))
        unstable += int(flags.sum())
    assert unstable == 0
    assert maxspeed <= 6.0


@pytest.mark.parametrize("kind", PREC)
def test_thread_determinism(kind):
    """test_robot.c:148-166: 1 vs 8 threads give identical states.  On the
    GPU the counterpart of the thread count is the form of the step kernel:
    one world per lane, per wavefront, four or two lanes per world, shared
    roots -- all give identical states, each the FP32 oracle's."""
    L = lib(kind)
    if kind == "gpu":
        sims = [TwinSim(8, grid=48, form=f) for f in ("lane", "wave", "leg", "pair", "shared")]
    else:
        sims = [L.sim(8, grid=48), L.sim(8, grid=48)]
    ids = np.arange(8)
    for s in sims:
        s.reset(ids, ids + 5, np.full(8, 0.7))
    u = P.Lib("o64").rng_uniform(3, 8 * 12 * 40)
    d, lo, hi, _ = sims[0].robot_info()
    for k in range(120):
        t = lo + (hi - lo) * u[(k // 3) * 96:(k // 3 + 1) * 96].reshape(8, 12)
        for i, s in enumerate(sims):
            s.step(t, 1 if i == 0 else 8)
    roots = [s.observe()[0] for s in sims]
    for r in roots[1:]:
        assert np.array_equal(roots[0], r)
