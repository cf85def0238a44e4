"""Pins the oracle's fall rules (orc_sim_falls, a restatement of the task's
rules in arc/task.py:180-203, which have no C counterpart in the reference).

Two kinds of check:
- the reference's own task (tests/golden/task_vectors.npz, its arc/task.py over
  the reference engine built from its sources): the oracle replays those 40
  steps and orc_sim_falls must give the recorded cause of every world at every
  step (the vectors hold flips and no-falls);
- states built so that each rule's answer is fixed by its definition, in both
  precisions: a torso past the terrain's half extent in x or y or below the
  void height (void; the bounds themselves are not void, the comparisons are
  strict), a torso tilted just past and just short of the flip threshold
  (R22 of a rotation about x by t is cos t), non-finite state and the
  engine's instability flag (bad), torso contact held for more than the
  allowed steps (beached; the counter counts consecutive steps with torso
  contact only, feet do not count, and restarts at zero without it), and the
  precedence bad > void > flip > beached on worlds where two rules hold.
A dropped rule, a swapped precedence, a non-strict bound, the wrong channel
or the wrong rotation entry fails one of them."""

from pathlib import Path

import numpy as np
import pytest

import pyoracle as P
from pytask import OracleTask
from test_task_golden import OFFSETS, Cfg

GOLDEN = Path(__file__).resolve().parent / "golden" / "task_vectors.npz"

NONE, BAD, VOID, FLIP, BEACHED = 0, 1, 2, 3, 4
pytestmark = pytest.mark.skipif(not P.available("o64"), reason="oracle not built (run build())")


def _qx(c: float) -> list:
    """unit quaternion (w, x, y, z) of a rotation about x whose cosine is c"""
    t = np.arccos(c)
    return [np.cos(t / 2), np.sin(t / 2), 0.0, 0.0]


class _Worlds:
    def __init__(self, kind, n):
        self.lib = P.Lib(kind)
        self.L = self.lib.lib
        self.sim = self.lib.sim(n)
        ids = np.arange(n, dtype=np.int32)
        self.sim.reset(ids, ids.astype(np.uint32), np.zeros(n))  # flat ground, standing robots
        self.default = self.sim.robot_info()[0]

    def body(self, i, b):
        out = np.zeros(13)
        self.L.orc_body_get(self.L.orc_sim_world(self.sim.p, i), b, out.ctypes.data)
        return out

    def set_body(self, i, b, s):
        s = np.ascontiguousarray(s, np.float64)
        self.L.orc_body_set(self.L.orc_sim_world(self.sim.p, i), b, s.ctypes.data)

    def torso(self, i, **kw):
        s = self.body(i, 0)
        for k, v in kw.items():
            if k == "q":
                s[3:7] = v
            else:
                s[{"x": 0, "y": 1, "z": 2, "vx": 7}[k]] = v
        self.set_body(i, 0, s)


def test_oracle_falls_reproduce_reference_task():
    """Replays the reference task's recorded episodes on the FP64 oracle and
    evaluates orc_sim_falls after every step, with its own torso-contact
    counter (cleared for the worlds the task resets)."""
    g = np.load(GOLDEN)
    cfg = type("C", (Cfg,), {})
    for k, v in zip(g["cfg_keys"], g["cfg_values"]):
        setattr(cfg, str(k), type(getattr(Cfg, str(k)))(v))
    n = int(g["n"])
    o = OracleTask(P.Lib("o64").sim(n, cfg.dt, cfg.substeps, cfg.grid, cfg.cell), cfg, OFFSETS, cfg.grid, cfg.cell,
                   rng=np.random.default_rng(int(g["seed"])))
    o.reset(np.arange(n))
    torso = np.zeros(n, np.int32)
    seen = set()
    for k in range(len(g["actions"])):
        o.step(g["actions"][k])
        cause = o.sim.falls(torso, int(cfg.beached_steps), float(cfg.flip_up_z), float(cfg.void_z))
        assert np.array_equal(cause, g["cause"][k]), f"step {k + 1}"
        seen.update(np.unique(cause).tolist())
        ended = np.flatnonzero(g["terminated"][k] | g["truncated"][k])
        torso[ended] = 0
        if len(ended):
            o.reset(ended)
    assert seen == {NONE, FLIP}


@pytest.mark.parametrize("kind", ["o64", "o32"])
def test_fall_rules_on_built_states(kind):
    half = 64 * 0.4 / 2  # grid x cell / 2: 12.8 m
    void_z, flip_z, beached = -2.0, 0.3, 30
    # the largest position of the oracle's precision at or below the bound,
    # and the next one up: |x| > half is strict
    dt = np.float64 if kind == "o64" else np.float32
    edge = dt(half) if float(dt(half)) <= half else np.nextafter(dt(half), dt(0))
    past = np.nextafter(edge, dt(np.inf))
    cases = [
        ({}, NONE),
        ({"x": half + 0.05}, VOID),
        ({"x": float(edge)}, NONE),
        ({"x": float(-past)}, VOID),
        ({"y": -(half + 0.05)}, VOID),
        ({"z": void_z - 0.01}, VOID),
        ({"z": void_z}, NONE),               # z < void_z is strict
        ({"q": _qx(flip_z - 1e-3)}, FLIP),
        ({"q": _qx(flip_z + 1e-3)}, NONE),
        ({"q": _qx(-1.0)}, FLIP),            # upside down
        ({"vx": np.nan}, BAD),
        ({"x": half + 1.0, "q": _qx(-1.0)}, VOID),  # void before flip
        ({"vx": np.inf, "x": half + 1.0}, BAD),     # bad before void
    ]
    w = _Worlds(kind, len(cases))
    for i, (kw, _) in enumerate(cases):
        w.torso(i, **kw)
    torso = np.full(len(cases), 100, np.int32)  # no torso contact: every counter restarts
    cause = w.sim.falls(torso, beached, flip_z, void_z)
    assert cause.tolist() == [c for _, c in cases]
    assert not torso.any()


@pytest.mark.parametrize("kind", ["o64", "o32"])
def test_beached_counts_torso_contact_only(kind):
    """Robots pushed 0.45 m into flat ground touch it with the torso during
    the next step; standing robots touch it with their feet only."""
    n, beached = 6, 2
    w = _Worlds(kind, n)
    for _ in range(3):  # the robots settle onto their feet
        w.sim.step(np.tile(w.default, (n, 1)))
    for i in range(1, n):
        for b in range(13):
            s = w.body(i, b)
            s[2] -= 0.45
            w.set_body(i, b, s)
    w.sim.step(np.tile(w.default, (n, 1)))
    root, rot, _, contacts, flags = w.sim.observe()
    # preconditions: torso contact only where lowered, feet touching on the
    # standing robot, and no other rule holds before world 5 is turned over
    assert contacts[0, 4] == 0 and contacts[0, :4].max() > 0
    assert (contacts[1:, 4] > 0).all()
    assert not flags.any() and np.isfinite(root).all()
    assert (rot[:, 2, 2] > 0.3).all() and (root[:, 2] > -2.0).all()
    w.torso(5, q=_qx(-1.0))  # turned over after its torso touched: flip before beached
    torso = np.array([7, beached - 1, beached, beached + 5, 0, beached], np.int32)
    cause = w.sim.falls(torso, beached, 0.3, -2.0)
    assert torso.tolist() == [0, beached, beached + 1, beached + 6, 1, beached + 1]
    assert cause.tolist() == [NONE, NONE, BEACHED, BEACHED, NONE, FLIP]


@pytest.mark.parametrize("kind", ["o64", "o32"])
def test_unstable_flag_is_bad(kind):
    """A NaN drive target passes the engine's clamp, makes the world
    non-finite and sets its instability flag: the bad rule."""
    n = 3
    w = _Worlds(kind, n)
    t = np.tile(w.default, (n, 1))
    t[1, 4] = np.nan
    w.sim.step(t)
    flags = w.sim.observe()[4]
    assert flags.tolist() == [0, 1, 0]
    cause = w.sim.falls(np.zeros(n, np.int32), 30, 0.3, -2.0)
    assert cause.tolist() == [NONE, BAD, NONE]
