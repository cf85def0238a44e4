"""Host-side parts of the vectorised env (arc/vecenv.py) that need no GPU:
the Box-like spaces and the VecEnv bookkeeping used when stable_baselines3 is
absent (SB3's VecEnv.seed gives env i the seed seed + i; a reset consumes
the pending seeds)."""

import numpy as np

from arc import vecenv


def test_box_fields():
    b = vecenv.Box(-1.0, 1.0, (12,))
    assert b.shape == (12,) and b.dtype == np.float32
    assert b.contains(np.zeros(12, np.float32)) and not b.contains(np.full(12, 2.0))
    o = vecenv.Box(-np.inf, np.inf, (195,))
    assert np.all(np.isinf(o.low)) and o.contains(np.zeros(195))


def test_standalone_bookkeeping():
    e = vecenv._StandaloneVecEnv(4, vecenv.Box(-np.inf, np.inf, (3,)), vecenv.Box(-1.0, 1.0, (2,)))
    assert e._seeds == [None] * 4
    assert e.seed(10) == [10, 11, 12, 13]
    e._reset_seeds()
    assert e._seeds == [None] * 4
    assert list(e._get_indices(None)) == [0, 1, 2, 3] and e._get_indices(2) == [2]
    e.set_options({"a": 1})
    assert e._options == [{"a": 1}] * 4


def test_device_env_has_the_vecenv_methods():
    # the methods ArcVecEnv defines (arc/env.py:86-140) and the ones SB3 callers use
    for name in ("reset", "step_async", "step_wait", "step", "close", "render", "get_images", "get_attr",
                 "set_attr", "env_method", "env_is_wrapped", "seed"):
        assert callable(getattr(vecenv.DeviceVecEnv, name))
