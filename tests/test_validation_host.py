"""The boundary refuses invalid world state before touching the device (no GPU
needed): the checks run on the host and return the status codes of include/exgame.h."""
import numpy as np
import pytest

from exgame_amd import _lib, scenes
from exgame_amd.world import DeviceWorld


def bad(scene, **fields):
    s = scenes.jointed_chains(n_chains=2, m=4, seed=1) if scene is None else scene
    for k, v in fields.items():
        arr = getattr(s, k).copy()
        idx, val = v
        arr.reshape(-1)[idx] = val
        object.__setattr__(s, k, arr) if hasattr(s, "__dataclass_fields__") else setattr(s, k, arr)
    return s


@pytest.mark.parametrize("field,idx,val,code", [
    ("mass", 3, -1.0, 6),          # negative mass: invalid
    ("mass", 2, np.nan, 6),
    ("mass", 1, 0.0, 7),           # zero mass = a kinematic body: unsupported
    ("inertia", 9 * 2 + 4, -0.5, 6),   # not positive definite
    ("inertia", 0, np.inf, 6),
    ("pose", 8 * 1 + 5, np.nan, 6),
    ("momentum", 6 * 3 + 2, np.inf, 6),
    ("size", 0, -1.0, 6),
    ("origin", 1, np.nan, 6),
    ("joint_params", 10 * 0 + 7, np.nan, 6),
])
def test_create_refuses_invalid_state(field, idx, val, code):
    s = bad(None, **{field: (idx, val)})
    with pytest.raises(_lib.ExgError) as e:
        DeviceWorld(s)
    assert e.value.code == code, str(e.value)


def test_create_refuses_a_rotorless_pose():
    s = scenes.jointed_chains(n_chains=1, m=2, seed=1)
    pose = s.pose.copy()
    pose[0, :4] = 0.0
    object.__setattr__(s, "pose", pose)
    with pytest.raises(_lib.ExgError) as e:
        DeviceWorld(s)
    assert e.value.code == 6 and "rotor" in str(e.value)
