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
    ("mass", 1, 0.0, 6),           # a kinematic body (mass 0) before dynamic ones in its island
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


def test_create_refuses_islands_beyond_the_large_record():
    # 49 bodies in one island (one more than the mechanism-sized kernel holds)
    s = scenes.cloth_grids(shapes=((7, 7),))
    with pytest.raises(_lib.ExgError) as e:
        DeviceWorld(s)
    assert e.value.code == 7 and "48" in str(e.value)
    # 129 joints between 10 bodies
    s = scenes.jointed_chains(n_chains=1, m=10, seed=1)
    extra = 129 - s.n_joints
    kind = np.concatenate([s.joint_kind, np.full(extra, scenes.TORSIONAL_DAMPER_WARPED, np.int32)])
    ja = np.concatenate([s.joint_a, np.zeros(extra, np.int32)])
    jb = np.concatenate([s.joint_b, np.ones(extra, np.int32)])
    prm = np.concatenate([s.joint_params, np.tile([[5.0] + [0.0] * 9], (extra, 1))])
    for k, v in (("joint_kind", kind), ("joint_a", ja), ("joint_b", jb), ("joint_params", prm),
                 ("island_joint_start", np.array([0, 129], np.int32))):
        object.__setattr__(s, k, v)
    with pytest.raises(_lib.ExgError) as e:
        DeviceWorld(s)
    assert e.value.code == 7 and "128" in str(e.value)


def test_kinematic_bodies_last_in_their_islands_pass_validation():
    # anchored chains: each island's anchor (mass 0, zero tensor) is its last body; the
    # host checks pass and creation goes on to the device (absent here: a backend error)
    s = scenes.anchored_chains(n_chains=3, m=8, seed=1)
    assert (s.mass == 0.0).sum() == 3
    with pytest.raises(_lib.ExgError) as e:
        DeviceWorld(s)
    assert e.value.code == 5, str(e.value)


def test_kinematic_bodies_first_are_refused():
    s = scenes.anchored_chains(n_chains=2, m=4, seed=1, kinematic_first=True)
    with pytest.raises(_lib.ExgError) as e:
        DeviceWorld(s)
    assert e.value.code == 6 and "kinematic bodies go last" in str(e.value)


def test_kinematic_body_in_a_mechanism_sized_island_passes_validation():
    s = scenes.anchored_chains(n_chains=1, m=12, seed=1)
    with pytest.raises(_lib.ExgError) as e:
        DeviceWorld(s)
    assert e.value.code == 5, str(e.value)
