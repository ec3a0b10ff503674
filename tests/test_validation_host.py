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
    ("mass", 1, np.inf, 6),
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


def test_kinematic_bodies_first_are_refused_by_the_c_abi():
    # the C ABI takes an island's kinematic bodies last (DeviceWorld reorders for it)
    s = scenes.anchored_chains(n_chains=2, m=4, seed=1, kinematic_first=True)
    L = _lib.lib()
    i32 = lambda x: np.ascontiguousarray(x, dtype=np.int32)
    f64 = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    arrs = [f64(s.pose), f64(s.momentum), f64(s.mass), f64(s.inertia), f64(s.size)]
    ist, org = i32(s.island_start), f64(s.origin)
    jk, ja, jb, jp = i32(s.joint_kind), i32(s.joint_a), i32(s.joint_b), f64(s.joint_params)
    import ctypes as C
    h = C.c_void_p()
    rc = L.exg_world_create(s.n_bodies, *[_lib.fptr(x) for x in arrs], s.n_islands, _lib.iptr(ist), _lib.fptr(org),
                            s.n_joints, _lib.iptr(jk), _lib.iptr(ja), _lib.iptr(jb), _lib.fptr(jp), 1.0, 1,
                            C.byref(h))
    assert rc == 6
    assert "kinematic bodies go last" in L.exg_last_error().decode()


def test_device_world_takes_kinematic_bodies_anywhere():
    # DeviceWorld uploads each island's dynamic bodies first: validation passes (and
    # creation goes on to the absent device)
    s = scenes.anchored_chains(n_chains=2, m=4, seed=1, kinematic_first=True)
    with pytest.raises(_lib.ExgError) as e:
        DeviceWorld(s)
    assert e.value.code == 5, str(e.value)


def test_kinematic_body_in_a_mechanism_sized_island_passes_validation():
    s = scenes.anchored_chains(n_chains=1, m=12, seed=1)
    with pytest.raises(_lib.ExgError) as e:
        DeviceWorld(s)
    assert e.value.code == 5, str(e.value)
