"""The structure stage on the GPU (structureReady, nelson/DoubleSection.hpp:83-246;
the (W, W) pairs of S, nelson/MatrixWWtMultiplier.hpp:18-120): every field
of the device build (edge orders, W blocks and their repeated observations,
camera-major lists, the reference's W order, S blocks with their pair
lists, the camera order) equals the host build's bit for bit
(nsn_structure_check), on the 10-camera problem with fixed cameras and
points, per-edge weights, repeated observations, unobserved cameras and
points and every camera order, on the ladybug graph, and at the
1000-camera size."""
import numpy as np
import pytest

from helpers import small_cfg
from nelson_amd import CONFIGS, BAProblem, generate_ba, generate_on_graph, load_visibility

pytestmark = pytest.mark.gpu


def _check(g, **kw):
    n, th, td = BAProblem.structure_check(g, **kw)
    print(f"structure: E={len(g['cam_idx'])} {kw.get('camera_order', 'given')}: host {th:.1f} ms, "
          f"device {td:.1f} ms", flush=True)
    assert n == 0
    return th, td


@pytest.mark.parametrize("order", ["given", "rcm", "nd"])
def test_structure_small_problem(order):
    g = generate_ba(dict(CONFIGS["cfg1"]))
    _check(g, camera_order=order)


@pytest.mark.parametrize("order", ["given", "nd"])
def test_structure_fixed_weights_repeats_unobserved(order):
    rng = np.random.default_rng(4)
    g = generate_ba(small_cfg(n_cameras=24, n_points=900, k_min=3, k_max=6, layout=1, arc=8, seed=11))
    E = len(g["cam_idx"])
    # repeated observations of some (camera, point) pairs, out of order
    rep = rng.choice(E, 60, replace=False)
    g = dict(g)
    g["cam_idx"] = np.concatenate([g["cam_idx"], g["cam_idx"][rep]])
    g["pt_idx"] = np.concatenate([g["pt_idx"], g["pt_idx"][rep]])
    g["obs"] = np.concatenate([np.asarray(g["obs"]).reshape(-1, 2), np.asarray(g["obs"]).reshape(-1, 2)[rep] + 0.5])
    perm = rng.permutation(len(g["cam_idx"]))
    for k in ("cam_idx", "pt_idx", "obs"):
        g[k] = np.asarray(g[k])[perm]
    # one more camera and point that nothing observes
    g["n_cameras"] += 1
    g["n_points"] += 1
    cam_fixed = np.zeros(g["n_cameras"], np.uint8)
    cam_fixed[[0, 5, 17]] = 1
    pt_fixed = np.zeros(g["n_points"], np.uint8)
    pt_fixed[rng.choice(g["n_points"], 40, replace=False)] = 1
    weight = rng.uniform(0.5, 2.0, len(g["cam_idx"]))
    _check(g, camera_order=order, cam_fixed=cam_fixed, pt_fixed=pt_fixed, weight=weight)


def test_structure_every_point_fixed():
    g = generate_ba(dict(CONFIGS["cfg1"]))
    _check(g, pt_fixed=np.ones(g["n_points"], np.uint8))


def test_structure_ladybug():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ladybug_49_7776.txt")
    g = generate_on_graph(load_visibility(path))
    for order in ("given", "rcm", "nd"):
        _check(g, camera_order=order)


@pytest.mark.parametrize("order", ["given", "nd"])
def test_structure_1000_cameras(order):
    g = generate_ba(dict(CONFIGS["cfg2"]))
    th, td = _check(g, camera_order=order)
    assert td < th  # the point of building it on the device
