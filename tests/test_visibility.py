"""The visibility-graph loader (nsn_visibility_load, the text format of
nelson/samples/Problems.cpp:8-30) and the generator on a given graph
(nsn_generate_on_graph, the geometry of the BA sample,
nelson/samples/BASchurPerfomances.cpp:40-92), host only.

The reference's one data fixture is the ladybug graph it embeds
(ProblemCollections::ladybug_49_7776, nelson/samples/Problems.cpp:38-...):
tests/golden/ladybug_49_7776.txt holds it in Problem::load's format
(written by tests/golden/make_ladybug.py).  Its known figures: 49 cameras,
7 776 points, 31 843 pairs, 3 449 points seen twice, 361-906 observations
per camera, no repeated pair."""
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from nelson_amd import LADYBUG_SAMPLE, NelsonError, generate_on_graph, load_visibility

LADYBUG = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ladybug_49_7776.txt")


@pytest.fixture(scope="module")
def ladybug():
    return load_visibility(LADYBUG)


def test_ladybug_graph(ladybug):
    g = ladybug
    assert (g["n_cameras"], g["n_points"], len(g["cam_idx"])) == (49, 7776, 31843)
    deg_p = np.bincount(g["pt_idx"], minlength=7776)
    deg_c = np.bincount(g["cam_idx"], minlength=49)
    assert deg_p.min() == 2 and (deg_p == 2).sum() == 3449
    assert deg_c.min() == 361 and deg_c.max() == 906
    assert len(np.unique(np.stack([g["cam_idx"], g["pt_idx"]], 1), axis=0)) == 31843
    # first and last pairs of the initializer
    assert list(zip(g["cam_idx"][:4], g["pt_idx"][:4])) == [(0, 0), (1, 0), (3, 0), (26, 0)]
    assert (g["cam_idx"][-1], g["pt_idx"][-1]) == (48, 7775)


def test_loader_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(0)
    c = rng.integers(0, 7, 40)
    q = rng.integers(0, 30, 40)
    f = tmp_path / "g.txt"
    # any whitespace separates the numbers, as with the reference's stream reads
    f.write_text("7 30\n40\n" + "\n".join(f"{a}\t{b}" for a, b in zip(c, q)) + "\n")
    g = load_visibility(f)
    assert g["n_cameras"] == 7 and g["n_points"] == 30
    assert np.array_equal(g["cam_idx"], c) and np.array_equal(g["pt_idx"], q)
    empty = tmp_path / "e.txt"
    empty.write_text("3 4 0\n")
    assert len(load_visibility(empty)["cam_idx"]) == 0
    bad = {
        "missing": None,
        "header": "3 4\n",
        "negative": "3 -4 1\n0 0\n",
        "short": "3 4 3\n0 0\n1 1\n",
        "malformed": "3 4 2\n0 0\n1 x\n",
        "camera range": "3 4 2\n0 0\n3 1\n",
        "point range": "3 4 2\n0 0\n1 4\n",
    }
    for name, text in bad.items():
        p = tmp_path / f"bad_{name.replace(' ', '_')}.txt"
        if text is not None:
            p.write_text(text)
        with pytest.raises(NelsonError):
            load_visibility(p)


def _project(cams, pts, ci, pi, f, cx, cy):
    R = Rotation.from_rotvec(cams.reshape(-1, 6)[ci, :3]).as_matrix()
    Xc = np.einsum("eij,ej->ei", R, pts.reshape(-1, 3)[pi]) + cams.reshape(-1, 6)[ci, 3:]
    return np.stack([f * Xc[:, 0] / Xc[:, 2] + cx, f * Xc[:, 1] / Xc[:, 2] + cy], 1), Xc


def test_generator_on_graph_geometry(ladybug):
    g = generate_on_graph(ladybug, obs_sigma=0.0)
    f, cx, cy = LADYBUG_SAMPLE["focal"], LADYBUG_SAMPLE["cx"], LADYBUG_SAMPLE["cy"]
    assert g["n_edges"] == 31843 and g["focal"] == f
    uv, Xc = _project(g["cams_gt"], g["pts_gt"], g["cam_idx"], g["pt_idx"], f, cx, cy)
    assert np.abs(g["obs"].reshape(-1, 2) - uv).max() <= 1e-9 * np.abs(uv).max()
    assert Xc[:, 2].min() > 2.0 - np.sqrt(3.0)  # every point in front of every camera
    assert np.abs(g["pts_gt"]).max() <= 1.0
    # cameras at radius 2 looking at the origin: centre = -R^T t, the origin on the optical axis
    cams = g["cams_gt"].reshape(-1, 6)
    R = Rotation.from_rotvec(cams[:, :3]).as_matrix()
    centre = -np.einsum("cji,cj->ci", R, cams[:, 3:])
    assert np.allclose(np.linalg.norm(centre, axis=1), 2.0, atol=1e-12)
    assert np.allclose(cams[:, 3:], [[0.0, 0.0, 2.0]] * 49, atol=1e-12)
    # camera 0 and point 0 unperturbed, the rest perturbed; seeded
    g2 = generate_on_graph(ladybug)
    assert np.array_equal(g2["cams0"][:6], g2["cams_gt"][:6]) and np.array_equal(g2["pts0"][:3], g2["pts_gt"][:3])
    d = (g2["cams0"] - g2["cams_gt"]).reshape(-1, 6)[1:]
    assert np.all(np.abs(d).max(axis=1) > 0) and abs(d.std() / LADYBUG_SAMPLE["rot_sigma"] - 1.0) < 0.2
    assert np.array_equal(generate_on_graph(ladybug)["obs"], g2["obs"])
    assert not np.array_equal(generate_on_graph(ladybug, seed=50)["obs"], g2["obs"])
    r = g2["obs"].reshape(-1, 2) - _project(g2["cams_gt"], g2["pts_gt"], g2["cam_idx"], g2["pt_idx"], f, cx, cy)[0]
    assert abs(r.std() / LADYBUG_SAMPLE["obs_sigma"] - 1.0) < 0.02


def test_generator_on_graph_rejects_bad_input(ladybug):
    with pytest.raises(NelsonError):
        generate_on_graph(ladybug, cam_radius=1.5)  # points could sit behind a camera
    bad = dict(ladybug, cam_idx=ladybug["cam_idx"].copy())
    bad["cam_idx"][7] = 49
    with pytest.raises(NelsonError):
        generate_on_graph(bad)
