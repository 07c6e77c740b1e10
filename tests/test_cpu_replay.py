"""deform.txt record / replay and the DeformScript trajectories (replay.py,
include/arapgs_history.hpp), on the CPU.

* format: the reference's recorded session (datasets/pinocchio deform.txt,
  committed as tests/golden/pinocchio_deform.txt) parses and writes back byte
  for byte, from Python and from the C++ header (compiled here with g++); a
  history recorded through the recording calls survives both;
* playback: the frames of the pinocchio session follow the reference's rules
  (four block additions, the first dragged and the others pinned, then 244 +
  51 + 45 translation steps with the recorded block types and movements); a
  synthetic history exercises removal, twist (0.005 int(x) rad) and scale
  (1 + 0.002 int(x)); a removal with no block left is an error;
* scripts: PointRotateByAxis agrees with Rodrigues' rotation in FP64, and the
  trajectories of LoadDeformScript0-6 agree with their closed forms (circle
  arcs about the z or y axis through (0, -+1.5, 0) plus the parabolic offset,
  a straight line for script 6), with 50 steps (39 for scripts 4 and 5).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from arapgs_amd import replay as R

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "pinocchio_deform.txt")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def fixture_text():
    with open(FIXTURE) as f:
        return f.read()


def test_pinocchio_history_round_trips_byte_for_byte():
    text = fixture_text()
    h = R.DeformHistory.parse(text)
    assert not h.nodes_on_mesh and len(h.node_indices) == 501
    assert h.total_operations == 7 and h.move_operations == 3
    assert h.operation_types == [0, 0, 0, 0, 4, 4, 4]
    assert [len(b) for b in h.block_nodes] == [211, 281, 7, 2]
    assert [m.shape[0] for m in h.mouse_movements] == [244, 51, 45]
    assert h.blocks_types_moves == [[-1, -1, -1, 1]] * 3 and h.energy_on_centers == [0, 0, 0]
    assert h.dumps() == text


@pytest.fixture(scope="module")
def history_check(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("hist") / "history_check")
    subprocess.check_call([gxx, "-O1", "-std=c++17", "-I", INCLUDE, os.path.join(HERE, "dropin", "history_check.cpp"),
                           "-o", exe])
    return exe


def test_cpp_header_round_trips_the_fixture(history_check, tmp_path):
    out = str(tmp_path / "out.txt")
    res = subprocess.run([history_check, FIXTURE, out], capture_output=True, text=True, check=True)
    assert res.stdout.strip() == "nodes 501 operations 7 moves 3 blocks 4 steps 340"
    with open(out) as f:
        assert f.read() == fixture_text()


def synthetic_history():
    h = R.DeformHistory(nodes_on_mesh=True, node_indices=list(range(0, 40, 2)))
    h.record_add_block([0, 1, 2])
    h.record_add_block([5, 6])
    h.record_add_block([9, 10, 11, 12])
    for v in (7.9, -2.2, 12.0):
        h.record_move_step((v, 0, 0), axis=(0.1, 0.9, -0.2, 0.0))
    h.record_move_end(2, [1, 0, -1], False)
    h.record_remove_block(1)
    for v in (-3.5, 4.0):
        h.record_move_step((v, 0, 0))
    h.record_move_end(3, [1, -1], False)
    h.record_move_step((0.0125, -0.25, 1e-7))
    h.record_move_end(1, [1, 0], True)
    return h


def test_recorded_history_round_trips_through_both_codecs(history_check, tmp_path):
    h = synthetic_history()
    p = str(tmp_path / "a.txt")
    h.save(p)
    back = R.DeformHistory.load(p)
    assert back.dumps() == h.dumps()
    assert back.operation_types == [0, 0, 0, 2, -2, 3, 1] and back.total_operations == 7
    assert back.move_operations == 3 and back.energy_on_centers == [0, 0, 1]
    out = str(tmp_path / "b.txt")
    subprocess.run([history_check, p, out], check=True, capture_output=True)
    with open(out) as f:
        assert f.read() == h.dumps()


def test_malformed_history_is_rejected(history_check, tmp_path):
    text = fixture_text().replace("Block_Nodes:", "Block_Nodez:")
    with pytest.raises(ValueError):
        R.DeformHistory.parse(text)
    p = str(tmp_path / "bad.txt")
    with open(p, "w") as f:
        f.write(text)
    res = subprocess.run([history_check, p, str(tmp_path / "o.txt")], capture_output=True, text=True)
    assert res.returncode == 1 and "Block_Nodes" in res.stderr


def test_pinocchio_playback_frames():
    h = R.DeformHistory.parse(fixture_text())
    frames = list(R.HistoryPlayer(h))
    assert len(frames) == 4 + 340
    adds = frames[:4]
    assert [f.kind for f in adds] == ["add"] * 4
    assert [f.types for f in adds] == [[1], [1, 0], [1, 0, 0], [1, 0, 0, 0]]
    assert adds[-1].blocks == h.block_nodes
    moves = frames[4:]
    assert all(f.kind == "move" and f.op == 4 and not f.center for f in moves)
    assert all(f.types == [-1, -1, -1, 1] for f in moves)
    assert [i for i, f in enumerate(moves) if f.changed] == [0]
    deltas = np.stack([f.delta for f in moves])
    assert np.array_equal(deltas, np.concatenate(h.mouse_movements))
    assert [f.step for f in moves[:3]] == [0, 1, 2] and moves[244].step == 0


def test_synthetic_playback_remove_twist_scale_centre():
    frames = list(R.HistoryPlayer(synthetic_history()))
    kinds = [(f.kind, f.op) for f in frames]
    assert kinds == [("add", 0)] * 3 + [("move", 2)] * 3 + [("remove", -2)] + [("move", 3)] * 2 + [("move", 1)]
    tw = frames[3:6]
    assert [f.radian for f in tw] == [float(np.float32(0.005) * np.float32(v)) for v in (7, -2, 12)]
    assert np.array_equal(tw[0].axis, np.float32([0.1, 0.9, -0.2]))
    assert tw[0].types == [1, 0, -1]
    rm = frames[6]
    assert rm.blocks == [[0, 1, 2], [9, 10, 11, 12]] and rm.types == [1, -1]
    sc = frames[7:9]
    assert [f.scale for f in sc] == [float(np.float32(0.002) * np.float32(v) + np.float32(1)) for v in (-3, 4)]
    c = frames[9]
    assert c.center and c.types == [1, 0] and np.array_equal(c.delta, np.float32([0.0125, -0.25, 1e-7]))


def test_removal_without_blocks_is_an_error():
    h = R.DeformHistory()
    h.record_remove_block(0)
    with pytest.raises(ValueError):
        list(R.HistoryPlayer(h))


def rodrigues(p, center, axis, theta):
    a = np.asarray(axis[:3], np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    Rm = np.eye(3) + np.sin(theta) * K + (1 - np.cos(theta)) * K @ K
    return (np.asarray(p, np.float64) - center) @ Rm.T + center


def test_point_rotate_by_axis_is_a_rotation():
    rng = np.random.default_rng(0)
    p = rng.normal(size=(50, 3)).astype(np.float32)
    for axis, theta in (((0, 0, 1, 0), 0.7), ((0.3, -1.2, 0.5, 0), -2.1), ((0, 1, 0, 0), 3.0)):
        c = np.float32([0.2, -1.5, 0.1])
        out = R.rotate_about_axis(p, c, axis, theta)
        np.testing.assert_allclose(out, rodrigues(p, c.astype(np.float64), axis, float(np.float32(theta))),
                                   atol=2e-6)


@pytest.fixture(scope="module")
def stripes_nodes():
    import json
    with open(os.path.join(HERE, "golden", "stripes_graph.json")) as f:
        d = json.load(f)
    return np.asarray(d["nodes"], np.float32)


@pytest.mark.parametrize("index", range(7))
def test_deform_scripts_closed_forms(stripes_nodes, index):
    s = R.deform_script(index, stripes_nodes)
    n_steps = 39 if index in (4, 5) else 50
    assert s.node_aims.shape[0] == n_steps
    if index == 6:
        assert s.blocks_types == [0, 1] and len(s.block_nodes[1]) == 64
        start = stripes_nodes[s.block_nodes[1]]
        for k in (1, 17, 50):
            exp = start - np.float32([0.01, 0, 0]) * np.float32(k)
            assert np.array_equal(s.node_aims[k - 1], exp)
        return
    drag = 0 if index == 3 else 1
    assert s.blocks_types == ([1, 0] if index == 3 else [0, 1])
    start = stripes_nodes[s.block_nodes[drag]].astype(np.float64)
    c = np.array([0, 1.5 if index == 3 else -1.5, 0])
    pi = np.pi
    for k in (1, 13, n_steps):
        if index in (1, 5):
            exp = rodrigues(start, c, (0, 1, 0), -k * 1.5 * pi / 50)
        else:
            kk = {0: 5.4, 2: 3.0, 3: 3.0, 4: 8.4}[index]
            x = -k * kk / (pi * 50)
            if index in (0, 4):
                y, theta = 3 * pi * pi / kk ** 2 * x * x, -k * pi / 50
            elif index == 2:
                y, theta = -(pi / kk - pi * pi / (2 * kk)) * x * x, -k * pi / 100
            else:
                y, theta = (pi / kk - pi * pi / (2 * kk)) * x * x, k * pi / 100
            exp = rodrigues(start, c, (0, 0, 1), theta) + np.array([x, y, 0])
        np.testing.assert_allclose(s.node_aims[k - 1], exp, atol=2e-5)


def test_script_player_aims(stripes_nodes):
    s = R.deform_script(3, stripes_nodes)
    start = stripes_nodes.copy()
    frames = list(R.ScriptPlayer(s, start))
    assert len(frames) == 50 and frames[0].changed and not frames[1].changed
    f = frames[7]
    b1, b2 = s.block_nodes
    assert list(f.aims_idx) == b1 + b2
    assert np.array_equal(f.aims_val[: len(b1)], s.node_aims[7])
    assert np.array_equal(f.aims_val[len(b1):], start[b2])
