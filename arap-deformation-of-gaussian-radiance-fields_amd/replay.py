"""Recorded and scripted drags.

* DeformHistory: the edit history of a session and its text file, deform.txt
  (class DeformHistory, helper.hpp:172-197; RecordDeformation and
  LoadDeformation, GaussianView.cpp:4859-4935 and 4961-5100). The file is one
  line of space-separated tokens, each value followed by a space:
    <nodes_on_mesh> Nodes: n i... Total_Operations: t Move_Operations: m
    Operation_Types: n op... Block_Nodes: n (size id...)...
    Mouse_Movements: n (size (x y z)...)... Blocks_Types_Moves: n (size type...)...
    Energy_on_Center: n v... Twist_Axis: n (x y z w)...
  Floats are written like C++ `ostream << float` (printf %g, 6 significant
  digits), so a file written by the reference reads back and writes out byte
  for byte.
* HistoryPlayer: the playback state machine of GaussianView::onUpdate
  (GaussianView.cpp:1757-1915), one Frame per update call: block removals
  (op < 0), block additions (op 0: the first block is dragged, type 1, later
  ones pinned, type 0), and the steps of a move (op > 0, with the block types
  recorded for it): 1 translate with the constraints on the block centres, 2
  twist by 0.005 * int(x) radians about the recorded axis, 3 scale by
  1 + 0.002 * int(x), 4 translate.
* deform_script: the DeformScript trajectories 0-6 of the stripes scene
  (LoadDeformScript0-6, GaussianView.cpp:2512-2791) and ScriptPlayer, their
  playback (RunDeformScript and the run_script branch, GaussianView.cpp:
  1917-1990, 2793-2816).
* Session: runs Frames on the device state of a scene (the arapgs frame):
  rebuilds the solver and the static masks (UpdateIndicies /
  CheckStaticSamples, GaussianView.cpp:1996-2090) when the blocks change, sets
  the aims with the aims kernels and runs arapgs_deform_frame.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import arapgs

PI = 3.1415926535  # helper.hpp:47


def _fmt(x: float) -> str:
    return "%g" % float(np.float32(x))


@dataclass
class DeformHistory:
    nodes_on_mesh: bool = False
    node_indices: list = field(default_factory=list)      # Vertex_index of every graph node
    total_operations: int = 0
    move_operations: int = 0
    operation_types: list = field(default_factory=list)   # 0 add, -(i+1) remove i, 1..4 moves
    block_nodes: list = field(default_factory=list)       # node ids of every added block
    mouse_movements: list = field(default_factory=list)   # per move: float32 (steps, 3)
    blocks_types_moves: list = field(default_factory=list)
    energy_on_centers: list = field(default_factory=list)
    twist_axis: list = field(default_factory=list)        # per move: float32 (4,)

    # ------------------------------------------------------------ text format
    @classmethod
    def parse(cls, text: str) -> "DeformHistory":
        toks = text.split()
        pos = 0

        def take():
            nonlocal pos
            if pos >= len(toks):
                raise ValueError("deform.txt: unexpected end of file")
            pos += 1
            return toks[pos - 1]

        def label(name):
            t = take()
            if t != name:
                raise ValueError(f"deform.txt: expected {name!r}, found {t!r}")

        h = cls()
        h.nodes_on_mesh = int(take()) == 1
        label("Nodes:")
        h.node_indices = [int(take()) for _ in range(int(take()))]
        label("Total_Operations:")
        h.total_operations = int(take())
        label("Move_Operations:")
        h.move_operations = int(take())
        label("Operation_Types:")
        h.operation_types = [int(take()) for _ in range(int(take()))]
        label("Block_Nodes:")
        h.block_nodes = [[int(take()) for _ in range(int(take()))] for _ in range(int(take()))]
        label("Mouse_Movements:")
        h.mouse_movements = [np.array([[float(take()) for _ in range(3)] for _ in range(int(take()))],
                                      np.float32).reshape(-1, 3) for _ in range(int(take()))]
        label("Blocks_Types_Moves:")
        h.blocks_types_moves = [[int(take()) for _ in range(int(take()))] for _ in range(int(take()))]
        label("Energy_on_Center:")
        h.energy_on_centers = [int(take()) for _ in range(int(take()))]
        label("Twist_Axis:")
        h.twist_axis = [np.array([float(take()) for _ in range(4)], np.float32) for _ in range(int(take()))]
        return h

    @classmethod
    def load(cls, path) -> "DeformHistory":
        with open(path) as f:
            return cls.parse(f.read())

    def dumps(self) -> str:
        out = [str(1 if self.nodes_on_mesh else 0), "Nodes:", str(len(self.node_indices))]
        out += [str(int(i)) for i in self.node_indices]
        out += ["Total_Operations:", str(self.total_operations), "Move_Operations:", str(self.move_operations)]
        out += ["Operation_Types:", str(len(self.operation_types))] + [str(int(v)) for v in self.operation_types]
        out += ["Block_Nodes:", str(len(self.block_nodes))]
        for b in self.block_nodes:
            out += [str(len(b))] + [str(int(v)) for v in b]
        out += ["Mouse_Movements:", str(len(self.mouse_movements))]
        for m in self.mouse_movements:
            m = np.asarray(m, np.float32).reshape(-1, 3)
            out += [str(m.shape[0])] + [_fmt(v) for v in m.reshape(-1)]
        out += ["Blocks_Types_Moves:", str(len(self.blocks_types_moves))]
        for b in self.blocks_types_moves:
            out += [str(len(b))] + [str(int(v)) for v in b]
        out += ["Energy_on_Center:", str(len(self.energy_on_centers))] + [str(int(v)) for v in self.energy_on_centers]
        out += ["Twist_Axis:", str(len(self.twist_axis))]
        for a in self.twist_axis:
            out += [_fmt(v) for v in np.asarray(a, np.float32).reshape(4)]
        return "".join(t + " " for t in out)

    def save(self, path) -> None:
        with open(path, "w") as f:
            f.write(self.dumps())

    # ------------------------------------------------------------ recording
    # The bookkeeping of the editing UI (GaussianView.cpp:1246-1248, 1328-1329,
    # 1440-1476, 1584-1610).
    def record_add_block(self, nodes) -> None:
        self.total_operations += 1
        self.operation_types.append(0)
        self.block_nodes.append([int(v) for v in nodes])

    def record_remove_block(self, index: int) -> None:
        self.total_operations += 1
        self.operation_types.append(-(int(index) + 1))

    def record_move_step(self, value, axis=(0.0, 0.0, 0.0, 0.0)) -> None:
        """One drag step: the translation (ops 1, 4) or (int offset, 0, 0)
        (twist, scale); axis is stored with the first step of a move."""
        if len(self.mouse_movements) == self.move_operations:
            self.mouse_movements.append(np.zeros((0, 3), np.float32))
            self.twist_axis.append(np.asarray(axis, np.float32).reshape(4))
        m = self.mouse_movements[self.move_operations]
        self.mouse_movements[self.move_operations] = np.concatenate(
            [m, np.asarray(value, np.float32).reshape(1, 3)])

    def record_move_end(self, op: int, blocks_type, on_center: bool) -> None:
        if op not in (1, 2, 3, 4):
            raise ValueError("move operations are 1 (centre), 2 (twist), 3 (scale), 4 (translate)")
        self.operation_types.append(int(op))
        self.blocks_types_moves.append([int(t) for t in blocks_type])
        self.energy_on_centers.append(1 if on_center else 0)
        self.total_operations += 1
        self.move_operations += 1


# ---------------------------------------------------------------- playback
@dataclass
class Frame:
    """One update call of a playback."""
    kind: str                   # "add", "remove", "move"
    op: int
    blocks: list                # block node lists after this call's edit
    types: list
    changed: bool               # blocks or types changed in this call
    center: bool = False        # constraints on the block centres (op 1)
    delta: np.ndarray | None = None    # ops 1, 4
    axis: np.ndarray | None = None     # op 2
    radian: float = 0.0                # op 2
    scale: float = 1.0                 # op 3
    aims_idx: np.ndarray | None = None   # scripts: nodes with explicit aims
    aims_val: np.ndarray | None = None
    step: int = 0


class HistoryPlayer:
    """Iterates the Frames of a DeformHistory (GaussianView.cpp:1757-1915)."""

    def __init__(self, history: DeformHistory):
        self.h = history

    def __iter__(self):
        h = self.h
        blocks, types = [], []
        step = add_idx = move_idx = cur = 0
        run_move = False
        current_control = 0
        while step < h.total_operations:
            if step >= len(h.operation_types):
                raise ValueError("deform.txt: fewer operation types than Total_Operations")
            op = h.operation_types[step]
            changed = False
            kind = None
            if op < 0:
                if not blocks:
                    raise ValueError(f"operation {step}: removes a block but none is left (the viewer would stall)")
                i = -(op + 1)
                current_control = i
                del blocks[i]
                del types[i]
                current_control = current_control % len(blocks) if blocks else 0
                changed, kind = True, "remove"
                step += 1
            elif op == 0:
                blocks.append(list(h.block_nodes[add_idx]))
                types.append(1 if not types else 0)
                if len(types) == 1:
                    current_control = 0
                add_idx += 1
                changed, kind = True, "add"
                step += 1
            elif not run_move:
                new_types = list(h.blocks_types_moves[move_idx])
                if len(new_types) != len(blocks):
                    raise ValueError(f"move {move_idx}: {len(new_types)} block types for {len(blocks)} blocks")
                changed = new_types != types
                types = new_types
                cur = 0
                run_move = True
            frame = None
            if kind is not None:
                frame = Frame(kind, op, [list(b) for b in blocks], list(types), changed)
            if run_move and cur < len(h.mouse_movements[move_idx]):
                mv = np.asarray(h.mouse_movements[move_idx][cur], np.float32)
                frame = Frame("move", op, [list(b) for b in blocks], list(types), changed, step=cur)
                if op in (1, 4):
                    frame.delta = mv.copy()
                    frame.center = op == 1
                elif op == 2:
                    frame.axis = np.asarray(h.twist_axis[move_idx], np.float32)[:3].copy()
                    frame.radian = float(np.float32(0.005) * np.float32(int(mv[0])))
                elif op == 3:
                    frame.scale = float(np.float32(0.002) * np.float32(int(mv[0])) + np.float32(1.0))
                cur += 1
            if run_move and cur == len(h.mouse_movements[move_idx]):
                run_move = False
                move_idx += 1
                step += 1
            if frame is not None:
                yield frame


# ---------------------------------------------------------------- scripts
STRIPES_BLOCK1 = [0, 20, 53, 59, 63, 64, 67, 68, 145, 167, 189, 190, 192, 196, 197, 199]
STRIPES_BLOCK2 = [1, 7, 21, 26, 54, 70, 80, 127, 176, 178, 179, 181, 183, 184, 185, 186]
STRIPES6_BLOCK1 = [0, 3, 4, 7, 9, 11, 13, 14, 16, 25, 30, 31, 35, 36, 43, 44, 45, 47, 56, 57, 58, 64, 79, 81, 84, 85,
                   86, 95, 98, 101, 105, 128, 129, 130, 131, 132, 147, 149, 151, 152, 153, 156, 157, 159, 160, 161,
                   162, 164, 165, 166, 169, 170, 172, 173, 174, 175, 183, 184, 186, 187, 195, 196, 198, 199]
STRIPES6_BLOCK2 = [1, 2, 5, 6, 8, 10, 12, 15, 17, 22, 23, 24, 28, 32, 38, 39, 48, 53, 54, 55, 61, 65, 67, 72, 73, 75,
                   82, 87, 90, 91, 92, 102, 106, 107, 108, 109, 111, 112, 117, 118, 120, 121, 123, 124, 133, 134, 135,
                   136, 137, 138, 139, 140, 141, 142, 143, 144, 176, 177, 179, 180, 188, 189, 191, 192]


def rotate_about_axis(p, center, axis, radian):
    """PointRotateByAxis (helper.cpp:1077-1100) in float32, term for term."""
    f = np.float32
    p = np.asarray(p, f).reshape(-1, 3)
    radian = f(radian)
    cost, sint = f(np.cos(radian)), f(np.sin(radian))
    ax = np.asarray(axis, f)
    norm = f(np.sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]))
    x, y, z = ax[0] / norm, ax[1] / norm, ax[2] / norm
    one = f(1.0)
    a, b, c = (f(v) for v in np.asarray(center, f))
    out = np.empty_like(p)
    out[:, 0] = (x * x * (one - cost) + cost) * p[:, 0] + (x * y * (one - cost) - z * sint) * p[:, 1] + \
        (x * z * (one - cost) + y * sint) * p[:, 2]
    out[:, 1] = (y * x * (one - cost) + z * sint) * p[:, 0] + (y * y * (one - cost) + cost) * p[:, 1] + \
        (y * z * (one - cost) - x * sint) * p[:, 2]
    out[:, 2] = (z * x * (one - cost) - y * sint) * p[:, 0] + (z * y * (one - cost) + x * sint) * p[:, 1] + \
        (z * z * (one - cost) + cost) * p[:, 2]
    out[:, 0] += (a * (y * y + z * z) - x * (b * y + c * z)) * (one - cost) + (b * z - c * y) * sint
    out[:, 1] += (b * (x * x + z * z) - y * (a * x + c * z)) * (one - cost) + (c * x - a * z) * sint
    out[:, 2] += (c * (x * x + y * y) - z * (a * x + b * y)) * (one - cost) + (a * y - b * x) * sint
    return out


@dataclass
class DeformScript:
    """class DeformScript (helper.hpp:200-208)."""
    block_nodes: list
    blocks_types: list
    node_aims: np.ndarray  # (steps, dragged nodes, 3) float32


def deform_script(index: int, node_pos) -> DeformScript:
    """LoadDeformScript0-6 (GaussianView.cpp:2512-2791) for a graph whose node
    positions are node_pos (the stripes graph: 200 nodes). The arithmetic keeps
    the reference's types: Pi is a double, k and the outputs are floats."""
    f = np.float32
    node_pos = np.asarray(node_pos, f)
    steps = 50
    if index == 6:
        b1, b2, types = STRIPES6_BLOCK1, STRIPES6_BLOCK2, [0, 1]
        start = node_pos[b2]
        aims = [start - np.array([0.01, 0.0, 0.0], f) * f(s) for s in range(1, steps + 1)]
        return DeformScript([b1, b2], types, np.stack(aims).astype(f))
    b1, b2 = STRIPES_BLOCK1, STRIPES_BLOCK2
    drag_first = index == 3
    types = [1, 0] if drag_first else [0, 1]
    start = node_pos[b1 if drag_first else b2]
    center = np.array([0.0, 1.5 if drag_first else -1.5, 0.0], f)
    n_steps = steps - 11 if index in (4, 5) else steps   # df_step < inter_steps - 10, else <= inter_steps
    aims = []
    for s in range(1, n_steps + 1):
        if index in (1, 5):
            radian = f(-float(f(s)) * 1.5 * PI / float(f(steps)))
            aims.append(rotate_about_axis(start, center, (0, 1, 0, 0), radian))
            continue
        if index in (0, 4):
            k = f(5.4) if index == 0 else f(8.4)
            a = f(3.0 * PI * PI / float(k * k))
            x = f(float(f(s)) * (-float(k) / PI) / float(f(steps)))
            y = a * x * x
            radian = f(-float(f(s)) * PI / float(f(steps)))
        elif index == 2:
            k = f(3.0)
            a = f(PI / float(k) - PI * PI / (2 * float(k)))
            x = f(float(f(s)) * (-float(k) / PI) / float(f(steps)))
            y = -a * x * x
            radian = f(-float(f(s)) * (PI / 2.0) / float(f(steps)))
        elif index == 3:
            k = f(3.0)
            a = f(PI / float(k) - PI * PI / (2 * float(k)))
            x = f(float(f(s)) * (-float(k) / PI) / float(f(steps)))
            y = a * x * x
            radian = f(float(f(s)) * (PI / 2.0) / float(f(steps)))
        else:
            raise ValueError("deform scripts are 0..6")
        rot = rotate_about_axis(start, center, (0, 0, 1, 0), radian)
        aims.append(rot + np.array([x, y, 0.0], f))
    return DeformScript([b1, b2], types, np.stack(aims).astype(f))


class ScriptPlayer:
    """RunDeformScript + the run_script branch (GaussianView.cpp:1917-1990,
    2793-2816): the script's blocks and types are set once; every step aims
    the pinned (type 0) nodes at their positions when the script started and
    the dragged (type 1) nodes, in block order, at the script's aims."""

    def __init__(self, script: DeformScript, start_node_pos):
        self.s = script
        self.start = np.asarray(start_node_pos, np.float32)

    def __iter__(self):
        s = self.s
        blocks = [list(b) for b in s.block_nodes]
        types = list(s.blocks_types)
        # the writes of one step in the reference's order: block by block,
        # pinned nodes from the start positions, dragged ones from the script
        idx, src = [], []
        k = 0
        for b, t in zip(blocks, types):
            if t == 0:
                idx += b
                src += [("start", n) for n in b]
            elif t == 1:
                idx += b
                src += [("script", k + j) for j in range(len(b))]
                k += len(b)
        if k != s.node_aims.shape[1]:
            raise ValueError("script: the dragged blocks and the aims differ in size")
        for step in range(s.node_aims.shape[0]):
            val = np.array([self.start[i] if w == "start" else s.node_aims[step, i] for w, i in src],
                           np.float32).reshape(-1, 3)
            yield Frame("move", 0, blocks, types, step == 0, aims_idx=np.array(idx, np.int32), aims_val=val,
                        step=step)


# ---------------------------------------------------------------- device session
class Session:
    """Runs playback Frames on the device state of a scene (the dict of
    workload.prepare: pos, rot, scale, sh, scale0, ends, ends_idx, ends_w,
    node_pos, node_src_idx, node_src_w, samples, sample_idx, sample_w,
    sample_feat; gs_static and sample_static are set here)."""

    def __init__(self, ctx, state: dict, nbr, gn_iters=30, weighting_factor=1.0, stop_rule=True, **solver_kw):
        import torch
        self.torch = torch
        self.ctx = ctx
        self.st = state
        self.nbr = np.ascontiguousarray(nbr, np.uint32)
        self.src_idx = state["node_src_idx"].cpu().numpy()
        self.src_w = state["node_src_w"].cpu().numpy()
        self.iters, self.stop_rule = gn_iters, stop_rule
        self.wf = weighting_factor
        self.solver_kw = solver_kw
        self.solver = None
        self._layout = None
        self.M, self.K = self.nbr.shape
        dev = state["node_pos"].device
        self.aims = torch.empty_like(state["node_pos"])
        self.energies = torch.zeros(gn_iters + 1, dtype=torch.float64, device=dev)
        self.iters_done = torch.zeros(1, dtype=torch.int32, device=dev)
        self.active = torch.zeros(0, dtype=torch.int32, device=dev)

    def _relayout(self, frame: Frame):
        torch = self.torch
        key = (tuple(tuple(b) for b in frame.blocks), tuple(frame.types), frame.center)
        if key == self._layout:
            return
        self._layout = key
        if self.solver is not None:
            self.solver.close()
            self.solver = None
        dev = self.st["node_pos"].device
        node_pos = self.st["node_pos"].cpu().numpy()
        blocks = [np.asarray(b, np.int32) for b in frame.blocks]
        self.solver = arapgs.Solver(self.ctx, self.nbr, blocks, np.asarray(frame.types, np.int32), self.src_idx,
                                    self.src_w, node_pos, weighting_factor=self.wf,
                                    constraints_on_center=frame.center, **self.solver_kw)
        # UpdateIndicies + CheckStaticSamples: static nodes, samples and Gaussians
        node_static = np.zeros(self.M, np.uint8)
        for b, t in zip(frame.blocks, frame.types):
            if t < 0:
                node_static[list(b)] = 1
        ns = torch.from_numpy(node_static).to(dev)
        st = self.st
        if st.get("samples") is not None and st["samples"].shape[0] > 0:
            st["sample_static"] = self.ctx.static_rows(ns, st["sample_idx"])
        n = st["pos"].shape[0]
        st["gs_static"] = self.ctx.static_rows(ns, st["ends_idx"].view(n, 6 * self.K)) if frame.blocks else None
        act = [n for b, t in zip(frame.blocks, frame.types) if t == 1 for n in b]
        self.active = torch.tensor(act, dtype=torch.int32, device=dev)

    def run(self, frame: Frame):
        """Applies one Frame; returns (energies, iterations) of its solve, or
        None for an edit-only frame."""
        torch = self.torch
        if frame.kind != "move":
            return None
        self._relayout(frame)
        st, ctx = self.st, self.ctx
        npos = st["node_pos"]
        if frame.aims_idx is not None:
            idx = torch.from_numpy(frame.aims_idx).to(npos.device)
            val = torch.from_numpy(np.ascontiguousarray(frame.aims_val, np.float32)).to(npos.device)
            ctx.aims_set(npos, idx, val, self.aims)
        elif self.active.shape[0] == 0:
            ctx.aims_translate(npos, self.active, np.zeros(3, np.float32), self.aims)
        elif frame.op in (1, 4):
            ctx.aims_translate(npos, self.active, frame.delta, self.aims)
        elif frame.op == 2:
            ctx.aims_twist(npos, self.active, frame.axis, frame.radian, self.aims)
        elif frame.op == 3:
            ctx.aims_scale(npos, self.active, frame.scale, self.aims)
        else:
            ctx.aims_translate(npos, self.active, np.zeros(3, np.float32), self.aims)
        fb = arapgs.frame_buffers(st, self.K)
        self.solver.deform_frame(fb, self.aims, self.iters, self.stop_rule, energies=self.energies,
                                 iters_done=self.iters_done)
        return self.energies, self.iters_done

    def close(self):
        if self.solver is not None:
            self.solver.close()
            self.solver = None
