"""Batches of independent boxes stepped through captured step graphs, and the
tile-shape autotuner (needs an MI355X).

Every box of a batch must end bit for bit where the same box ends when it is
stepped alone with misa_step: the graph replays the same launches with the
same arguments.  Each test replays the graphs several times with eager steps,
downloads and host checks in between.
"""
import numpy as np
import pytest

import misa_eam as M
import oracle_py as O

pytestmark = pytest.mark.gpu

A = 2.855
F = 5.6 / 2.855
FIELDS = ("id", "type", "x", "v", "f", "rho", "df")


def _box(setfl, cells, seed, tset, ratio, precision, pka=None):
    sim = O.Sim(cells, A, F, setfl)
    sim.build_world(seed, tset, ratio, seed)
    atoms = sim.atoms().copy()
    out = []
    for _ in range(2):
        dev = M.EamDevice(cells)
        dev.make_offsets()
        dev.set_potential(setfl)
        dev.set_precision(precision)
        dev.upload(atoms)
        dev.set_timestep(0.0005)
        dev.compute_forces()
        if pka:
            dev.collision_step(*pka)
        out.append(dev)
    return out


def _same(a, b, what):
    sa, ia = a.download()
    sb, ib = b.download()
    va, vb = M.sub_box_view(sa, a), M.sub_box_view(sb, b)
    for k in FIELDS:
        np.testing.assert_array_equal(va[k], vb[k], err_msg="%s: %s" % (what, k))
        np.testing.assert_array_equal(ia[k], ib[k], err_msg="%s: inter %s" % (what, k))
    return len(ia)


def test_batch_steps_equal_single_box_steps(setfl_fe, setfl_fecuni):
    boxes = [
        _box(setfl_fe, (6, 6, 6), 11, 600.0, (1, 0, 0), M.PREC_FP64),
        _box(setfl_fecuni, (8, 7, 9), 12, 300.0, (985, 10, 5), M.PREC_MIXED),
        _box(setfl_fecuni, (10, 10, 10), 13, 300.0, (985, 10, 5), M.PREC_FP64,
             pka=((5, 5, 5, 0), (1.0, 3.0, 5.0), 400.0)),
    ]
    batch = M.EamBatch([b[0] for b in boxes])
    replays = 0
    for n in (7, 6, 1, 9):
        replays = batch.step(n)
        for q, (a, b) in enumerate(boxes):
            b.step(n)
            ni = _same(a, b, "box %d after %d" % (q, n))
            print("box %d: %d steps, inter %d, graph replays %d" % (q, n, ni, replays), flush=True)
    assert replays == 3 + 2 + 0 + 4
    # a change of step length re-captures the graph
    for a, b in boxes:
        a.set_timestep(0.0004)
        b.set_timestep(0.0004)
    batch.step(5)
    for q, (a, b) in enumerate(boxes):
        b.step(5)
        _same(a, b, "box %d after the new step length" % q)
    batch.close()


def test_autotune_leaves_the_trajectory_alone(setfl_fecuni):
    cells = (8, 8, 12)
    a, b = _box(setfl_fecuni, cells, 21, 300.0, (985, 10, 5), M.PREC_MIXED)
    ty, ms = a.autotune(2)
    # mixed precision: every y-tile height with 4 | 8 density and 2 | 4 force gathers in flight
    trials = 4 * len([t for t in range(1, min(32, cells[1]) + 1) if cells[1] % t == 0])
    b.step(2 * trials)
    print("autotune: ty %d, shape %s, %.3f ms/step over %d candidates" % (ty, a.launch_shape(), ms, trials),
          flush=True)
    assert cells[1] % ty == 0 and a.launch_shape()[0] == ty
    _same(a, b, "after autotune")
    a.step(4)
    b.step(4)
    _same(a, b, "after autotune + 4 steps")


@pytest.mark.gpu
def test_launch_shapes_give_one_trajectory(setfl_fecuni):
    """Every launch shape of the mixed passes (y-tile, gathers in flight)
    computes every site the same way: 30 steps of an alloy box at 900 K
    agree bit for bit, and so do their forces."""
    cells = (8, 8, 12)
    ref, _ = _box(setfl_fecuni, cells, 33, 900.0, (900, 70, 30), M.PREC_MIXED_KAHAN)
    ref.set_launch_shape(8, 4, 4)
    ref.step(30)
    for shape in [(1, 8, 2), (4, 4, 2), (2, 8, 4)]:
        dev, _ = _box(setfl_fecuni, cells, 33, 900.0, (900, 70, 30), M.PREC_MIXED_KAHAN)
        dev.set_launch_shape(*shape)
        assert dev.launch_shape() == shape
        dev.step(30)
        _same(ref, dev, "shape %s" % (shape,))
        dev.close()
