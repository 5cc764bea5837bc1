"""GPU: the fused one-step kernel with each padded row in two workgroups
(row segments; pm_set_row_split), against whole rows and against O1.

A segment's first cell takes E(n) of its left neighbour from memory and its
last cell's E update takes H(n+1/2) of the cell to its right, recomputed
with the arithmetic the neighbouring segment uses for that cell.  Cases:
256-cell rows split into 128-cell segments (FP32 with FP16 Cb, FP64), CPML
on all faces, two dipoles, a dispersive box across the segment boundary,
several tile heights, z-chunks and tile orders; 512-cell FP64 rows, which
only the split instances step fused, against O1 (1e-10).
"""

import numpy as np
import pytest

from tests.conftest import rel_l2
from tests.test_gpu_fused import D, _all, _drive, _fields, _oracle, _sim

pytestmark = pytest.mark.gpu


def _run(shape, dtype, storage, split, steps, ty=2, kchunk=8, order=3, ade=False, vec=0):
    import prismo_amd as pa

    sim, eps = _sim(*shape, t=5, dtype=dtype, storage=storage, nsrc=2)
    if ade:
        si = pa.LorentzMaterial(7.98, [pa.LorentzPole(omega_0=1.1e16, delta_epsilon=3.69,
                                                      gamma=1e13)])
        # x from 100 to 150 physical nodes: across the segment boundary at 128
        sim.add_dispersive_box(si, (95 * D, 4 * D, 4 * D), (145 * D, 14 * D, 12 * D))
    eng = sim.solver.updater.engine_for(sim.fields)
    eng.ctx.set_tuning(True, ty, kchunk, vec, order)
    eng.ctx.set_row_split(split)
    assert eng.ctx.row_split() == split, (split, eng.ctx.row_split())
    sim.run_steps(steps)
    return sim, eps


@pytest.mark.parametrize("dtype,storage", [(np.float32, "fp16"), (np.float64, "auto")])
@pytest.mark.parametrize("ty,kchunk,order", [(2, 8, 3), (3, 5, 2), (4, 12, 0)])
def test_split_rows_equal_whole_rows(mi355x, dtype, storage, ty, kchunk, order):
    shape = (200, 30, 28)
    whole, _ = _run(shape, dtype, storage, 1, 61, ty, kchunk, order)
    split, _ = _run(shape, dtype, storage, 2, 61, ty, kchunk, order)
    a, b = _all(_fields(whole)), _all(_fields(split))
    print(f"{np.dtype(dtype).name} ty {ty}: max |diff| {np.max(np.abs(a - b)):.3e}, "
          f"rel L2 {rel_l2(b, a):.2e}")
    assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype,storage", [(np.float32, "auto"), (np.float64, "auto")])
def test_split_rows_dispersive_box_across_boundary(mi355x, dtype, storage):
    shape = (200, 30, 28)
    whole, _ = _run(shape, dtype, storage, 1, 57, ade=True)
    split, _ = _run(shape, dtype, storage, 2, 57, ade=True)
    assert np.array_equal(_all(_fields(whole)), _all(_fields(split)))


def test_fp64_512_cell_rows_fused_vs_oracle(mi355x):
    """Rows of 480 cells (512 padded) in FP64: the automatic choice is two
    segments of 256, the only fused shape for them; against O1."""
    sim, eps = _sim(480, 22, 20, t=5, dtype=np.float64, nsrc=2)
    eng = sim.solver.updater.engine_for(sim.fields)
    assert eng.ctx.fused_steps() == 1 and eng.ctx.row_split() == 2
    steps = 80
    sim.run_steps(steps)
    o1 = _oracle(sim, eps, 5)
    _drive(o1, sim, steps)
    err = rel_l2(_all(_fields(sim)), _all(o1.F))
    print(f"FP64 480-cell rows, two segments, vs O1: {err:.2e}")
    assert err <= 1e-10


@pytest.mark.parametrize("vec,ty,ade", [(1, 3, False), (2, 4, False), (1, 3, True), (2, 4, True)])
def test_fp32_512_cell_rows_split_vs_whole(mi355x, vec, ty, ade):
    """FP32 480-cell rows: two 256-cell segments (one cell per lane and
    3-row tiles, or two cells per lane and 4-row tiles) against the
    whole-row two-cell-per-lane kernel (different instances, same
    arithmetic per cell), with and without a dispersive box."""
    shape = (480, 22, 20)
    whole, _ = _run(shape, np.float32, "auto", 1, 40, ty=2, ade=ade)
    split, _ = _run(shape, np.float32, "auto", 2, 40, ty=ty, vec=vec, ade=ade)
    assert np.array_equal(_all(_fields(whole)), _all(_fields(split)))


@pytest.mark.parametrize("dtype,storage,bar", [(np.float32, "fp16", 1e-5),
                                               (np.float64, "auto", 1e-10)])
@pytest.mark.parametrize("ty,kchunk,order", [(5, 7, 3), (7, 12, 2)])
def test_tall_tiles_of_segments_vs_oracle(mi355x, dtype, storage, bar, ty, kchunk, order):
    """5- and 7-row tiles of 128-cell segments (16 waves), which whole
    256-cell rows cannot take, against O1 (fed the same FP16-rounded Cb in
    FP32).  Whole rows are not a bitwise reference here: another tile height
    moves cells between the kernel's slab and interior instances, which may
    round differently."""
    from oracle import compat as O0

    steps = 61
    sim, eps = _run((200, 30, 28), dtype, storage, 2, steps, ty, kchunk, order)
    eng = sim.solver.updater.engine_for(sim.fields)
    cb = None
    if storage == "fp16":
        _, Cb, _, _ = O0.coefficients(eps, np.ones_like(eps), np.zeros_like(eps),
                                      np.zeros_like(eps), sim.dt)
        scale = eng.cb_scale
        cb = {c: (f(Cb) / scale).astype(np.float16).astype(np.float64) * scale
              for c, f in (("Ex", O0._ex), ("Ey", O0._ey), ("Ez", O0._ez))}
    o1 = _oracle(sim, eps, 5, cb=cb)
    _drive(o1, sim, steps)
    err = rel_l2(_all(_fields(sim)), _all(o1.F))
    print(f"{np.dtype(dtype).name} ty {ty} segments vs O1: {err:.2e}")
    assert err <= bar
