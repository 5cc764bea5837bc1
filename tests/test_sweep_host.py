"""CPU: the sweep driver and S-parameter host code (no GPU)."""

import numpy as np

from prismo_amd.analysis.sparameters import (SParameterAnalyzer, compute_group_delay,
                                             export_touchstone)
from prismo_amd.optimization.sweep import BatchedRunner, ParameterSweep, SweepParameter


class _FakeBase:
    """What the sweep needs of BatchEngine beyond start/finish: one group key
    for every design, and no design diverged."""

    class _Never:
        def __getitem__(self, d):
            return False

    diverged = _Never()

    @staticmethod
    def design_key(sim, nsteps):
        return 0


def test_combinations_are_the_meshgrid_in_order():
    sw = ParameterSweep([SweepParameter("a", [1, 2]), SweepParameter("b", [10, 20, 30])],
                        lambda p: {"m": p["a"] * p["b"]})
    assert [(c["a"], c["b"]) for c in sw.parameter_combinations] == [
        (1, 10), (1, 20), (1, 30), (2, 10), (2, 20), (2, 30)]
    res = sw.run()
    assert [r["m"] for r in res] == [10, 20, 30, 20, 40, 60]
    assert sw.get_result_array("m").shape == (2, 3)
    best, r = sw.find_optimal("m")
    assert best == {"a": 2, "b": 30} and r["m"] == 60


def test_errors_are_captured_per_design():
    def f(p):
        if p["a"] == 2:
            raise RuntimeError("boom")
        return {"m": 1}

    res = ParameterSweep([SweepParameter("a", [1, 2, 3])], f).run()
    assert [r["status"] for r in res] == ["success", "error", "success"]
    assert res[1]["error"] == "boom"


def test_batched_build_errors_do_not_stop_the_sweep(monkeypatch):
    import prismo_amd.core.batch as B

    ran = {}

    class FakeEngine(_FakeBase):
        def __init__(self, sims, **kw):
            ran["n"] = len(sims)

        def start(self, n, graph=True):
            ran["steps"] = n

        def finish(self):
            pass

    monkeypatch.setattr(B, "BatchEngine", FakeEngine)

    def build(p):
        if p["a"] < 0:
            raise ValueError("bad design")
        return object()

    sw = ParameterSweep.from_samples(
        [{"a": 1}, {"a": -1}, {"a": 2}],
        batched=BatchedRunner(build=build, analyze=lambda s, p: {"x": p["a"]}, steps=7))
    res = sw.run()
    assert ran == {"n": 2, "steps": 7}
    assert [r["status"] for r in res] == ["success", "error", "success"]
    assert res[2]["x"] == 2 and "bad design" in res[1]["error"]


def test_batched_groups_are_pipelined(monkeypatch):
    """Group g+1 is built and queued before group g is waited for and
    analysed; a group whose run fails reports every design as an error."""
    import prismo_amd.core.batch as B

    log = []

    class FakeEngine(_FakeBase):
        def __init__(self, sims, **kw):
            self.ids = [s["a"] for s in sims]

        def start(self, n, graph=True):
            log.append(("start", tuple(self.ids)))

        def finish(self):
            log.append(("finish", tuple(self.ids)))
            if 4 in self.ids:
                raise RuntimeError("device error")

    monkeypatch.setattr(B, "BatchEngine", FakeEngine)

    def build(p):
        log.append(("build", p["a"]))
        return dict(p)

    def analyze(sim, p):
        log.append(("analyze", p["a"]))
        return {"x": p["a"]}

    sw = ParameterSweep.from_samples(
        [{"a": k} for k in range(5)],
        batched=BatchedRunner(build=build, analyze=analyze, steps=3, max_batch=2))
    res = sw.run()
    assert log == [("build", 0), ("build", 1), ("start", (0, 1)),
                   ("build", 2), ("build", 3), ("start", (2, 3)),
                   ("finish", (0, 1)), ("analyze", 0), ("analyze", 1),
                   ("build", 4), ("start", (4,)),
                   ("finish", (2, 3)), ("analyze", 2), ("analyze", 3),
                   ("finish", (4,))]
    assert [r["status"] for r in res] == ["success"] * 4 + ["error"]
    assert "device error" in res[4]["error"]


def test_s_parameters_and_touchstone(tmp_path):
    f = np.linspace(1e14, 2e14, 5)
    sa = SParameterAnalyzer(2, f)
    pin = np.full(5, 4.0)
    sa.add_port_data(0, 0, pin, np.full(5, 0.04))
    sa.add_port_data(1, 0, np.full(5, 1.0), np.zeros(5))
    assert np.allclose(sa.get_s_parameter(1, 0), 0.5)
    assert np.allclose(sa.get_s_parameter(0, 0), 0.1)
    assert np.allclose(sa.get_insertion_loss_db(1, 0), -20 * np.log10(0.5))
    path = tmp_path / "dev.s2p"
    export_touchstone(path, f, sa.s_matrix, comments=["test"])
    lines = path.read_text().splitlines()
    assert lines[1] == "! test" and lines[2] == "# Hz S RI R 50.0"
    row = [float(v) for v in lines[3].split()]
    # frequency, then S11 re/im, S12, S21, S22 (row-major)
    assert row[0] == f[0] and row[1] == 0.1 and row[5] == 0.5
    # group delay of a pure delay tau: exp(-i w tau)
    tau = 3e-15
    gd = compute_group_delay(f, np.exp(-1j * 2 * np.pi * f * tau))
    assert np.allclose(gd, tau)


def test_batched_first_group_size(monkeypatch):
    """``first_batch`` sizes the first group; the rest are ``max_batch``."""
    import prismo_amd.core.batch as B

    groups = []

    class FakeEngine(_FakeBase):
        def __init__(self, sims, **kw):
            groups.append([s["a"] for s in sims])

        def start(self, n, graph=True):
            pass

        def finish(self):
            pass

    monkeypatch.setattr(B, "BatchEngine", FakeEngine)
    sw = ParameterSweep.from_samples(
        [{"a": k} for k in range(7)],
        batched=BatchedRunner(build=dict, analyze=lambda s, p: {}, steps=1, max_batch=3,
                              first_batch=1))
    assert all(r["status"] == "success" for r in sw.run())
    assert groups == [[0], [1, 2, 3], [4, 5, 6]]


def test_batched_diverged_design_is_an_error_alone(monkeypatch):
    """A design the kernel flags as diverged (Inf/NaN) reports 'error' and is
    not analysed; the other designs of its group are."""
    import prismo_amd.core.batch as B

    analysed = []

    class FakeEngine(_FakeBase):
        def __init__(self, sims, **kw):
            self.ids = [s["a"] for s in sims]

        def start(self, n, graph=True):
            pass

        def finish(self):
            self.diverged = [a == 2 for a in self.ids]

    monkeypatch.setattr(B, "BatchEngine", FakeEngine)
    sw = ParameterSweep.from_samples(
        [{"a": k} for k in range(4)],
        batched=BatchedRunner(build=dict, analyze=lambda s, p: analysed.append(p["a"]) or {},
                              steps=1))
    res = sw.run()
    assert [r["status"] for r in res] == ["success", "success", "error", "success"]
    assert "diverged" in res[2]["error"]
    assert analysed == [0, 1, 3]


def test_unbatchable_design_runs_alone(monkeypatch):
    """A design the batch cannot take (design_key raises) or of another key
    runs through its own Simulation; its group still runs batched, and a
    single that fails is the only error."""
    import prismo_amd.core.batch as B

    groups, alone = [], []

    class FakeEngine(_FakeBase):
        def __init__(self, sims, **kw):
            groups.append([s.a for s in sims])

        def start(self, n, graph=True):
            pass

        def finish(self):
            pass

        @staticmethod
        def design_key(sim, nsteps):
            if sim.a in (1, 5):
                raise NotImplementedError("lossy media")
            return "other" if sim.a == 3 else "main"

    class Sim:
        def __init__(self, p):
            self.a = p["a"]

        def run_steps(self, n):
            if self.a == 5:
                raise RuntimeError("single failed")
            alone.append((self.a, n))

    monkeypatch.setattr(B, "BatchEngine", FakeEngine)
    sw = ParameterSweep.from_samples(
        [{"a": k} for k in range(6)],
        batched=BatchedRunner(build=Sim, analyze=lambda s, p: {"a": s.a}, steps=9))
    res = sw.run()
    assert groups == [[0, 2, 4]]
    assert alone == [(1, 9), (3, 9)]
    assert [r["status"] for r in res] == ["success"] * 5 + ["error"]
    assert [r.get("a") for r in res[:5]] == [0, 1, 2, 3, 4]
    assert "single failed" in res[5]["error"]


def test_mode_data_keeps_the_reference_quotients():
    """add_mode_data: S_ii = backward/forward; S_ij = forward/backward as the
    reference defines it (sparameters.py:113-122) unless an 'incident'
    coefficient is given, then forward/incident."""
    f = np.array([1e14, 2e14])
    sa = SParameterAnalyzer(2, f)
    fwd = np.array([2.0 + 1.0j, 4.0])
    bwd = np.array([0.5, 1.0 - 1.0j])
    sa.add_mode_data(0, 0, {"forward": fwd, "backward": bwd})
    assert np.allclose(sa.get_s_parameter(0, 0), bwd / fwd)
    sa.add_mode_data(1, 0, {"forward": fwd, "backward": bwd})
    assert np.allclose(sa.get_s_parameter(1, 0), fwd / bwd)
    inc = np.array([8.0, 8.0j])
    sa.add_mode_data(1, 0, {"forward": fwd, "backward": bwd, "incident": inc})
    assert np.allclose(sa.get_s_parameter(1, 0), fwd / inc)
