"""The batched-march autotuner's staged search (host logic) against a
synthetic plan whose step time is a known function of the launch shape: the
search finds the optimum within its stages, leaves the plan tuned to it, and
never asks for a row group larger than the rows per wavefront."""
import numpy as np

from bristolfe_v2_amd import _lib, autotune

KEYS = {_lib.TUNE_MULTI_ROWS_PER_BLOCK: "rows", _lib.TUNE_MULTI_ROWS_PER_WAVE: "rpw",
        _lib.TUNE_XCD_REMAP: "remap", _lib.TUNE_MULTI_GROUP: "group",
        _lib.TUNE_MULTI_OPERATOR: "op", _lib.TUNE_MULTI_NODES_PER_WAVE: "npw"}


class SyntheticPlan:
    """Stands in for _lib.Plan: run_batched returns zeros, stats() a loop
    time from cost(shape, columns); group 0 means automatic (pairs)."""

    def __init__(self, cost):
        self.cost = cost
        self.shape = {"rows": 4, "rpw": 1, "remap": 1, "group": 0, "op": 0, "npw": 4}
        self.requests = []
        self._ms = 0.0

    def set_tuning(self, key, value):
        self.shape[KEYS[key]] = value

    def used_group(self):
        g = self.shape["group"] or 2
        return min(g, self.shape["rpw"])

    def info(self):
        return {"multi_group": self.used_group()}

    def run_batched(self, n_t, S, frc_dof, frc_src, frcs, hist_dof):
        shape = dict(self.shape, group=self.used_group(), columns=S)
        self.requests.append(shape)
        self._ms = self.cost(shape) * (n_t - 1)
        self._steps = n_t - 1
        return np.zeros((len(hist_dof), S, n_t)), 1

    def stats(self):
        return {"loop_ms": self._ms, "steps": self._steps}


def _cost(s):
    # optimum: 2 wavefronts per workgroup, 4 rows per wavefront, groups of 4,
    # plain order, all 64 columns in one pass
    c = 1.0 + 0.1 * abs(s["rows"] - 2) + 0.05 * abs(np.log2(s["rpw"]) - 2)
    c -= {1: 0.0, 2: 0.2, 4: 0.3, 8: 0.25}[s["group"]]
    c += 0.02 * s["remap"]
    return c * (64 / s["columns"])


def test_staged_search_finds_the_optimum_and_tunes_the_plan():
    plan = SyntheticPlan(_cost)
    fd = np.arange(64)
    best, table = autotune.tune_batched(plan, 64, fd, np.arange(64), np.ones((1, 400)), fd,
                                        probe_steps=20, repeats=1)
    assert (best["rows"], best["rpw"], best["group"], best["remap"], best["tile"]) == (2, 4, 4, 0, 64)
    assert {k: plan.shape[k] for k in ("rows", "rpw", "remap", "group", "op")} == \
        {"rows": 2, "rpw": 4, "remap": 0, "group": 4, "op": 0}
    assert best["operator"] == "csr"   # no element operator on this plan
    assert min(r["ms_per_step"] for r in table) == best["ms_per_step"]
    # every request kept the group within the rows per wavefront
    assert all(r["group"] <= r["rpw"] for r in plan.requests)
    # the 32-column tile was tried, in two passes of 32 columns
    assert any(r["columns"] == 32 for r in plan.requests)


def test_probe_starts_at_the_first_forced_sample():
    seen = []

    class P(SyntheticPlan):
        def run_batched(self, n_t, S, frc_dof, frc_src, frcs, hist_dof):
            seen.append((n_t, np.asarray(frcs).shape[1]))
            return super().run_batched(n_t, S, frc_dof, frc_src, frcs, hist_dof)

    F = np.zeros((1, 500))
    F[0, 37:] = 1.0
    fd = np.arange(8)
    autotune.tune_batched(P(_cost), 8, fd, np.arange(8), F, fd, probe_steps=20, repeats=1,
                          rows=(1, 2), rpws=(1, 2), tiles=(64,))
    assert all(n == 37 + 20 for n, _ in seen)


def test_run_tiled_passes_each_pass_its_own_signal_rows():
    """One signal row per forcing entry (p rows) with a tile narrower than the
    sources: every pass gets exactly the rows of its own entries, and the
    assembled history puts each pass's columns in place."""
    calls = []

    class P:
        def run_batched(self, n_t, S, frc_dof, frc_src, frcs, hist_dof):
            frcs = np.atleast_2d(frcs)
            assert frcs.shape[0] in (1, len(frc_dof))
            calls.append((S, np.array(frc_dof), np.array(frc_src), frcs.copy()))
            h = np.zeros((len(hist_dof), S, n_t))
            for k, (d, s_) in enumerate(zip(frc_dof, frc_src)):
                h[0, s_, :] += d * frcs[k if frcs.shape[0] > 1 else 0]
            return h, 1

        def stats(self):
            return {"loop_ms": 1.0, "steps": 1}

    n_src, n_t = 10, 6
    fd = np.arange(20) + 100
    fs = np.repeat(np.arange(n_src), 2)
    F = np.arange(20 * n_t, dtype=float).reshape(20, n_t)
    h, ms = autotune.run_tiled(P(), n_t, n_src, fd, fs, F, [0], 4)
    assert [c[0] for c in calls] == [4, 4, 2] and ms == 3.0
    for S, d, s_, rows in calls:
        np.testing.assert_array_equal(rows, F[np.isin(fd, d)])
    ref = np.zeros((1, n_src, n_t))
    for k in range(20):
        ref[0, fs[k]] += fd[k] * F[k]
    np.testing.assert_array_equal(h, ref)
    # one broadcast row goes to every pass unchanged
    calls.clear()
    autotune.run_tiled(P(), n_t, n_src, fd, fs, F[:1], [0], 4)
    assert all(c[3].shape[0] == 1 for c in calls)


def test_element_product_stage_when_an_element_operator_is_attached():
    """With an element operator whose groups have pattern kernels, the last
    stages time the multi-source element product and its node-block form at
    2, 4 and 8 nodes per wavefront, and keep the fastest."""

    class P(SyntheticPlan):
        def info(self):
            return {"multi_group": self.used_group(), "operator_kind": 1, "ebe_pattern": 2,
                    "multi_operator": self.shape["op"]}

    def cost(sh):
        if sh.get("op") == 1:
            return 0.5 + 0.01 * abs(sh["npw"] - 4)
        if sh.get("op") == 2:
            return 0.45 + 0.01 * abs(sh["npw"] - 2)
        return _cost(sh)

    class Q(P):
        def run_batched(self, n_t, S, frc_dof, frc_src, frcs, hist_dof):
            shape = dict(self.shape, group=self.used_group(), columns=S)
            self.requests.append(shape)
            self._ms = cost(shape) * (n_t - 1)
            self._steps = n_t - 1
            return np.zeros((len(hist_dof), S, n_t)), 1

    plan = Q(_cost)
    fd = np.arange(64)
    best, table = autotune.tune_batched(plan, 64, fd, np.arange(64), np.ones((1, 400)), fd,
                                        probe_steps=20, repeats=1)
    assert best["operator"] == "node" and best["npw"] == 2
    assert plan.shape["op"] == 2 and plan.shape["npw"] == 2
    assert sorted(r["npw"] for r in table if r["operator"] == "ebe") == [2, 4, 8]
    assert sorted(r["npw"] for r in table if r["operator"] == "node") == [2, 4, 8]
