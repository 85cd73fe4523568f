"""hipGraph replay of a whole search (tga_plan_set_graph): on every path the plan can take,
graph replays of new observations, through host and device input in any order, return the
same peaks, bit for bit, as eager searches; timed searches report kernel times (graphs with
event-record nodes, or eager launches where the runtime refuses such nodes in a capture);
plans with a Bluestein transform stay eager.  The eager peaks themselves are pinned against the
oracle in test_gpu_plan.py, test_gpu_fp16.py and test_gpu_correlate.py."""
import os
import re

import numpy as np
import pytest
import torch

import tga
from tga import synth

pytestmark = pytest.mark.gpu

FS = 16.368e6
IF = 4.092e6


def grid(center, band, step):
    n = int(round(2 * band / step)) + 1
    return [center - band + i * step for i in range(n)]


FC = grid(IF, 1000.0, 500.0)
FC2 = grid(IF, 900.0, 300.0)
CASES = {
    # name: (fs, L, n_ant, fmt, [(system, prns, fc, nc_epochs)], precision, path in describe)
    "fused": (FS, 2 * 16368 + 100, 2, 2, [("gps", [1, 2, 3], FC, 0), ("sbas", [131], FC, 0)],
              "fp32", "fused16368"),
    "fp16": (FS, 2 * 16368, 2, 1, [("gps", [1, 2, 3], FC, 0)], "fp16", "fused16368-fp16"),
    "noncoherent": (FS, 3 * 16368, 2, 1, [("gps", [1, 2], FC, 3)], "fp32",
                    "fused16368-noncoherent"),
    "split4": (FS, 2 * 65472, 1, 1, [("galileo", [3, 11], grid(IF, 500.0, 500.0), 0)], "fp32",
               "split4x16368"),
    "split10_unfolded": (FS, 163680, 1, 1, [("beidou", [30], grid(IF, 500.0, 500.0), 0)],
                         "fp32", "split10x16368"),
    "subrows4": (FS, 2 * 65472, 1, 1, [("galileo", [3, 11], grid(IF, 500.0, 500.0), 0)],
                   "fp32", "split4x16368"),
    "subrows10_unfolded": (FS, 163680, 1, 1, [("beidou", [30], grid(IF, 500.0, 500.0), 0)],
                             "fp32", "split10x16368"),
    "subrows10": (FS, 2 * 163680, 1, 2, [("l1c", [5, 9], grid(IF, 500.0, 500.0), 0)],
                  "fp32", "split10x16368"),
    "half4": (FS, 2 * 65472, 1, 1, [("galileo", [3, 11], grid(IF, 500.0, 500.0), 0)],
              "fp32", "split4x16368"),
    "half10_unfolded": (FS, 163680, 1, 1, [("beidou", [30], grid(IF, 500.0, 500.0), 0)],
                        "fp32", "split10x16368"),
    "half10": (FS, 2 * 163680, 1, 2, [("l1c", [5, 9], grid(IF, 500.0, 500.0), 0)],
               "fp32", "split10x16368"),
    "fourstep4": (FS, 2 * 65472, 1, 1, [("galileo", [3, 11], grid(IF, 500.0, 500.0), 0)],
                  "fp32", "split132x496"),
    "fourstep10_unfolded": (FS, 163680, 1, 1, [("beidou", [30], grid(IF, 500.0, 500.0), 0)],
                            "fp32", "split330x496"),
    "generic_folded": (2.046e6, 3 * 2046, 2, 2, [("gps", [3, 4], FC2, 0),
                                                  ("qzss", [193], FC2, 0)], "fp32",
                       "generic-folded"),
    "generic_unfolded": (2.046e6, 3000, 2, 1, [("gps", [9], FC2, 0)], "fp32",
                         "generic-unfolded"),
}


def signals_for(systems, fs, seed):
    rng = np.random.default_rng(seed)
    out = []
    for name, prns, fc, _ in systems:
        P = int(round(fs * {"gps": 1e-3, "sbas": 1e-3, "qzss": 1e-3, "galileo": 4e-3,
                            "beidou": 1e-2, "l1c": 1e-2}[name]))
        out.append(dict(system=name, prn=prns[0], cn0=52.0, doppler=float(rng.choice(fc)) - IF,
                        delay=int(rng.integers(0, P)), phase=float(rng.uniform(0, 6.28))))
    return out


# the split cases pin their FP32 form: the sub-row transforms of k_rows (matrix-core radix 31),
# k_subrows, the half sub-rows of k_halfrows, or the four-step form
SPLIT_FORM = {"split4": "1", "split10_unfolded": "1", "subrows4": "2",
              "subrows10_unfolded": "2", "subrows10": "2", "half4": "5", "half10_unfolded": "5",
              "half10": "5", "fourstep4": "3", "fourstep10_unfolded": "3"}


def make_plan(case):
    fs, L, n_ant, fmt, systems, precision, _ = CASES[case]
    old = os.environ.get("TGA_SPLIT_VARIANT")
    if case in SPLIT_FORM:
        os.environ["TGA_SPLIT_VARIANT"] = SPLIT_FORM[case]
    try:
        return _plan(fs, n_ant, L, systems, fmt, precision)
    finally:
        if old is None:
            os.environ.pop("TGA_SPLIT_VARIANT", None)
        else:
            os.environ["TGA_SPLIT_VARIANT"] = old


def _plan(fs, n_ant, L, systems, fmt, precision):
    return tga.Plan(fs, n_ant, L, [tga.SystemSearch(n, f, p, nc_epochs=e)
                                   for n, p, f, e in systems],
                    input_format=fmt, precision=precision)


def observations(case, n=3):
    fs, L, n_ant, fmt, systems, _, _ = CASES[case]
    return [synth.synth_observation(n_ant, L, fs, signals_for(systems, fs, 10 + k), fmt=fmt,
                                    seed=100 + k) for k in range(n)]


def same_peaks(a, b):
    return all(np.array_equal(x[f], y[f]) for x, y in zip(a, b) for f in x.dtype.names)


def graph_launches(plan):
    m = re.search(r"graph=replay launches=(\d+)", plan.describe())
    return int(m.group(1)) if m else 0


@pytest.mark.parametrize("case", list(CASES))
def test_graph_replay_matches_eager(gpu, case):
    obs = observations(case)
    plan = make_plan(case)  # one plan: the same autotuned kernels in both modes
    assert CASES[case][6] in plan.describe() and "graph=pending" in plan.describe()
    plan.set_graph(False)
    assert "graph=off" in plan.describe()
    want = [plan.search(o) for o in obs]
    assert not same_peaks(want[0], want[1])  # the observations differ
    plan.set_graph(True)
    order = [0, 1, 2, 0, 2, 1]
    for k, i in enumerate(order):
        assert same_peaks(plan.search(obs[i]), want[i]), (case, k, i)
    assert graph_launches(plan) == len(order)  # the plan was warm: the first search captures

    # device input at other addresses: copied into the plan's buffer before each replay
    dev = [torch.from_numpy(o).cuda() for o in obs]
    torch.cuda.synchronize()
    for i in (2, 0, 1):
        assert same_peaks(plan.search_device(dev[i].data_ptr()), want[i]), (case, "device", i)
    assert graph_launches(plan) == len(order) + 3


def test_first_search_runs_eagerly(gpu):
    obs = observations("fused", 2)
    plan = make_plan("fused")
    plan.search(obs[0])
    assert "graph=pending" in plan.describe()
    plan.search(obs[1])
    assert graph_launches(plan) == 1


@pytest.mark.parametrize("case", ["fused", "split4", "fourstep4", "generic_folded"])
def test_timing_with_graph(gpu, case):
    # timing: a graph with event-record nodes around the timed stages; its replays set the
    # timers like eager launches do, and switching timing off captures the plain graph again
    obs = observations(case, 2)
    plan = make_plan(case)
    plan.set_graph(False)
    want = [plan.search(o) for o in obs]
    plan.set_graph(True)
    plan.set_timing(True)
    for i in (1, 0):
        assert same_peaks(plan.search(obs[i]), want[i])
    ms, n, _ = plan.timing(0)
    assert n == 2 and ms > 0.0
    # a HIP runtime that refuses event nodes in captures runs timed searches eagerly
    timed_replays = 0 if "timed=eager" in plan.describe() else 2
    plan.set_timing(False)
    assert same_peaks(plan.search(obs[1]), want[1])
    assert graph_launches(plan) == timed_replays + 1


def test_graph_switch_off_and_on(gpu):
    obs = observations("fused", 2)
    plan = make_plan("fused")
    want = [plan.search(o) for o in obs]  # eager, then capture + replay
    plan.set_graph(False)
    assert "graph=off" in plan.describe()
    assert same_peaks(plan.search(obs[0]), want[0])
    plan.set_graph(True)
    for i in (1, 0, 1):
        assert same_peaks(plan.search(obs[i]), want[i])
    assert graph_launches(plan) == 1 + 3  # the replay before the switch, then a new capture


def test_bluestein_plan_stays_eager(gpu):
    # fs = 4.099 MHz: P = 4099 (prime), N = 2P: the generic engine's Bluestein form
    fs, L = 4.099e6, 2 * 4099
    obs = [synth.synth_observation(1, L, fs, [dict(system="gps", prn=5, cn0=55.0,
                                                   doppler=0.0, delay=100 * k + 7)],
                                   fmt=1, seed=k) for k in range(2)]
    plan = tga.Plan(fs, 1, L, [tga.SystemSearch("gps", [IF], [5])])
    assert "graph=off (Bluestein" in plan.describe()
    got = [plan.search(o) for o in obs + obs]
    assert same_peaks(got[0], got[2]) and same_peaks(got[1], got[3])
    assert int(got[1][0][0, 0]["best_codephase"]) == (4099 - 107) % 4099
