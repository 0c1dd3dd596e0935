"""GPU checks of the autotuner (starsh_hip_autotune_batch,
starsh_hip_autotune) and of the schedules it records.

A schedule changes only how a compression is run (GEMM block shapes,
Gram-Schmidt panel width, which rank bucket a tile starts in, and whether a
tile known to be false far-field is stored densely at once), never the
accuracy contract: every compression with a schedule must meet the
requested tolerance (global and per-tile relative Frobenius errors from
FP64 regenerated tiles).  A per-tile schedule replayed on the problem it
was recorded on must reproduce the recorded run: the same rank for every
tile, each tile computed in one round.

Problems use kernel parameters that no other test uses, and the schedule
table is emptied before and after each test, so tuned knobs never reach
other tests.
"""
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import starsh_hip as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(msg):
    print(msg, flush=True)


def _opt(name, value):
    assert S.lib().starsh_hip_set_option(name.encode(),
                                         C.c_double(value)) == 0, name


@pytest.fixture(autouse=True)
def _clean_table():
    _opt("schedule_clear", 0)
    yield
    _opt("schedule_clear", 0)
    _opt("tile_schedule", 1)
    _opt("mixed_first_bucket", -1)
    _opt("mixed_ladder", -1)


def _problem(beta, seed=7, N=20000, nb=1000):
    return S.TLRProblem("spatial", N, 2, nb, symm="S", seed=seed,
                        kernel="exp", beta=beta)


def _check_accuracy(p, tol):
    g, terr, tnorm = p.dfe_tiles()
    big = tnorm > 1e-12 * tnorm.max()
    assert g <= tol, g
    assert np.all(terr[big] <= 2 * tol), terr.max()
    return g


def _run(make, fn, maxrank, tol):
    """Compresses a fresh problem (a compression moves false far-field
    tiles to the near field of its format); returns the far-field tiles
    before, {tile: rank} of those still far-field after, and the stats."""
    p = make()
    before = p.tiles()
    assert p.approximate(maxrank, tol, 0, fn) == 0
    st = S.stats()
    far = dict(zip(p.tiles(), p.ranks().tolist()))
    _check_accuracy(p, tol)
    p.free_matrix()
    return before, far, st


def test_mixed_tuning_records_and_replays_tile_schedule():
    """Mixed form: first bucket, per-tile schedule and panel width; the
    replay gives the recorded run's ranks with one round per tile."""
    maxrank, tol = 120, 1e-7
    fn = S.init("HIP_MIXED", oversample=16)

    def make():
        return _problem(0.2731)
    _say("case: mixed autotune, 2D exp beta 0.2731 N=20000 nb=1000 "
         "maxrank 120")
    t = S.autotune_batch([make()], maxrank, tol, form=1)
    _say(f"  tuning {t}")
    assert t["ladder"] in (0, 1)
    assert 0 <= t["first_bucket"] <= 4 + t["ladder"]
    assert t["qr_panel"] in (0, 4, 8)
    assert t["untuned_seconds"] > 0 and t["schedule_seconds"] > 0
    assert t["ladder_seconds"] > 0
    assert t["best_seconds"] <= t["untuned_seconds"]
    assert t["best_seconds"] <= t["ladder_seconds"]
    if t["ladder"] == 0:
        assert t["first_seconds"][t["first_bucket"]] > 0
    assert t["ntrials"] >= 5
    # the recorded run: chosen first bucket and ladder, no per-tile schedule
    _opt("mixed_first_bucket", t["first_bucket"])
    _opt("mixed_ladder", t["ladder"])
    tiles0, far0, st0 = _run(make, fn, maxrank, tol)
    _opt("mixed_first_bucket", -1)
    _opt("mixed_ladder", -1)
    ndense = len(tiles0) - len(far0)
    assert ndense > 0 and st0.nfalse_far == ndense
    # through the table: the per-tile schedule when the tuner kept it
    _, far1, st1 = _run(make, fn, maxrank, tol)
    _say(f"  rounds recorded run {st0.tile_rounds}, replay "
         f"{st1.tile_rounds}; dense tiles {ndense}")
    if t["tile_schedule"]:
        assert far1 == far0
        assert st1.tile_rounds == len(tiles0) - ndense
        assert st1.early_dense == ndense
        assert st0.tile_rounds >= st1.tile_rounds
    # with per-tile schedules off, the table's first bucket only
    _opt("tile_schedule", 0)
    _, far2, st2 = _run(make, fn, maxrank, tol)
    assert far2 == far0
    assert st2.tile_rounds == st0.tile_rounds


def test_tile_schedule_for_other_kernel_parameters():
    """A per-tile schedule recorded for beta=0.2467 applied to beta=0.2 and
    0.35 on the same points (parameter sweep): accuracy holds."""
    maxrank, tol = 120, 1e-7
    fn = S.init("HIP_MIXED", oversample=16)
    t = S.autotune_batch([_problem(0.2467, seed=8)], maxrank, tol, form=1)
    _say(f"case: tuned beta 0.2467: {t}")
    for beta in (0.2, 0.35):
        _say(f"case: schedule of beta 0.2467 on beta {beta}")
        _opt("tile_schedule", 0)
        tiles, farq, stq = _run(lambda: _problem(beta, seed=8), fn, maxrank,
                                tol)
        _opt("tile_schedule", 1)
        _, far, st = _run(lambda: _problem(beta, seed=8), fn, maxrank, tol)
        _say(f"  far tiles {len(far)} (without schedule {len(farq)}); "
             f"rounds {st.tile_rounds} ({stq.tile_rounds})")


def test_plain_tuning_keeps_accuracy():
    maxrank, tol = 200, 1e-7
    fn = S.init("HIP", oversample=16)

    def make():
        return _problem(0.1873, seed=9, N=16000, nb=1600)
    _say("case: plain autotune, 2D exp beta 0.1873 N=16000 nb=1600")
    _, far0, _ = _run(make, fn, maxrank, tol)
    t = S.autotune_batch([make()], maxrank, tol, form=0)
    _say(f"  tuning {t}")
    assert t["first_bucket"] == -1 and t["schedule_seconds"] < 0
    assert t["sketch_bn"] in (0, 128, 256) and t["sketch_rw"] in (0, 1, 2)
    assert t["proj_bn"] in (0, 64, 128) and t["proj_rw"] in (0, 1, 2)
    assert t["qr_panel"] in (0, 4, 8)
    assert t["best_seconds"] <= t["untuned_seconds"]
    _, far1, _ = _run(make, fn, maxrank, tol)
    # GEMM block shapes change the order of the sums only
    assert far1.keys() == far0.keys()
    assert max(abs(far1[k] - far0[k]) for k in far0) <= 2


def test_batch_tuning():
    """A batch of four problems tuned as one, then compressed by
    starsh_blrm_approximate_batch with the recorded schedules."""
    maxrank, tol = 150, 1e-7
    S.init("HIP_MIXED", oversample=16)
    specs = [(6000, 768, 0.0913), (9000, 960, 0.1617), (12000, 1088, 0.2213),
             (7000, 832, 0.2871)]

    def make():
        return [S.TLRProblem("spatial", N, 2, nb, symm="S", seed=200 + i,
                             kernel="exp", beta=b)
                for i, (N, nb, b) in enumerate(specs)]
    _say("case: batch autotune, 4 2D exp problems")
    t = S.autotune_batch(make(), maxrank, tol, form=1)
    _say(f"  tuning {t}")
    assert 0 <= t["first_bucket"] <= 4 + t["ladder"]
    ps = make()
    nfar = sum(len(p.tiles()) for p in ps)
    assert S.approximate_batch(ps, maxrank, tol) == 0
    st = S.stats()
    ndense = nfar - sum(len(p.tiles()) for p in ps)
    _say(f"  rounds {st.tile_rounds}, far tiles {nfar}, dense {ndense}")
    if t["tile_schedule"]:
        assert st.tile_rounds == nfar - ndense
    for p in ps:
        _check_accuracy(p, tol)
        p.free_matrix()


_CHILD = r"""
import json, sys
sys.path[:0] = [sys.argv[1] + "/stars-h_amd", sys.argv[1]]
import ctypes as C
import starsh_hip as S
fn = S.init("HIP_MIXED", oversample=16)


def make():
    return S.TLRProblem("spatial", 20000, 2, 1000, symm="S", seed=10,
                        kernel="exp", beta=0.2589)


def run():
    p = make()
    assert p.approximate(120, 1e-7, 0, fn) == 0
    st = S.stats()
    g, terr, tnorm = p.dfe_tiles()
    far = [[i, j, r] for (i, j), r in zip(p.tiles(), p.ranks().tolist())]
    p.free_matrix()
    return {"tile_rounds": st.tile_rounds, "early_dense": st.early_dense,
            "err": g, "far": far}


out = {}
if sys.argv[2] == "tune":
    p = make()
    out["tuning"] = S.autotune_batch([p], 120, 1e-7, form=1)
    out["tiles"] = [list(t) for t in p.tiles()]
else:
    out["replay"] = run()
    L = S.lib()
    assert L.starsh_hip_set_option(b"tile_schedule", C.c_double(0)) == 0
    assert L.starsh_hip_set_option(b"mixed_first_bucket",
                                   C.c_double(4)) == 0
    out["bucket4"] = run()
print("RESULT " + json.dumps(out), flush=True)
"""


def _child(mode, path):
    env = dict(os.environ, STARSH_HIP_SCHEDULE=str(path))
    r = subprocess.run([sys.executable, "-c", _CHILD, ROOT, mode], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")][-1]
    return json.loads(line[7:])


def test_schedule_file_round_trip(tmp_path):
    """The tuner appends its entry to STARSH_HIP_SCHEDULE; the entry with
    its per-tile schedule replaced by a hand-made one (ladder 0, every tile
    in the widest bucket, every 7th stored densely) is read by a new process,
    which
    must follow it: the marked tiles move to the near field, the others get
    the ranks of a compression that starts every tile in the widest bucket,
    each tile in one round."""
    path = tmp_path / "schedule.txt"
    _say("case: schedule file, tuning process")
    res = _child("tune", path)
    t, tiles = res["tuning"], [tuple(x) for x in res["tiles"]]
    nfar = len(tiles)
    entry = [x for x in path.read_text().split("\n") if x.strip()]
    assert len(entry) == 1
    fields = entry[0].split()
    assert len(fields) == 24
    assert int(fields[13]) == 1 and int(fields[14]) == t["first_bucket"]
    assert int(fields[15]) == t["ladder"]
    if t["tile_schedule"]:
        assert int(fields[22]) == len(fields[23]) == nfar
        assert set(fields[23]) <= set("012345d"[:5 + t["ladder"]] + "d")
    else:
        assert fields[22] == "0" and fields[23] == "-"
    sched = "".join("d" if i % 7 == 3 else "4" for i in range(nfar))
    fields[15] = "0"
    fields[22], fields[23] = str(nfar), sched
    path.write_text(" ".join(fields) + "\n")
    _say("case: schedule file, replaying process")
    r = _child("replay", path)
    rep, b4 = r["replay"], r["bucket4"]
    assert rep["err"] <= 1e-7
    marked = {tiles[i] for i in range(nfar) if sched[i] == "d"}
    far_rep = {(i, j): k for i, j, k in rep["far"]}
    far_b4 = {(i, j): k for i, j, k in b4["far"]}
    assert rep["early_dense"] == len(marked)
    assert rep["tile_rounds"] == nfar - len(marked)
    assert far_rep == {k: v for k, v in far_b4.items() if k not in marked}


def test_schedule_file_malformed_is_ignored(tmp_path):
    """A schedule file the loader cannot read (an entry in an older
    22-field layout, then text) is ignored: a new process compresses as it
    does without a file (same ranks, same rounds, error within the
    bound)."""
    _say("case: schedule file, absent")
    none = _child("replay", tmp_path / "absent.txt")
    bad = tmp_path / "bad.txt"
    bad.write_text("0 2 0.2589 0.5 0 1 1 1 20000 1000 120 16 1e-07 1 1 4 "
                   "0 0 0 0 12345 0 -\nnot a schedule\n")
    _say("case: schedule file, malformed")
    r = _child("replay", bad)
    assert r["replay"]["err"] <= 1e-7
    assert r["replay"]["far"] == none["replay"]["far"]
    assert r["replay"]["tile_rounds"] == none["replay"]["tile_rounds"]
    assert r["replay"]["early_dense"] == none["replay"]["early_dense"]
