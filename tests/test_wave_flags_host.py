"""Host model of the wave-flag phase ordering used by the one-cell z tiles
(pyminiweather_amd/csrc/pmw_fused.hip: wave_flags_init, wave_phase_done).

Each of NW waves runs phases 1..P in order. After finishing phase p a wave
publishes p in its flag, then waits until both neighbouring waves (the edge
waves see a sentinel above every phase) have published p before it starts
phase p+1. A wave's phase reads and writes only its own rows and up to 2
rows of each neighbouring wave, so the kernel is correct if (a) every
schedule of the waves finishes (no deadlock), (b) a wave starts phase p only
when its neighbours have finished phase p-1 (read-after-write), and (c) no
neighbour has started phase p+1 while a wave runs phase p (write-after-read:
no wave gets two phases ahead of a neighbour). The model checks these under
random, adversarial and (for small blocks) all schedules.
"""
import random

import pytest

SENTINEL = 1 << 20


def _run(nw, phases, pick, stop=None, one_sided=False):
    """Simulate one block. A wave alternates two atomic events: run its next
    phase p and publish p; then leave its spin once both neighbours have
    published p. `pick` chooses the next wave among those that can move.
    `stop` = (wave, phase): that wave quits after the phase without
    publishing it; `one_sided`: waves wait for the lower neighbour only
    (both protocol mistakes). Returns the (wave, phase) run log."""
    flag = [SENTINEL] + [0] * nw + [SENTINEL]
    nxt = [1] * nw             # the phase each wave runs next
    state = ["run"] * nw       # "run" | "wait" | "done"
    log = []
    while True:
        movable = []
        for w in range(nw):
            if state[w] == "run":
                movable.append(w)
            elif state[w] == "wait":
                p = nxt[w] - 1
                if flag[w] >= p and (one_sided or flag[w + 2] >= p):
                    movable.append(w)
        if not movable:
            if not all(s == "done" for s in state):
                raise RuntimeError(f"deadlock: states {state}, flags {flag[1:-1]}")
            return log
        w = pick(movable)
        if state[w] == "wait":  # the spin ends
            state[w] = "run"
            continue
        p = nxt[w]
        for nb in (w - 1, w + 1):
            if 0 <= nb < nw:
                assert nxt[nb] >= p, f"wave {w} runs {p} before wave {nb} finished {p - 1}"
                assert nxt[nb] <= p + 1, f"wave {nb} started {p + 1} while wave {w} runs {p}"
        log.append((w, p))
        nxt[w] += 1
        if stop == (w, p):
            state[w] = "done"
            continue
        flag[w + 1] = p
        state[w] = "done" if p == phases else "wait"


@pytest.mark.parametrize("nw", [1, 2, 3, 8, 12])
def test_random_schedules_finish_in_phase_order(nw):
    rng = random.Random(1234 + nw)
    for _ in range(300):
        log = _run(nw, 5, rng.choice)
        assert len(log) == 5 * nw
        for w in range(nw):
            assert [p for v, p in log if v == w] == [1, 2, 3, 4, 5]


@pytest.mark.parametrize("nw", [2, 5, 12])
def test_adversarial_schedules_finish(nw):
    # always move the lowest, the highest or the middle movable wave
    for pick in (min, max, lambda r: r[len(r) // 2]):
        assert len(_run(nw, 5, pick)) == 5 * nw


def _all_schedules(nw, phases):
    """Depth-first over every sequence of choices; returns how many
    complete schedules were run (each checks the invariants)."""
    done = 0
    pending = [[]]
    while pending:
        prefix = pending.pop()
        taken = []

        def pick(movable):
            k = len(taken)
            if k < len(prefix):
                i = prefix[k]
            else:
                i = 0
                for j in range(1, len(movable)):
                    pending.append(taken + [j])
            taken.append(i)
            return movable[i]

        _run(nw, phases, pick)
        done += 1
    return done


def test_every_schedule_of_small_blocks():
    # two waves, three phases, and three waves, two phases
    assert _all_schedules(2, 3) > 10
    assert _all_schedules(3, 2) > 50


def test_a_wave_that_skips_its_flag_stalls_the_block():
    # the mistake the kernel avoids: a wave that stops publishing makes its
    # neighbours wait for ever, which the model reports
    with pytest.raises(RuntimeError, match="deadlock"):
        _run(3, 3, min, stop=(1, 2))


def test_waiting_for_one_neighbour_only_breaks_the_order():
    # waiting on one side is not enough: some schedule lets a wave run a
    # phase before its other neighbour finished the one before
    rng = random.Random(7)
    with pytest.raises(AssertionError):
        for _ in range(200):
            _run(4, 4, rng.choice, one_sided=True)
