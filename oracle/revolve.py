"""Oracle restatement of the binomial checkpointing schedule ("revolve",
Griewank & Walther, ACM TOMS Algorithm 799) as used by the reference's
Revolve<T> class (revolve.cpp:8-305).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinned against the compiled reference's action sequences
(tests/golden/revolve_*.npz) and against the closed form of the number of
forward steps, numforw = reps*steps - C(reps+snaps, snaps)*reps/(snaps+1).
"""
from __future__ import annotations

ADVANCE, TAKESHOT, RESTORE, FIRSTURN, YOUTURN, TERMINATE, ERROR = range(7)
NAMES = ("advance", "takeshot", "restore", "firsturn", "youturn", "terminate", "error")
CHECKUP = 512   # revolve.h:17
REPSUP = 64     # revolve.h:18
MAXINT = 2147483647


def _cdiv(a: int, b: int) -> int:
    """C integer division (truncation toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def numforw(steps: int, snaps: int) -> int:
    """Forward steps used by the schedule (revolve.cpp:31-62)."""
    if snaps < 1 or snaps > CHECKUP:
        return -1
    reps, rng = 0, 1
    while rng < steps:
        reps += 1
        rng = _cdiv(rng * (reps + snaps), reps)
    if reps > REPSUP:
        return -1
    return reps * steps - _cdiv(rng * reps, snaps + 1)


def maxrange(ss: int, tt: int) -> int:
    """C(ss+tt, tt) evaluated in double, clamped to MAXINT (revolve.cpp:86-111)."""
    if tt < 0 or ss < 0:
        return -1
    res = 1.0
    for i in range(1, tt + 1):
        res *= (ss + i)
        res /= i
        if res > MAXINT:
            return MAXINT
    return int(res)


def adjust(steps: int) -> int:
    """Snapshot count balancing memory and recomputation (revolve.cpp:114-147)."""
    snaps, reps, s = 1, 1, 0
    while maxrange(snaps + s, reps + s) > steps:
        s -= 1
    while maxrange(snaps + s, reps + s) < steps:
        s += 1
    snaps += s
    reps += s
    s = -1
    while maxrange(snaps, reps) >= steps:
        if snaps > reps:
            snaps -= 1
            s = 0
        else:
            reps -= 1
            s = 1
    if s == 0:
        snaps += 1
    if s == 1:
        reps += 1
    return snaps


class Revolve:
    """State machine returning one action per call, like Revolve::revolve()."""

    def __init__(self, steps: int, snaps: int):
        self.steps = steps
        if snaps > steps or snaps < 2:
            snaps = adjust(steps)
        self.snaps = snaps
        self.capo, self.fine, self.check = 0, steps, -1
        self.ch = [0] * CHECKUP
        self.turn = 0
        self.oldfine = 0
        self.oldsnaps = snaps
        self.info = 0

    def _advance_target(self) -> int:
        """Where the next advance stops: binomial split of [capo, fine) for
        the checkpoints still free (ds = snaps - check)."""
        ds = self.snaps - self.check
        span = self.fine - self.capo
        reps, rng = 0, 1
        while rng < span:
            reps += 1
            rng = _cdiv(rng * (reps + ds), reps)
        if reps > REPSUP:
            return -1
        b1 = _cdiv(rng * reps, ds + reps)
        b2 = _cdiv(b1 * ds, ds + reps - 1) if ds > 1 else 1
        if ds == 1:
            b3 = 0
        else:
            b3 = _cdiv(b2 * (ds - 1), ds + reps - 2) if ds > 2 else 1
        b4 = _cdiv(b2 * (reps - 1), ds)
        if ds < 3:
            b5 = 0
        else:
            b5 = _cdiv(b3 * (ds - 2), reps) if ds > 3 else 1
        if span <= b1 + b3:
            target = self.capo + b4
        elif span >= rng - b5:
            target = self.capo + b1
        else:
            target = self.fine - b2 - b3
        return target if target != self.capo else self.capo + 1

    def step(self) -> int:
        if self.check < -1 or self.capo > self.fine:
            return ERROR
        if self.check == -1 and self.capo < self.fine:
            self.turn = 0
            self.ch[0] = self.capo - 1
        span = self.fine - self.capo
        if span == 0:
            if self.check == -1 or self.capo == self.ch[0]:
                self.check -= 1
                return TERMINATE
            self.capo = self.ch[self.check]
            self.oldfine = self.fine
            return RESTORE
        if span == 1:
            self.fine -= 1
            if self.check >= 0 and self.ch[self.check] == self.capo:
                self.check -= 1
            self.oldfine = self.fine
            if self.turn == 0:
                self.turn = 1
                return FIRSTURN
            return YOUTURN
        if self.check == -1 or self.ch[self.check] != self.capo:
            self.check += 1
            if self.check >= CHECKUP or self.check + 1 > self.snaps:
                return ERROR
            self.ch[self.check] = self.capo
            if self.check == 0:
                self.oldsnaps = self.snaps
            self.oldfine = self.fine
            return TAKESHOT
        if self.oldfine < self.fine and self.snaps == self.check + 1:
            return ERROR
        if self.snaps - self.check < 1:
            return ERROR
        target = self._advance_target()
        if target < 0:
            return ERROR
        self.capo = target
        self.oldfine = self.fine
        return ADVANCE


def schedule(steps: int, snaps: int):
    """Full action list [(action, capo, check), ...] up to terminate/error."""
    r = Revolve(steps, snaps)
    out = []
    while True:
        a = r.step()
        out.append((a, r.capo, r.check))
        if a in (TERMINATE, ERROR):
            return out
