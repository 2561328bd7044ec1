"""A plain-Python statement of the reference's best-ant selection and stop
rules, used by the tests as an independent check of the oracle and the GPU
bookkeeping (written from src/cuda/aco_cuda.cu, not from either of them).

* significant(prev, new, min_rel): is_significant_improvement
  (aco_cuda.cu:83-94), ACO_EPS = 1e-9 (include/aco.h:8).
* Spec.step(cost, unvisited): one epoch of the loop at aco_cuda.cu:322-399:
  the iteration best is the feasible ant of lowest cost, ties to the lower
  ant (select_iter_best_host, :104-126); a strictly lower cost replaces the
  global best, and the stagnation counter restarts only when that gain is
  significant; an epoch without a better feasible ant, or without any
  feasible ant, adds one to the counter; the loop stops once the counter
  reaches the stagnation limit (:337) or the epoch count reaches
  ACO_SOLVER_FIXED_EPOCHS.
"""
import sys

DBL_MAX = sys.float_info.max
EPS = 1e-9
INT64_MAX = 2 ** 63 - 1


def significant(prev, new, min_rel):
    if prev >= DBL_MAX or new >= DBL_MAX:
        return new < prev
    if new >= prev - EPS:
        return False
    return (prev - new) / max(prev, EPS) + EPS >= min_rel


class Spec:
    def __init__(self, stagnation=0, min_rel=1e-3, fixed=0):
        self.stagnation, self.min_rel, self.fixed = stagnation, min_rel, fixed
        self.epoch = 0
        self.iter_since_best = 0
        self.global_best = None  # None = no feasible tour yet
        self.halted = False
        self.found = False
        self.best_ant = -1
        self.improved = False
        self.iter_best = None

    def step(self, cost, unvisited):
        assert not self.halted
        best = -1
        for a, (c, u) in enumerate(zip(cost, unvisited)):
            if u == 0 and (best < 0 or c < cost[best]):
                best = a
        self.found = best >= 0
        self.best_ant = best
        self.improved = False
        if best >= 0:
            c = int(cost[best])
            self.iter_best = c
            prev = DBL_MAX if self.global_best is None else float(self.global_best)
            if self.global_best is None or c < self.global_best:
                self.improved = True
                self.iter_since_best = 0 if significant(prev, float(c), self.min_rel) \
                    else self.iter_since_best + 1
                self.global_best = c
            else:
                self.iter_since_best += 1
        else:
            self.iter_best = None
            self.iter_since_best += 1
        self.epoch += 1
        if self.fixed > 0:
            self.halted = self.epoch >= self.fixed
        elif self.stagnation > 0 and self.iter_since_best >= self.stagnation:
            self.halted = True

    def as_state(self):
        return {"epoch": self.epoch, "iter_since_best": self.iter_since_best,
                "halted": int(self.halted), "found": int(self.found),
                "best_ant": self.best_ant, "improved": int(self.improved),
                "iter_best_cost": INT64_MAX if self.iter_best is None else self.iter_best,
                "global_best_cost": INT64_MAX if self.global_best is None else self.global_best}
