"""The run pipeline behind main.py: one solver over a dataset's instances and a range of seeds.

Behaviour kept from the reference's pipeline (utils/core.py:70-118), because scripts and result
files depend on it:

* instances run in increasing order of the first number in their file name ("G7" before "G12");
* the solver is built as ``cls(instance_name=, dataset=, run_name=, random_seed=, **params)`` and
  ``solve(J)`` returns (energy, spins, seconds); params the constructor does not name are dropped
  with a warning;
* seeds are random_seed, random_seed + 1, ... (num_seeds of them);
* an optional warm-up solve of a small synthetic instance comes first, and its result files are
  removed;
* a failed load counts as a failure for every seed of that instance; a failed solve is recorded as
  "out of memory" or its exception type, and the run goes on;
* the summary carries run_name, dataset, total_runs, successful_runs, failed_runs and failures
  (instance, seed, message), which main.py turns into exit status 2 when anything failed.

The structure is this module's own: a RunPlan resolved once from the config, a list of Jobs, and a
RunLog that counts outcomes.
"""
from __future__ import annotations

import inspect
import re
import traceback
import warnings
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Callable

from .data_loader import DataLoader
from .failures import is_out_of_memory
from .selector import get_solver_class
from .solution_io import save_solution, solution_path

# keys the pipeline consumes itself; every other config key is offered to the solver's constructor
PIPELINE_KEYS = frozenset({'solver', 'dataset', 'instance', 'num_seeds', 'random_seed', 'run_name',
                           'save_spins', 'warmup', 'datasets_path'})
WARMUP_INSTANCE = '__warmup__.txt'
_FIRST_NUMBER = re.compile(r"\d+")


def instance_size_key(name: str) -> int:
    """Sort key of an instance file: the first number in its name without the extension, 0 if none."""
    stem = name.rsplit('.', 1)[0]
    found = _FIRST_NUMBER.search(stem)
    return int(found.group()) if found else 0


def constructor_params(solver_class, offered: dict[str, Any]) -> dict[str, Any]:
    """The offered params the solver's constructor names explicitly (a **kwargs catch-all does not
    count); the others trigger one warning listing them."""
    accepted = set()
    for name, p in inspect.signature(solver_class.__init__).parameters.items():
        if p.kind is not inspect.Parameter.VAR_KEYWORD:
            accepted.add(name)
    dropped = sorted(set(offered) - accepted)
    if dropped:
        warnings.warn(f"config keys {dropped} are not parameters of {solver_class.__name__} and are ignored",
                      UserWarning)
    return {k: v for k, v in offered.items() if k in accepted}


def synthetic_warmup_couplings(n: int = 64, density: float = 0.2, seed: int = 0):
    """A small symmetric +-1 coupling matrix (no diagonal) for the warm-up solve."""
    import numpy as np
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    mask = np.triu(rng.random((n, n)) < density, 1)
    signs = rng.choice(np.array([-1.0, 1.0]), size=(n, n))
    upper = sp.csr_matrix(np.where(mask, signs, 0.0))
    return (upper + upper.T).tocsr()


@dataclass
class RunPlan:
    """Everything the pipeline needs, resolved once from the merged config."""
    solver_class: type
    params: dict[str, Any]
    dataset: str
    run_name: str
    seeds: list[int]
    datasets_path: str
    instances: list[str] | None
    save_spins: bool
    warmup: bool

    @classmethod
    def from_config(cls, config: dict[str, Any]) -> "RunPlan":
        solver_class = get_solver_class(config['solver'])
        offered = {k: v for k, v in config.items() if k not in PIPELINE_KEYS}
        first = int(config.get('random_seed', 42))
        count = int(config.get('num_seeds', 1))
        chosen = config.get('instance')
        if isinstance(chosen, str):
            chosen = [chosen]
        return cls(solver_class=solver_class, params=constructor_params(solver_class, offered),
                   dataset=config['dataset'], run_name=config.get('run_name', config['solver']),
                   seeds=list(range(first, first + count)), datasets_path=config.get('datasets_path', 'datasets'),
                   instances=list(chosen) if chosen else None, save_spins=bool(config.get('save_spins')),
                   warmup=bool(config.get('warmup')))

    @property
    def out_dir(self) -> Path:
        return Path("results") / self.dataset / self.run_name

    def make_solver(self, instance_name: str, seed: int):
        return self.solver_class(instance_name=instance_name, dataset=self.dataset, run_name=self.run_name,
                                 random_seed=seed, **self.params)


@dataclass
class RunLog:
    """Outcome counts of the (instance, seed) runs."""
    plan: RunPlan
    total: int = 0
    failures: list = field(default_factory=list)

    def ok(self) -> None:
        self.total += 1

    def fail(self, instance: str, seed: int, message: str) -> None:
        self.total += 1
        self.failures.append((instance, seed, message))

    def summary(self) -> dict[str, Any]:
        return {'run_name': self.plan.run_name, 'dataset': self.plan.dataset, 'total_runs': self.total,
                'successful_runs': self.total - len(self.failures), 'failed_runs': len(self.failures),
                'failures': list(self.failures)}


def _files_matching(root: Path, stem: str) -> set:
    return {p for p in root.rglob(f'*{stem}*') if p.is_file()} if root.exists() else set()


def warm_up(plan: RunPlan, solve: Callable | None = None) -> None:
    """One untimed solve of a synthetic instance, so that one-time device set-up is not charged to the
    first real instance; whatever result files it writes are deleted again."""
    stem = Path(WARMUP_INSTANCE).stem
    existing = _files_matching(plan.out_dir, stem)
    try:
        plan.make_solver(WARMUP_INSTANCE, plan.seeds[0]).solve(synthetic_warmup_couplings())
    except Exception as exc:  # a warm-up failure only costs timing accuracy
        print(f"  warm-up skipped after {type(exc).__name__}: {exc} "
              f"(the first instance's time then includes one-time set-up)")
    finally:
        for leftover in sorted(_files_matching(plan.out_dir, stem) - existing):
            leftover.unlink()


def _failure_label(exc: BaseException) -> str:
    return "out of memory" if is_out_of_memory(exc) else type(exc).__name__


def _solve_one(plan: RunPlan, log: RunLog, instance: str, seed: int, J, index) -> None:
    try:
        _, spins, _ = plan.make_solver(instance, seed).solve(J)
        if plan.save_spins:
            save_solution(solution_path(plan.out_dir, instance, seed), spins, J, index, dataset=plan.dataset,
                          instance_name=instance, run_name=plan.run_name, seed=seed)
    except Exception as exc:
        traceback.print_exc()
        message = f"{_failure_label(exc)}: {exc}"
        log.fail(instance, seed, message)
        print(f"  FAILED: {instance} seed={seed}: {message}")
        return
    log.ok()


def run_solver(config: dict[str, Any]) -> dict[str, Any]:
    """Run the configured solver on every instance and seed; returns the run summary."""
    plan = RunPlan.from_config(config)
    loader = DataLoader(plan.dataset, plan.datasets_path)
    instances = sorted(plan.instances if plan.instances else loader.list_instances(), key=instance_size_key)
    log = RunLog(plan)
    if plan.warmup:
        print("Warm-up solve on a small synthetic instance (not recorded)")
        warm_up(plan)
    for instance in instances:
        print(f"Processing instance: {instance}")
        try:
            J, index = loader.load_instance(instance, return_index=True)
        except Exception as exc:
            traceback.print_exc()
            for seed in plan.seeds:
                log.fail(instance, seed, f"load: {type(exc).__name__}: {exc}")
            continue
        for number, seed in enumerate(plan.seeds, start=1):
            print(f"  Seed {number}/{len(plan.seeds)} (seed={seed})")
            _solve_one(plan, log, instance, seed, J, index)
    return log.summary()
