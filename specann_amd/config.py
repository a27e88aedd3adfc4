"""Command-line flags and the YAML run config behind main.py.

Interface kept from the reference (utils/utils.py:25-81): the flag names, the config file
configs/<name>.yml, and the precedence top-level values < the run's ``params`` < command-line flags;
``--seed`` alone means that seed only. YAML numbers written as ``1e-6`` (no dot) read as floats.
The merge itself is a table of flag overrides applied over two config layers.
"""
from __future__ import annotations

import argparse
import re
from dataclasses import dataclass
from pathlib import Path
from typing import Any, Callable

import yaml

CONFIG_DIR = Path(__file__).resolve().parent.parent / "configs"


class _ConfigLoader(yaml.SafeLoader):
    """yaml.SafeLoader plus exponent-only floats (YAML 1.1 wants ``1.0e-6``; configs write ``1e-6``)."""


_ConfigLoader.add_implicit_resolver(
    'tag:yaml.org,2002:float',
    re.compile(r'^[-+]?(?:[0-9][0-9_]*)(?:\.[0-9_]*)?[eE][-+]?[0-9]+$'),
    list('-+0123456789'),
)

# (flag, dest, argparse keywords)
_FLAGS = (
    ('--solver', 'solver', dict(required=True, help='run name in the config file: SpecAnn, SpecAnn-64, SpecAdj, ...')),
    ('--dataset', 'dataset', dict(required=True, help='dataset folder under the datasets path')),
    ('--instance', 'instance', dict(nargs='+', default=None, help='instance files to run (default: the whole dataset)')),
    ('--config', 'config', dict(default='default', help='configs/<name>.yml to read (default: default)')),
    ('--num-seeds', 'num_seeds', dict(type=int, default=None, help='seeds per instance, counting up from the first')),
    ('--seed', 'seed', dict(type=int, default=None, help='first seed; given alone, only this seed runs')),
    ('--datasets-path', 'datasets_path', dict(default='datasets', help='folder that holds the dataset folders')),
)


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="Spectral Ising solvers (SpecAnn, SpecAdj, SpecSignedLap) on one MI355X")
    for flag, dest, kw in _FLAGS:
        p.add_argument(flag, dest=dest, **kw)
    spins = p.add_mutually_exclusive_group()
    spins.add_argument('--save-spins', dest='save_spins', action='store_true', default=None,
                       help='write each run\'s best spins to results/<dataset>/<run>/solutions/')
    spins.add_argument('--no-save-spins', dest='save_spins', action='store_false',
                       help='do not write spin files, whatever the config says')
    return p


def parse_arguments(argv=None) -> argparse.Namespace:
    return build_parser().parse_args(argv)


def load_config(config_name: str) -> dict[str, Any]:
    """configs/<config_name>.yml from the working directory, else the one shipped with the package."""
    for folder in (Path("configs"), CONFIG_DIR):
        candidate = folder / f"{config_name}.yml"
        if candidate.is_file():
            return yaml.load(candidate.read_text(), Loader=_ConfigLoader)
    raise FileNotFoundError(f"no config configs/{config_name}.yml (looked in ./configs and {CONFIG_DIR})")


def _at_least(bound: int, flag: str) -> Callable[[int], int]:
    def check(value: int) -> int:
        if value < bound:
            raise ValueError(f"{flag} must be at least {bound} (got {value})")
        return value
    return check


@dataclass(frozen=True)
class _Override:
    """A command-line value that replaces a config key when the flag was given."""
    dest: str
    key: str
    check: Callable[[Any], Any] = lambda v: v


_OVERRIDES = (
    _Override('datasets_path', 'datasets_path'),
    _Override('save_spins', 'save_spins'),
    _Override('num_seeds', 'num_seeds', _at_least(1, '--num-seeds')),
    _Override('seed', 'random_seed', _at_least(0, '--seed')),
)


def _run_entries(config: dict[str, Any]) -> dict[str, dict]:
    return {name: body for name, body in config.items() if isinstance(body, dict)}


def merge_args_with_config(args: argparse.Namespace, config: dict[str, Any]) -> dict[str, Any]:
    """Resolve one run's settings: the config's top-level scalars, overlaid by the named run's params,
    then its solver / run / dataset / instance identity, then the flags that were given."""
    runs = _run_entries(config)
    run = runs.get(args.solver)
    if run is None:
        raise ValueError(f"run {args.solver!r} is not defined in the config; defined runs: {sorted(runs)}")
    settings = {k: v for k, v in config.items() if k not in runs}
    settings.update(run.get('params') or {})
    settings.update(solver=run['solver_name'], run_name=args.solver, dataset=args.dataset)
    if args.instance:
        settings['instance'] = list(args.instance)
    for o in _OVERRIDES:
        value = getattr(args, o.dest, None)
        if value is not None:
            settings[o.key] = o.check(value)
    if getattr(args, 'seed', None) is not None and getattr(args, 'num_seeds', None) is None:
        settings['num_seeds'] = 1
    return settings
