"""Command-line arguments and the YAML config (utils/utils.py:1-81).

Same flags, same precedence (top-level values < the run's params < command-line flags), and the
same YAML loader that reads numbers such as 1e-6 as floats.
"""
import argparse
import re
from pathlib import Path
from typing import Any, Dict

import yaml

CONFIG_DIR = Path(__file__).resolve().parent.parent / "configs"


class _ConfigLoader(yaml.SafeLoader):
    """SafeLoader that also reads 1e-6 (no dot) as a float; quoted values stay strings."""


_ConfigLoader.add_implicit_resolver(
    'tag:yaml.org,2002:float',
    re.compile(r'^[-+]?(?:[0-9][0-9_]*)(?:\.[0-9_]*)?[eE][-+]?[0-9]+$'),
    list('-+0123456789'),
)


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="SpecAnn on MI355X: run a spectral solver on the instances of a dataset")
    p.add_argument('--solver', required=True, help='Run name from the config file (e.g. SpecAnn, SpecAdj)')
    p.add_argument('--dataset', required=True, help='Dataset folder in datasets/')
    p.add_argument('--instance', nargs='+', default=None, help='Instance file name(s) (default: all)')
    p.add_argument('--config', default='default', help='Config file name in configs/ (default: default)')
    p.add_argument('--num-seeds', dest='num_seeds', type=int, default=None, help='Seeds per instance')
    p.add_argument('--seed', dest='seed', type=int, default=None, help='First seed (alone: this seed only)')
    p.add_argument('--save-spins', dest='save_spins', action='store_true', default=None,
                   help='Write the best spin vector of each run to results/<dataset>/<run>/solutions/')
    p.add_argument('--no-save-spins', dest='save_spins', action='store_false', help='Do not write spin vectors')
    p.add_argument('--datasets-path', dest='datasets_path', default='datasets', help='Folder holding the datasets')
    return p


def parse_arguments(argv=None) -> argparse.Namespace:
    return build_parser().parse_args(argv)


def load_config(config_name: str) -> Dict[str, Any]:
    """Load configs/<config_name>.yml (the working directory's configs/ first, then the package's)."""
    for base in (Path("configs"), CONFIG_DIR):
        path = base / f"{config_name}.yml"
        if path.exists():
            with open(path) as f:
                return yaml.load(f, Loader=_ConfigLoader)
    raise FileNotFoundError(f"Config file not found: configs/{config_name}.yml")


def merge_args_with_config(args: argparse.Namespace, config: Dict[str, Any]) -> Dict[str, Any]:
    """Top-level config values, then the params of the run named by --solver, then the flags."""
    if args.solver not in config or not isinstance(config[args.solver], dict):
        runs = sorted(k for k, v in config.items() if isinstance(v, dict))
        raise ValueError(f"Unknown run name {args.solver!r}; the config file has: {', '.join(runs)}")
    run = config[args.solver]
    merged = {k: v for k, v in config.items() if not isinstance(v, dict)}
    merged.update(run.get('params') or {})
    merged['solver'] = run['solver_name']
    merged['run_name'] = args.solver
    merged['dataset'] = args.dataset
    if args.instance:
        merged['instance'] = list(args.instance)
    if getattr(args, 'datasets_path', None):
        merged['datasets_path'] = args.datasets_path
    if getattr(args, 'save_spins', None) is not None:
        merged['save_spins'] = args.save_spins
    if getattr(args, 'num_seeds', None) is not None:
        if args.num_seeds < 1:
            raise ValueError(f"--num-seeds must be >= 1, got {args.num_seeds}")
        merged['num_seeds'] = args.num_seeds
    if getattr(args, 'seed', None) is not None:
        if args.seed < 0:
            raise ValueError(f"--seed must be >= 0, got {args.seed}")
        merged['random_seed'] = args.seed
        if getattr(args, 'num_seeds', None) is None:
            merged['num_seeds'] = 1
    return merged
