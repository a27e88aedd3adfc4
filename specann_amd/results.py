"""Results CSVs with the reference's columns and upsert rules.

Per-run rows: results/<dataset>/<run>/<instance>.csv, replaced on (instance_name, dataset, seed,
solver_name) (models/spectral_annealing.py:722-778, spec_adj.py:190-238). Per-alpha rows:
results/<dataset>/<run>/alphas/<instance>_alphas.csv; a rerun replaces the whole previous sweep of
that (instance, dataset, seed, run) and rows are sorted by (seed, alpha) (:780-814).
"""
from __future__ import annotations

from pathlib import Path

import pandas as pd


def _base(instance_name: str) -> str:
    return instance_name.replace('.txt', '')


def store_general(dataset: str, run_name: str, instance_name: str, row: dict) -> Path:
    results_dir = Path(f"results/{dataset}/{run_name}")
    results_dir.mkdir(parents=True, exist_ok=True)
    csv_file = results_dir / f"{_base(instance_name)}.csv"
    df = pd.read_csv(csv_file, float_precision='round_trip') if csv_file.exists() else pd.DataFrame()
    if not df.empty:
        mask = ((df['instance_name'] == row['instance_name']) & (df['dataset'] == row['dataset'])
                & (df['seed'] == row['seed']) & (df['solver_name'] == row['solver_name']))
        if mask.any():
            for key in row:
                if key not in df.columns:
                    df[key] = None
            df.loc[mask, list(row.keys())] = list(row.values())
        else:
            df = pd.concat([df, pd.DataFrame([row])], ignore_index=True)
    else:
        df = pd.DataFrame([row])
    df.to_csv(csv_file, index=False, encoding='utf-8')
    return csv_file


def store_alphas(dataset: str, run_name: str, instance_name: str, seed, rows: list) -> Path | None:
    if not rows:
        return None
    out_dir = Path(f"results/{dataset}/{run_name}/alphas")
    out_dir.mkdir(parents=True, exist_ok=True)
    csv_file = out_dir / f"{_base(instance_name)}_alphas.csv"
    df_new = pd.DataFrame(rows)
    if csv_file.exists():
        old = pd.read_csv(csv_file, float_precision='round_trip')
        same = ((old['instance_name'] == instance_name) & (old['dataset'] == dataset)
                & (old['seed'].isna() if seed is None else old['seed'] == seed) & (old['solver_name'] == run_name))
        df = pd.concat([old[~same], df_new], ignore_index=True)
    else:
        df = df_new
    df = df.sort_values(['seed', 'alpha']).reset_index(drop=True)
    df.to_csv(csv_file, index=False, encoding='utf-8')
    return csv_file
