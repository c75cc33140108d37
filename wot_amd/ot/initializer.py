"""Per-pair OT configuration (wot/ot/initializer.py:42-183).

A configuration gives epsilon / lambda1 / lambda2 per day pair, either per
time pair (columns t0, t1) or per time point (column t; neighbouring values
are averaged onto the pair between them).
"""
import io
import os

import pandas as pd

_PAIR_FIELDS = ('epsilon', 'lambda1', 'lambda2')


def read_day_pairs(day_pairs):
    """A file path, or inline CSV with ';' line terminators (wot/io/io.py:548-556)."""
    if os.path.isfile(day_pairs):
        return pd.read_csv(day_pairs, engine='python', sep=None)
    return pd.read_csv(io.StringIO(day_pairs), sep=',', lineterminator=';')


def parse_configuration(config):
    """None, a path / inline string, a DataFrame (column 't' or columns
    't0','t1') -> dict {(t0, t1): {epsilon, lambda1, lambda2}} or None."""
    if config is None:
        return None
    if isinstance(config, str):
        return parse_configuration(read_day_pairs(config))
    if isinstance(config, pd.DataFrame):
        if 't' in config.columns:
            return parse_per_timepoint_configuration(config)
        if 't0' in config.columns and 't1' in config.columns:
            return parse_per_timepair_configuration(config)
        raise ValueError("Configuration must have at least a column 't' or two columns 't0' and 't1'")
    if isinstance(config, dict):
        raise ValueError("Not implemented")
    raise ValueError("Unrecognized argument type for configuration. Use DataFrame, dict, str or None")


def parse_per_timepoint_configuration(config):
    """Per-time-point values averaged onto consecutive pairs (initializer.py:84-123)."""
    if isinstance(config, pd.DataFrame):
        if 't' not in config.columns:
            raise ValueError("Invalid per-timepoint configuration : must have column t")
        types = {'t': float, 'epsilon': float, 'lambda1': float, 'lambda2': float}
        config = config.sort_values(by='t').astype({x: types[x] for x in types if x in config.columns})
        fields = [x for x in config.columns if x != 't' and x in types]
        pairs = {}
        for i in range(len(config) - 1):
            a, b = config.iloc[i], config.iloc[i + 1]
            pairs[(a['t'], b['t'])] = {x: (a[x] + b[x]) / 2 for x in fields}
        return pairs
    if isinstance(config, dict):
        raise ValueError("Not implemented")
    raise ValueError("Unrecognized argument type for per-timepoint configuration. Use DataFrame, str or None")


def parse_per_timepair_configuration(config):
    """Rows (t0, t1, params...) or a dict keyed by (t0, t1) (initializer.py:135-183);
    only epsilon / lambda1 / lambda2 are kept per pair."""
    if isinstance(config, pd.DataFrame):
        if 't0' not in config.columns or 't1' not in config.columns:
            raise ValueError("Invalid per-timepair configuration : must have columns t0 and t1")
        result = {}
        for i in range(len(config)):
            h = config.loc[i].to_dict()
            t0, t1 = h.pop('t0'), h.pop('t1')
            result[(t0, t1)] = h
        return parse_per_timepair_configuration(result)
    if isinstance(config, dict):
        try:
            if not all(isinstance(z, (int, float)) for x, y in config for z in [x, y]):
                raise ValueError("Dictionary keys for config must be pairs of floats")
        except (TypeError, ValueError):
            raise ValueError("Dictionnary keys for config must be pairs")
        for key in config:
            if not isinstance(config[key], dict):
                raise ValueError("Dictionnary values for config must be dictionnaries")
            config[key] = {x: config[key][x] for x in _PAIR_FIELDS if x in config[key]}
        return config
    raise ValueError("Unrecognized argument type for config. Use DataFrame or dict")


def parse_parameter_file(path):
    """Two-column file: parameter name, value (initializer.py:126-132)."""
    df = pd.read_csv(path, engine='python', sep=None, header=None)
    return {df.iloc[i, 0]: df.iloc[i, 1] for i in range(len(df))}
