"""Config composition with the reference's Hydra schema, without Hydra.

The reference composes ``conf/config_fp.yaml`` with one option of each group
(``exp_type_fp``, ``physics_fp``) through Hydra 1.3 / OmegaConf 2.3
(src/class/simulate.py:11, conf/config_fp.yaml:1-11).  This module reads the same
files and the same command-line grammar:

    key=value            override an existing key (``exp_type_fp.e_ratio_start=0.05``)
    +key=value           add a key that does not exist yet (``+tag=TLwithDose``)
    ++key=value          add or override
    group=option         pick a config group option (``exp_type_fp=TL12``)

Scalars are typed like OmegaConf's YAML loader, which — unlike plain PyYAML —
reads ``3e-4``, ``1e12`` or ``150e-9`` as floats (YAML 1.1 would keep them as
strings).  ``Config`` gives attribute access and ``.get`` like a DictConfig.
"""
from __future__ import annotations

import copy
import re
from pathlib import Path
from typing import Any, Dict, Iterable, List, Optional, Sequence

import yaml

PKG_CONF = Path(__file__).resolve().parent / "conf"

_FLOAT_RE = re.compile(
    r"""^(?:[-+]?(?:[0-9][0-9_]*)\.[0-9_]*(?:[eE][-+]?[0-9]+)?
        |[-+]?(?:[0-9][0-9_]*)(?:[eE][-+]?[0-9]+)
        |\.[0-9_]+(?:[eE][-+]?[0-9]+)?
        |[-+]?\.(?:inf|Inf|INF)
        |\.(?:nan|NaN|NAN))$""", re.X)


class _Loader(yaml.SafeLoader):
    """SafeLoader whose float resolver also accepts exponent-only forms (1e12)."""


_Loader.yaml_implicit_resolvers = {
    k: [(tag, rx) for tag, rx in v if tag != "tag:yaml.org,2002:float"]
    for k, v in yaml.SafeLoader.yaml_implicit_resolvers.copy().items()
}
_Loader.add_implicit_resolver("tag:yaml.org,2002:float", _FLOAT_RE, list("-+0123456789."))


def load_yaml(text_or_path) -> Any:
    if isinstance(text_or_path, Path):
        text = text_or_path.read_text()
    else:
        text = text_or_path
    return yaml.load(text, Loader=_Loader)


def parse_value(text: str) -> Any:
    """A command-line override value, typed like a YAML scalar / flow sequence."""
    if text == "":
        return ""
    return load_yaml(text)


class Config(dict):
    """Nested mapping with attribute access (the subset of DictConfig the drivers use)."""

    def __init__(self, data: Optional[Dict[str, Any]] = None):
        super().__init__()
        for k, v in (data or {}).items():
            self[k] = v

    def __setitem__(self, key, value):
        super().__setitem__(key, _wrap(value))

    def __getattr__(self, key):
        try:
            return self[key]
        except KeyError:
            raise AttributeError(f"Key '{key}' is not in struct") from None

    def __setattr__(self, key, value):
        self[key] = value

    def __deepcopy__(self, memo):
        return Config({k: copy.deepcopy(v, memo) for k, v in self.items()})

    def copy(self) -> "Config":   # DictConfig.copy in the optimizer is used as a deep copy
        return copy.deepcopy(self)

    def to_dict(self) -> Dict[str, Any]:
        return {k: (v.to_dict() if isinstance(v, Config) else list(v) if isinstance(v, list) else v)
                for k, v in self.items()}


def _wrap(v):
    if isinstance(v, Config):
        return v
    if isinstance(v, dict):
        return Config(v)
    if isinstance(v, (list, tuple)):
        return [_wrap(x) for x in v]
    return v


def _defaults(primary: Dict[str, Any]) -> List:
    d = primary.get("defaults", [])
    return list(d) if d else []


def compose(overrides: Sequence[str] = (), config_dir: Optional[Path] = None,
            config_name: str = "config_fp") -> Config:
    """Compose config_name.yaml and its defaults list, then apply overrides."""
    config_dir = Path(config_dir) if config_dir is not None else PKG_CONF
    primary = load_yaml(config_dir / f"{config_name}.yaml") or {}
    group_choice: Dict[str, str] = {}
    order: List[str] = []
    for item in _defaults(primary):
        if item == "_self_":
            order.append("_self_")
        elif isinstance(item, dict):
            for g, opt in item.items():
                group_choice[g] = opt
                order.append(g)
    plain: List[str] = []
    for ov in overrides:
        key, _, val = ov.lstrip("+").partition("=")
        if not ov.startswith("+") and "." not in key and key in group_choice:
            group_choice[key] = val
        else:
            plain.append(ov)
    cfg = Config()
    for item in order:
        if item == "_self_":
            for k, v in primary.items():
                if k not in ("defaults", "hydra"):
                    cfg[k] = v
        else:
            path = config_dir / item / f"{group_choice[item]}.yaml"
            if not path.exists():
                raise FileNotFoundError(f"Could not find '{item}/{group_choice[item]}' in {config_dir}")
            cfg[item] = load_yaml(path) or {}
    for ov in plain:
        apply_override(cfg, ov)
    return cfg


def apply_override(cfg: Config, ov: str) -> None:
    add = ov.startswith("+")
    force = ov.startswith("++")
    key, eq, val = ov.lstrip("+").partition("=")
    if not eq:
        raise ValueError(f"override '{ov}' is not key=value")
    parts = key.split(".")
    node = cfg
    for p in parts[:-1]:
        if p not in node:
            if not add:
                raise KeyError(f"Could not override '{key}': '{p}' is not in the config")
            node[p] = Config()
        node = node[p]
    last = parts[-1]
    if last in node and add and not force:
        raise KeyError(f"Could not append to config. An item is already at '{key}'")
    if last not in node and not add:
        raise KeyError(f"Could not override '{key}'. To append to your config use +{key}={val}")
    node[last] = parse_value(val)
