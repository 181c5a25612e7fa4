"""`.in` input files: the reference's key = value grammar and defaults.

Grammar (src/utils/read_config.cpp:18-46): each line is trimmed; empty lines and lines whose
first character is '#' are skipped; the first '=' becomes a space; the line is then read as
``key value`` with ``value`` parsed as a floating-point number (anything after it, e.g. a
trailing ``# comment``, is ignored).  A line that does not parse is reported on stderr and
skipped.  ``get(key, default)`` returns ``default`` (0.0 when not given) for a missing key
(src/utils/read_config.cpp:48-50).
"""
from __future__ import annotations

import re
import sys
from dataclasses import dataclass, field

# istringstream >> double accepts an optional sign, digits, '.', exponent, inf/nan.
_NUM = re.compile(r"[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan)", re.I)


@dataclass
class Config:
    data: dict = field(default_factory=dict)

    @classmethod
    def from_text(cls, text: str) -> "Config":
        cfg = cls()
        for raw in text.splitlines():
            line = raw.strip()
            if not line or line[0] == "#":
                continue
            line = line.replace("=", " ", 1)
            parts = line.split(None, 1)
            m = _NUM.match(parts[1].lstrip()) if len(parts) == 2 else None
            if m is None:
                print(f"Warning: Invalid line in config file: {line}", file=sys.stderr)
                continue
            cfg.data[parts[0]] = float(m.group(0))
        return cfg

    @classmethod
    def load(cls, path: str) -> "Config":
        try:
            with open(path) as f:
                text = f.read()
        except OSError as e:  # src/utils/read_config.cpp:21-23 throws invalid_argument
            raise ValueError("Could not open the input file") from e
        return cls.from_text(text)

    def get(self, key: str, default: float = 0.0) -> float:
        return self.data.get(key, default)

    def set(self, key: str, value: float) -> None:
        self.data[key] = float(value)

    def to_text(self) -> str:
        return "".join(f"{k} = {v!r}\n" for k, v in self.data.items())


def write_in(path: str, values: dict) -> None:
    """Write an `.in` file holding ``values`` (repr keeps every double exact)."""
    with open(path, "w") as f:
        for k, v in values.items():
            f.write(f"{k:<16}= {float(v)!r}\n")
