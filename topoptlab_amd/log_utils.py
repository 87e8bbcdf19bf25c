"""Iteration log of the optimisation loop, in topoptlab's SimpleLogger format.

``SimpleLogger`` mirrors topoptlab/log_utils.py:147-350: the file name gets
``.log`` appended unless it already ends with it, the file is truncated on
creation, every message is appended as one line and echoed to stdout, and
messages below the verbosity threshold are dropped (10 debug ``[DEBUG]``,
15 performance ``[PERF]``, 20 info (no prefix), 30 ``[WARNING]``, 40
``[ERROR]``, 50 ``[CRITICAL]``).  ``EmptyLogger`` has the same methods and
writes nothing (log_utils.py:352-440).

``parse_simple_logfile`` reads such a file back into the dictionary layout of
log_utils.py:449-560: ``{"header": {...}, "history": {key: array}}`` plus
one list of messages per tag (``"perf"``, ``"debug"``, ...).  The header ends
at the first iteration line (``it.: ...``) or tagged line.
"""
from __future__ import annotations

import re
from typing import Any, Dict, List

import numpy as np

# verbosity threshold below which each method is silent, and its line prefix
_LEVELS = {"debug": (10, "[DEBUG] "), "perf": (15, "[PERF] "), "info": (20, ""),
           "warning": (30, "[WARNING] "), "error": (40, "[ERROR] "),
           "critical": (50, "[CRITICAL] ")}

# "key: value" pairs of header and iteration lines; a key may end in "."
# ("it.", "obj."), a value is a decimal number, an integer or one word
_PAIR = re.compile(r"([A-Za-z_]+\.?):\s*([-+]?\d*\.\d+|\d+|[A-Za-z_./-]+)")
_TAG = re.compile(r"^\[([A-Z]+)\]\s*(.*)")
_ELEMENTS = re.compile(r"elements:\s*(\d+)\s*x\s*(\d+)(?:\s*x\s*(\d+))?")


class SimpleLogger:
    def __init__(self, file: str, verbosity: int = 20) -> None:
        self.file = file if file.endswith(".log") else f"{file}.log"
        open(self.file, "w").close()
        self.verbosity = verbosity

    def _write(self, msg: str) -> None:
        if not msg.endswith("\n"):
            msg += "\n"
        with open(self.file, "a") as fh:
            fh.write(msg)
        print(msg, end="")

    def __call__(self, msg: str) -> None:
        """Append ``msg`` whatever the verbosity."""
        self._write(msg)

    def _emit(self, level: str, msg: str) -> None:
        threshold, prefix = _LEVELS[level]
        if self.verbosity <= threshold:
            self._write(prefix + msg)

    def debug(self, msg: str) -> None:
        self._emit("debug", msg)

    def perf(self, msg: str) -> None:
        self._emit("perf", msg)

    def info(self, msg: str) -> None:
        self._emit("info", msg)

    def warning(self, msg: str) -> None:
        self._emit("warning", msg)

    def error(self, msg: str) -> None:
        self._emit("error", msg)

    def critical(self, msg: str) -> None:
        self._emit("critical", msg)


class EmptyLogger:
    def __init__(self, **kwargs: Any) -> None:
        pass

    def __call__(self, msg: str) -> None:
        pass

    debug = perf = info = warning = error = critical = __call__


def _number_or_word(v: str):
    try:
        return float(v)
    except ValueError:
        return v


def parse_simple_logfile(path: str) -> Dict[str, Any]:
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    lines = [ln for ln in lines if ln]
    body = next((i for i, ln in enumerate(lines)
                 if ln.startswith("it.:") or _TAG.match(ln)), len(lines))
    header: Dict[str, Any] = {}
    for ln in lines[:body]:
        if ln.lower().startswith("elements:"):
            m = _ELEMENTS.search(ln)
            if m:
                for name, val in zip(("nelx", "nely", "nelz"), m.groups()):
                    if val is not None:
                        header[name] = int(val)
            continue
        if ln.lower().startswith("filter method:"):
            header["filter method"] = ln.split(":", 1)[1].strip()
            continue
        pairs = _PAIR.findall(ln)
        if not pairs:
            header[f"line_{len(header)}"] = ln
        for k, v in pairs:
            header[k.strip(":")] = _number_or_word(v)
    rows: List[Dict[str, Any]] = []
    tagged: Dict[str, List[str]] = {}
    for ln in lines[body:]:
        if ln.startswith("it.:"):
            rows.append({k.strip(".: "): _number_or_word(v) for k, v in _PAIR.findall(ln)})
        m = _TAG.match(ln)
        if m:
            tagged.setdefault(m.group(1).lower(), []).append(m.group(2))
    history: Dict[str, np.ndarray] = {}
    for key in set().union(*rows) if rows else ():
        vals = [r.get(key, np.nan) for r in rows]
        try:
            history[key] = np.array([float(v) for v in vals])
        except ValueError:
            history[key] = np.array([r.get(key, "") for r in rows])
    return {"header": header, "history": history, **tagged}
