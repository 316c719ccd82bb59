"""Run metadata with the reference's keys (wakis/logger.py:14-97).

Three dictionaries, ``grid``, ``solver`` and ``wakeSolver``, filled by
GridFIT3D, SolverFIT3D and WakeSolver with the keys the reference uses, and
written by ``save_logs`` as ``<results_folder>/wakis.log``: a title, then one
section per dictionary ("## WakeSolver Logs ##", "## Solver Logs ##", "## Grid
Logs ##") holding the dictionary as indented JSON, arrays as lists and values
JSON cannot hold as their ``str``, so that tools reading the reference's log
read this one.
"""

import json
import os

import numpy as np

__all__ = ["Logger"]


def _plain(obj):
    """obj with arrays as lists and anything JSON cannot encode as str."""
    if isinstance(obj, dict):
        return {k: _plain(v) for k, v in obj.items()}
    if isinstance(obj, (list, tuple)):
        return [_plain(v) for v in obj]
    if isinstance(obj, np.ndarray):
        return obj.tolist()
    try:
        json.dumps(obj)
    except (TypeError, ValueError):
        return str(obj)
    return obj


class Logger:
    def __init__(self):
        self.grid = {}
        self.solver = {}
        self.wakeSolver = {}

    def save_logs(self, results_folder=None):
        """Write wakis.log into ``results_folder`` (default: the folder the
        WakeSolver logged), creating the folder if needed."""
        if results_folder is not None:
            self.wakeSolver["results_folder"] = results_folder
        folder = self.wakeSolver["results_folder"]
        os.makedirs(folder, exist_ok=True)
        lines = ["Simulation Parameters\n", "=====================\n\n"]
        for title, data in (("WakeSolver Logs", self.wakeSolver), ("Solver Logs", self.solver),
                            ("Grid Logs", self.grid)):
            lines.append(f"\n## {title} ##\n")
            if not data:
                lines.append("(empty)\n")
                continue
            lines.append(json.dumps(_plain(data), indent=2, ensure_ascii=False) + "\n")
        with open(os.path.join(folder, "wakis.log"), "w", encoding="utf-8") as fh:
            fh.write("".join(lines))

    @staticmethod
    def read_logs(path):
        """The three dictionaries of a wakis.log file, as
        {'wakeSolver': ..., 'solver': ..., 'grid': ...} (None for '(empty)')."""
        with open(path, encoding="utf-8") as fh:
            text = fh.read()
        out = {}
        for key, title in (("wakeSolver", "WakeSolver Logs"), ("solver", "Solver Logs"),
                           ("grid", "Grid Logs")):
            head = f"## {title} ##\n"
            i = text.index(head) + len(head)
            j = text.find("\n## ", i)
            body = text[i:j if j >= 0 else len(text)].strip()
            out[key] = None if body == "(empty)" else json.loads(body)
        return out
