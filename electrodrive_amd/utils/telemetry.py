"""CSV telemetry of solver traces (mirror of electrodrive/utils/telemetry.py:10-81).

``TelemetryWriter(path, headers)``: rows are dicts keyed by ``headers``; the
header line is written when the first row arrives, non-finite floats are
written as their string form so the CSV stays parseable, the number of data
rows can be capped with EDE_SOLVER_TRACE_MAX_ROWS, and writing never raises
to the solver (a failed write drops the row).
"""
import csv
import math
import os
import threading
from pathlib import Path
from typing import Any, Dict, List


class TelemetryWriter:
    def __init__(self, path, headers: List[str]):
        self.path = Path(path)
        self.headers = list(headers)
        self._lock = threading.Lock()
        self._ready = False
        self._rows = 0
        cap = os.environ.get("EDE_SOLVER_TRACE_MAX_ROWS")
        try:
            self._max_rows = int(cap) if cap else None
        except ValueError:
            self._max_rows = None

    def _start(self) -> None:
        self._ready = True
        try:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            with self.path.open("w", newline="", encoding="utf-8") as f:
                csv.writer(f).writerow(self.headers)
        except OSError:
            pass

    def append_row(self, row: Dict[str, Any]) -> None:
        with self._lock:
            if not self._ready:
                self._start()
            if self._max_rows is not None and self._rows >= self._max_rows:
                return
            vals = []
            for h in self.headers:
                v = row.get(h)
                if isinstance(v, float) and not math.isfinite(v):
                    v = str(v)
                vals.append(v)
            try:
                with self.path.open("a", newline="", encoding="utf-8") as f:
                    csv.writer(f).writerow(vals)
                self._rows += 1
            except OSError:
                return


__all__ = ["TelemetryWriter"]
