"""``JsonlLogger`` (electrodrive/utils/logging.py:136-240): one JSON object per
line in <out_dir>/events.jsonl with ts, level, msg and the caller's fields;
never raises into the caller.  ``out_dir=None`` logs nowhere."""
from __future__ import annotations

import datetime as _dt
import json
import math
import threading
from pathlib import Path
from typing import Any, Optional


def _clean(v: Any) -> Any:
    if isinstance(v, float) and not math.isfinite(v):
        return str(v)
    if isinstance(v, (str, int, float, bool)) or v is None:
        return v
    if isinstance(v, dict):
        return {str(k): _clean(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_clean(x) for x in v]
    try:
        import torch
        if isinstance(v, torch.Tensor):
            if v.numel() == 1:
                return _clean(v.item())
            return {"tensor_shape": list(v.shape), "dtype": str(v.dtype)}
    except ImportError:
        pass
    return str(v)


class JsonlLogger:
    def __init__(self, out_dir: Optional[Path | str] = None):
        self.dir = Path(out_dir) if out_dir is not None else None
        self.path = None
        self._lock = threading.Lock()
        self._stream = None
        if self.dir is not None:
            try:
                self.dir.mkdir(parents=True, exist_ok=True)
                self.path = self.dir / "events.jsonl"
                self._stream = self.path.open("a", encoding="utf-8")
            except OSError:
                self._stream = None

    def __enter__(self) -> "JsonlLogger":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def close(self) -> None:
        with self._lock:
            if self._stream is not None:
                try:
                    self._stream.close()
                except OSError:
                    pass
                self._stream = None

    def _emit(self, level: str, msg: str, **fields: Any) -> None:
        if self._stream is None:
            return
        rec = {"ts": _dt.datetime.now(_dt.timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"),
               "level": level, "msg": msg}
        rec.update({str(k): _clean(v) for k, v in fields.items()})
        try:
            line = json.dumps(rec, ensure_ascii=False)
        except (TypeError, ValueError):
            line = json.dumps({k: str(v) for k, v in rec.items()})
        with self._lock:
            try:
                self._stream.write(line + "\n")
                self._stream.flush()
            except OSError:
                pass

    def info(self, msg: str, **fields: Any) -> None:
        self._emit("INFO", msg, **fields)

    def debug(self, msg: str, **fields: Any) -> None:
        self._emit("DEBUG", msg, **fields)

    def warning(self, msg: str, **fields: Any) -> None:
        self._emit("WARN", msg, **fields)

    def error(self, msg: str, **fields: Any) -> None:
        if fields.pop("exc_info", False):
            import traceback
            fields["trace"] = traceback.format_exc()
        self._emit("ERROR", msg, **fields)
