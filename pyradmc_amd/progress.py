"""Progress ticks for long ``run``/``run_dij`` calls (``pyradmc/progress.py:29-96``).

One :class:`ProgressEvent` per tick; a callback that raises is logged and ignored,
so an observer can never break a run. Ticks are serialized under a lock.
"""

from __future__ import annotations

import logging
import threading
import time
from collections.abc import Callable
from dataclasses import dataclass

logger = logging.getLogger(__name__)


@dataclass(frozen=True)
class ProgressEvent:
    histories_done: int
    histories_total: int
    elapsed_s: float
    rate_hz: float


ProgressCallback = Callable[[ProgressEvent], None]


class ProgressEmitter:
    def __init__(self, callback: ProgressCallback | None, histories_total: int) -> None:
        self._callback = callback
        self._total = histories_total
        self._done = 0
        self._lock = threading.Lock()
        self._start = time.monotonic()
        self._last_t = self._start

    def tick(self, histories: int) -> None:
        with self._lock:
            self._done += histories
            now = time.monotonic()
            dt = now - self._last_t
            rate = histories / dt if dt > 0 else 0.0
            self._last_t = now
            event = ProgressEvent(self._done, self._total, now - self._start, rate)
            logger.debug("progress: %d/%d histories (%.1f Hz)", event.histories_done,
                         event.histories_total, event.rate_hz)
            if self._callback is None:
                return
            try:
                self._callback(event)
            except Exception:  # an observer must never break a run
                logger.exception("Progress callback raised; ignoring.")
