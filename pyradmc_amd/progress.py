"""Progress reports for long ``run`` / ``run_dij`` calls.

The event type and the callback contract follow the reference's
``pyradmc/progress.py:29-96``: a callback receives one :class:`ProgressEvent` per
tick (a block of batches in ``run``, a beamlet group in ``run_dij``), never
concurrently with itself and with ``histories_done`` increasing, and an observer
that raises cannot stop the transport. Every tick is also logged at DEBUG level on
this module's logger.
"""

from __future__ import annotations

import logging
import threading
import time
from collections.abc import Callable
from dataclasses import dataclass

_log = logging.getLogger(__name__)

__all__ = ["ProgressCallback", "ProgressEmitter", "ProgressEvent"]


@dataclass(frozen=True)
class ProgressEvent:
    """State of a transport at one tick.

    ``rate_hz`` is the rate over the interval since the previous tick (histories per
    second), not the run average; ``elapsed_s`` counts from the emitter's creation.
    """

    histories_done: int
    histories_total: int
    elapsed_s: float
    rate_hz: float


ProgressCallback = Callable[[ProgressEvent], None]


class ProgressEmitter:
    """Turns completed-history counts into :class:`ProgressEvent` notifications.

    ``tick`` may be called from several host threads; one lock covers both the
    bookkeeping and the notification, so events reach the observer one at a time and
    in the order their counts were added.
    """

    __slots__ = ("_observer", "_expected", "_count", "_guard", "_t0", "_t_prev")

    def __init__(self, callback: ProgressCallback | None, histories_total: int) -> None:
        self._observer = callback
        self._expected = int(histories_total)
        self._count = 0
        self._guard = threading.Lock()
        self._t0 = self._t_prev = time.perf_counter()

    def tick(self, histories: int) -> None:
        """Add ``histories`` completed histories and notify the observer."""
        with self._guard:
            self._notify(self._advance(int(histories)))

    def _advance(self, n: int) -> ProgressEvent:
        t = time.perf_counter()
        span, self._t_prev = t - self._t_prev, t
        self._count += n
        return ProgressEvent(self._count, self._expected, t - self._t0, n / span if span > 0.0 else 0.0)

    def _notify(self, event: ProgressEvent) -> None:
        if _log.isEnabledFor(logging.DEBUG):
            _log.debug("%d of %d histories transported, %.4g per second",
                       event.histories_done, event.histories_total, event.rate_hz)
        if self._observer is None:
            return
        try:
            self._observer(event)
        except Exception as exc:  # the transport goes on whatever the observer does
            _log.warning("progress callback failed and is ignored: %r", exc, exc_info=True)
