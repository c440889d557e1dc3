"""Per-action wall-clock timers with runko's naming rules
(runko/runko_timer.py:16-137): a repeated name becomes name-1, name-2, ...;
stop() closes the first still-open timer of that family."""

import time

import numpy as np


class Timer:
    def __init__(self):
        self._t = {}

    def _free_name(self, name):
        if name not in self._t:
            return name
        n = 1
        while f"{name}-{n}" in self._t:
            n += 1
        return f"{name}-{n}"

    def start(self, name):
        self._t[self._free_name(name)] = [time.time(), None]

    def stop(self, name):
        now = time.time()
        if name in self._t and self._t[name][1] is None:
            self._t[name][1] = now
            return
        n = 1
        while True:
            cand = f"{name}-{n}"
            if cand not in self._t:
                raise RuntimeError("Trying to stop time which has not been started.")
            if self._t[cand][1] is None:
                self._t[cand][1] = now
                return
            n += 1

    def get_elapsed_times(self):
        return {k: v[1] - v[0] for k, v in self._t.items()}


class TimerStatistics:
    def __init__(self, data, measured_laps):
        self.data = np.asarray(data, dtype=np.float64)
        self.measured_laps = measured_laps

    total = property(lambda self: float(np.sum(self.data)))
    average = property(lambda self: float(np.mean(self.data)))
    minimum = property(lambda self: float(np.min(self.data)))
    maximum = property(lambda self: float(np.max(self.data)))
    std_dev = property(lambda self: float(np.std(self.data)))
    count = property(lambda self: len(self.data))


def timer_statistics(timers):
    acc = {}
    for t in timers:
        for k, v in t.get_elapsed_times().items():
            acc.setdefault(k, []).append(v)
    return {k: TimerStatistics(v, len(timers)) for k, v in acc.items()}
