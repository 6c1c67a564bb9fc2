"""Launch-shape autotuner: times candidate configurations of a kernel on the
device and keeps the fastest per key (in memory, and in a JSON file when a
path is given), so later calls with the same key launch the tuned shape
without timing again.

Used for the persistent expression-program bound, whose free parameter is
the number of resident wavefronts: too few leave compute units idle, too
many thrash the L2 with per-box state.
"""
import json
import os

import torch


class Autotuner:
    def __init__(self, path=None):
        self.path = path
        self.cache = {}
        if path and os.path.exists(path):
            try:
                self.cache = json.load(open(path))
            except (OSError, ValueError):
                self.cache = {}

    def _save(self):
        if not self.path:
            return
        tmp = self.path + ".tmp"
        os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
        with open(tmp, "w") as f:
            json.dump(self.cache, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def time_ms(self, run, reps=3, stream=None):
        """Mean milliseconds of run() over reps, after one warm-up call,
        measured with events on `stream` (torch's current stream by
        default)."""
        stream = stream or torch.cuda.current_stream()
        run()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            run()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    def pick(self, key, candidates, run, reps=3):
        """The fastest candidate c for `key`, timing run(c) for each unless
        the key is cached.  Returns (choice, {candidate: ms} or None)."""
        key = str(key)
        if key in self.cache:
            return self.cache[key]["choice"], None
        times = {}
        for c in candidates:
            times[c] = self.time_ms(lambda: run(c), reps=reps)
        best = min(times, key=times.get)
        self.cache[key] = {"choice": best, "ms": {str(c): round(t, 4) for c, t in times.items()}}
        self._save()
        return best, times


_default = None


def default_tuner():
    """Process-wide tuner; persisted under $GEAR_AMD_TUNE_CACHE when set."""
    global _default
    if _default is None:
        _default = Autotuner(os.environ.get("GEAR_AMD_TUNE_CACHE"))
    return _default
