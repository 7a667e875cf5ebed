"""Run log of sketch-and-precondition solvers (drop-in for
parla/comps/determiter/logging.py:4-94)."""
import numpy as np


class SketchAndPrecondLog:
    """Phase times and the per-iteration error log (errors[0] is the error at x = 0)."""

    def __init__(self):
        self.time_sketch = 0.0
        self.time_factor = 0.0
        self.time_presolve = 0.0
        self.time_convert = 0.0
        self._time_setup = 0.0
        self.time_iterate = 0.0
        self.times = None
        self.errors = None
        self.error_desc = """Fill in."""

    @property
    def time_setup(self):
        self._time_setup = self.time_sketch + self.time_factor + self.time_convert
        return self._time_setup

    def wrap_up(self, iter_errors, init_error):
        """logging.py:72-94: times amortised linearly over the iterations."""
        iter_errors = np.atleast_1d(np.asarray(iter_errors, dtype=np.float64))
        setup = self.time_setup
        ramp = np.linspace(0, self.time_iterate, iter_errors.size, endpoint=True)
        self.times = np.concatenate(([setup], setup + self.time_presolve + ramp))
        self.errors = np.concatenate(([init_error], iter_errors))
