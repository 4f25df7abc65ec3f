"""Named phase timers with the reference's interface and printout
(python-bh/timer/timer.py:4-57): ``with Timer.get_handle(name):`` prints
``starting <name>`` / ``stopping <name>  <seconds>``, ``Timer.print()`` reports
the median per name, ``Timer.reset()`` clears.  Phases that launch GPU work
synchronise before they return, so the wall-clock times are device times."""
import statistics
from timeit import default_timer as timer


class TimerHandle:
    def __init__(self, parent, name):
        self.name = name
        self.parent = parent
        self.t0 = None

    def start(self):
        self.t0 = timer()

    def stop(self):
        elapsed = timer() - self.t0
        self.parent._report_time(self.name, elapsed)
        if self.parent.verbose:
            print(f"stopping {self.name}  {elapsed}")
        return elapsed

    def __enter__(self):
        if self.parent.verbose:
            print(f"starting {self.name}")
        self.start()
        return self

    def __exit__(self, exc_type, exc, tb):
        self.stop()


class TimerClass:

    def __init__(self):
        self.times = {}
        self.verbose = True

    def get_handle(self, name):
        return TimerHandle(self, name)

    def _report_time(self, name, elapsed):
        self.times.setdefault(name, []).append(elapsed)

    def medians(self, nwarm=0):
        return {k: statistics.median(v[nwarm:]) for k, v in self.times.items() if v[nwarm:]}

    def print(self, nwarm=0):
        print("*" * 50)
        print(f"{'name':<15}, {'median':<7}")
        for name, med in self.medians(nwarm).items():
            print(f"{name:<15}, {med:.5f}")
        print("*" * 50)

    def reset(self):
        self.times = {}


Timer = TimerClass()
