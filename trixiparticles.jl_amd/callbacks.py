"""``SortingCallback`` (src/callbacks/sorting.jl) for the drivers of this package.

The callback keeps the reference's interface and decisions: sort every ``interval``
accepted steps or at the first step after ``dt`` of simulation time has passed since the
last sort, optionally once at the start, and never when the run is finished.  Its
``affect`` sorts the fluid particles by cell through the library
(:meth:`Semidiscretization.sort_particles`).

``solve`` accepts it as ``callback`` (alone or in a list with plain step functions).  The
integrator handed to ``initialize`` / ``condition`` / ``affect`` is an
:class:`Integrator`: ``p`` (the ODE problem's parameters, holding ``semi``), ``u`` (the
pair ``(v_ode, u_ode)``, the reference's ``integrator.u.x``), ``t``, ``stats.naccept`` and
``isfinished``."""
from __future__ import annotations

from dataclasses import dataclass, field
from types import SimpleNamespace

__all__ = ["SortingCallback", "Integrator"]


@dataclass
class Integrator:
    p: dict
    u: tuple
    t: float
    stats: SimpleNamespace = field(default_factory=lambda: SimpleNamespace(naccept=0))
    isfinished: bool = False


class SortingCallback:
    """``SortingCallback(; interval=-1, dt=0.0, initial_sort=true)`` (sorting.jl:12-57).

    ``interval``: sort after every ``interval`` accepted steps; ``dt``: sort at the first
    step after ``dt`` of simulation time has passed since the last sort (no extra stop
    times); ``initial_sort``: sort once when the run starts."""

    def __init__(self, *, interval=-1, dt=0.0, initial_sort=True):
        if dt > 0 and interval > 0:
            raise ValueError("setting both `interval` and `dt` is not supported")
        if dt > 0:
            self.interval = float(dt)
        elif interval <= 0:
            raise ValueError("either `interval` or `dt` must be set to a positive value")
        else:
            self.interval = int(interval)
        self.last_t = 0.0
        self.initial_sort = bool(initial_sort)
        self.sorts = 0  # sorts done by affect (not in the reference; for reports)

    @property
    def by_time(self):
        return isinstance(self.interval, float)

    # `initialize` (sorting.jl:60-75)
    def initialize(self, integrator):
        if self.initial_sort:
            self.affect(integrator)
        else:
            self.last_t = integrator.t

    # `condition` (sorting.jl:78-93)
    def condition(self, u, t, integrator):
        if integrator.isfinished:
            return False
        if self.by_time:
            return (t - self.last_t) > self.interval
        # condition_integrator_interval (callbacks.jl:17-26) without the final-solution clause,
        # which `isfinished` has already excluded
        return self.interval > 0 and integrator.stats.naccept % self.interval == 0

    # `affect!` (sorting.jl:96-116)
    def affect(self, integrator):
        semi = integrator.p["semi"]
        v_ode, u_ode = integrator.u
        semi.sort_particles(v_ode, u_ode)
        self.last_t = integrator.t
        self.sorts += 1
        return integrator

    def __repr__(self):
        key = "dt" if self.by_time else "interval"
        return (f"SortingCallback({key}={self.interval}, "
                f"initial_sort={'true' if self.initial_sort else 'false'})")
