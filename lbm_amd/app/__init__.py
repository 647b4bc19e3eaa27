"""Apps: per-case parameters, inlet schedules, boundary-condition order,
observables and outputs (restating lbm/src/app/*.py), plus the benchmark
configurations of this package (``configs``)."""
