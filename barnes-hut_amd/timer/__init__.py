from .timer import Timer  # noqa: F401
