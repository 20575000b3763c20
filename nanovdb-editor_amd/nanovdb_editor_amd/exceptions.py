"""Exception types of the Python package (names follow pymodule/nanovdb_editor/exceptions.py)."""


class InvalidArgumentError(ValueError):
    """A call received arguments it cannot accept."""


class PipelineError(RuntimeError):
    """A library call failed."""
