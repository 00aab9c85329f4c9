"""Eager PyTorch restatement of the reference POM step (test infrastructure only)."""
from .explicit import State, run_time_loop, step  # noqa: F401
from .state import load_state, SCALAR_NAMES  # noqa: F401
