from os import sched_getaffinity


def default_num_threads() -> int:
    """Usable cores minus one, at least one (reference: utils/misc.py:4-18)."""
    n = len(sched_getaffinity(0))
    return n - 1 if n > 1 else 1
