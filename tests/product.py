"""Loads the product binding (low-witch-finder_amd/whf.py) for the tests."""
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

import __graft_entry__  # noqa: E402

whf = __graft_entry__.load_whf()
