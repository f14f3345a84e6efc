"""``import ncvis`` for code written against the reference.

The reference builds its extension as the module ``ncvis`` (setup.py:133-134) and its tests
``import ncvis`` (recipe/test.py:2). This module exposes the same names, backed by the MI355X
implementation in ``ncvis_amd``.
"""
from ncvis_amd import NCVis, NCVisWrapper, find_ab_params  # noqa: F401

__all__ = ["NCVis", "NCVisWrapper", "find_ab_params"]
