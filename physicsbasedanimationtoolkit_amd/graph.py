"""Vertex colouring strategies (bindings/pypbat/graph/Color.cpp:15-27; graph/Enums.h)."""
from __future__ import annotations

import enum


class GreedyColorOrderingStrategy(enum.IntEnum):
    Natural = 0
    SmallestDegree = 1
    LargestDegree = 2


class GreedyColorSelectionStrategy(enum.IntEnum):
    LeastUsed = 0
    FirstAvailable = 1


Natural = GreedyColorOrderingStrategy.Natural
SmallestDegree = GreedyColorOrderingStrategy.SmallestDegree
LargestDegree = GreedyColorOrderingStrategy.LargestDegree
LeastUsed = GreedyColorSelectionStrategy.LeastUsed
FirstAvailable = GreedyColorSelectionStrategy.FirstAvailable
