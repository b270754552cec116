"""Sizes of the vertex-clustered rules (src/3d/quadrature.jl:70-144)."""
import math


def rule_size(rule: int, n: int, m: int) -> int:
    """Points of rule 1..6: n*m for corner rules, 2*n*m' (m' = m rounded up to even)
    for the mid-edge split rules."""
    if rule <= 3:
        return n * m
    return 2 * n * int(math.ceil(m / 2.0)) * 2
