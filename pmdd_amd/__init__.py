"""pmdd_amd -- exact Pareto-frontier enumeration over decision diagrams on MI355X.

Host surface (instance parsers, DD constructors) in libpmdd_host.so, HIP engine
for gfx950 in libpmdd_engine.so, both behind the C ABI in include/.
"""
from .dd import DecisionDiagram, PROBLEM_KNAPSACK, PROBLEM_SETPACKING, PROBLEM_TSP  # noqa: F401
