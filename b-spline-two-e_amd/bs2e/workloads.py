"""Named workloads of the benchmark and parity suite.

Knots come from generate_grid(k, m=3, Z=2, h_max, r_max) (grid_tools.f90:6-55);
h_max is the middle of the plateau of values that gives exactly the named
number of B-splines n_b.
"""
from dataclasses import dataclass


@dataclass(frozen=True)
class Workload:
    name: str
    k: int          # B-spline order
    n_b: int        # retained B-splines
    r_max: float
    h_max: float
    k_GL: int
    max_k: int
    lmax: int
    m: int = 3
    Z: int = 2
    description: str = ""


WORKLOADS = {
    "tiny": Workload("tiny", 4, 36, 8.0, 1.0, 6, 2, 1, description="order 4, 36 B-splines, r<=8, GL 6, k<=2"),
    "t0": Workload("t0", 8, 46, 15.0, 1.5, 14, 4, 2,
                   description="the reference test's basis: order 8, r<=15, h_max 1.5, GL 14, k<=4"),
    "c1": Workload("c1", 7, 60, 50.0, 2.3384, 10, 4, 2,
                   description="He-like Z=2, order 7, 60 B-splines, r<=50, GL 10, l<=2, k<=4"),
    "c2": Workload("c2", 8, 200, 200.0, 1.2046, 12, 10, 5,
                   description="order 8, 200 B-splines, r<=200, GL 12, lmax 5, k<=10"),
    "c3": Workload("c3", 8, 300, 300.0, 1.1252, 12, 16, 8,
                   description="order 8, 300 B-splines, r<=300, GL 12, lmax 8, k<=16"),
    "c4": Workload("c4", 9, 400, 400.0, 1.0933, 15, 16, 8,
                   description="order 9, 400 B-splines, r<=400, GL 15, k<=16 (largest serving job)"),
    "c5": Workload("c5", 9, 600, 1000.0, 1.7802, 14, 24, 12,
                   description="order 9, 600 B-splines, r<=1000, GL 14, lmax 12, k<=24"),
}
