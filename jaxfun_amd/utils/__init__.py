from .common import Domain, lambdify, n, ulp

__all__ = ["Domain", "lambdify", "n", "ulp"]
