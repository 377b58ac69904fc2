"""Physical constants (electrodrive/utils/config.py:34-35).

Test infrastructure only (see oracle/__init__.py).
"""
import math

EPS_0 = 8.854187817e-12
K_E = 1.0 / (4.0 * math.pi * EPS_0)
