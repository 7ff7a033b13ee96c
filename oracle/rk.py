"""Explicit Runge-Kutta restatements (FP64, test infrastructure).

``TSIT5``: the 5(4) pair of Ch. Tsitouras, "Runge-Kutta pairs of order 5(4)
satisfying only the first column simplifying assumption", Computers &
Mathematics with Applications 62 (2011) 770-775 -- the default solver of
``SomaxModel.integrate`` (``somax/_src/core/model.py:75``, ``diffrax.Tsit5``,
a third-party dependency, diffrax >= 0.6.0, absent here).  Used with
``ConstantStepSize`` the embedded error estimate plays no part, so only the
stage matrix ``a``, the nodes ``c`` and the weights ``b`` (= the last row of
``a``: first same as last) are restated.

Pins (tests/test_oracle_pins.py): the 17 order conditions of orders 1-5 hold,
the nodes are the row sums, and the global error on y' = lambda y falls as
dt^5.
"""

from __future__ import annotations

import numpy as np

TSIT5_C = np.array([0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0])
TSIT5_A = np.zeros((7, 7))
TSIT5_A[1, :1] = [0.161]
TSIT5_A[2, :2] = [-0.008480655492356989, 0.335480655492357]
TSIT5_A[3, :3] = [2.897153057105493, -6.359448489975075, 4.3622954328695815]
TSIT5_A[4, :4] = [5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525]
TSIT5_A[5, :5] = [5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401,
                  -0.028269050394068383]
TSIT5_A[6, :6] = [0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081,
                  2.324710524099774]
TSIT5_B = TSIT5_A[6].copy()


def explicit_step(f, y, dt, a=TSIT5_A, b=TSIT5_B):
    """One explicit RK step of ``y' = f(y)`` for a tuple of arrays ``y``:
    ``k_i = f(y + dt sum_j a_ij k_j)``, ``y + dt sum_i b_i k_i``."""
    y = tuple(np.asarray(x) * 1.0 for x in y)  # FP64 (complex kept)
    ks = []
    for i in range(len(b)):
        yi = tuple(x + dt * sum(a[i, j] * ks[j][n] for j in range(i)) for n, x in enumerate(y))
        ks.append(f(yi))
    return tuple(x + dt * sum(b[i] * ks[i][n] for i in range(len(b))) for n, x in enumerate(y))
