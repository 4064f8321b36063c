"""Butcher / Rosenbrock tableaus and the algorithm alias registries.

The coefficient tables are the published ones (Dormand-Prince 1980,
Bogacki-Shampine 1989, Cash-Karp 1990, Fehlberg 1969, Hairer-Norsett-Wanner
DOP853, Tsitouras 2011, Verner 7(6), Kvaerno 2004, Alexander 1977, the MOOSE
L-stable DIRK3, Hairer-Wanner SDIRK4 and Radau IIA, Gauss-Legendre, ROS3P of
Rang-Angermann 2005, RODAS3P and the SciML Rosenbrock23 variant); they carry
the same values and aliases as the reference's tableau modules
(integrators/algorithms/generic_{erk,dirk,firk,rosenbrockw}_tableaus.py) so
that `algorithm="dormand-prince-54"` etc. select the same method.

Derived properties (d = b - b_hat, first-same-as-last, b / b_hat matching a
row of A) follow base_algorithm_step.py:290-560; the device kernels read
them as compile-time constants.
"""

from __future__ import annotations

import math
from typing import Dict, Optional, Sequence, Tuple

import numpy as np


class ButcherTableau:
    """Runge-Kutta style coefficient table.

    Parameters
    ----------
    a
        Stage coupling matrix (rows padded to ``stage_count``).
    b
        Solution weights.
    c
        Stage nodes (fractions of the step).
    order
        Classical order of the method.
    b_hat
        Embedded weights; ``None`` for a method without an error estimate.
    """

    family = "generic"

    def __init__(self, a, b, c, order, b_hat=None, name="custom",
                 dense_prediction_ratio_float32=0.0,
                 dense_prediction_ratio_float64=0.0):
        s = len(b)
        rows = []
        for row in a:
            row = [float(x) for x in row] + [0.0] * (s - len(row))
            rows.append(tuple(row[:s]))
        self.a: Tuple[Tuple[float, ...], ...] = tuple(rows)
        self.b: Tuple[float, ...] = tuple(float(x) for x in b)
        self.c: Tuple[float, ...] = tuple(float(x) for x in c)
        self.order = int(order)
        self.b_hat = None if b_hat is None else tuple(float(x) for x in b_hat)
        if self.b_hat is not None and len(self.b_hat) != s:
            raise ValueError("b_hat must match the number of stages in b")
        self.name = name
        self.dense_prediction_ratio_float32 = float(dense_prediction_ratio_float32)
        self.dense_prediction_ratio_float64 = float(dense_prediction_ratio_float64)

    # -- structure -----------------------------------------------------
    @property
    def stage_count(self) -> int:
        return len(self.b)

    @property
    def d(self) -> Optional[Tuple[float, ...]]:
        """Error weights b - b_hat."""
        if self.b_hat is None:
            return None
        return tuple(bv - bh for bv, bh in zip(self.b, self.b_hat))

    @property
    def has_error_estimate(self) -> bool:
        d = self.d
        return d is not None and any(w != 0.0 for w in d)

    @property
    def first_same_as_last(self) -> bool:
        return bool(
            self.c and self.c[0] == 0.0 and self.c[-1] == 1.0
            and self.a[-1] == self.b and not any(self.a[0])
        )

    def _matching_row(self, weights) -> Optional[int]:
        if weights is None:
            return None
        match = None
        s = self.stage_count
        for r, row in enumerate(self.a):
            if all(abs(row[i] - weights[i]) <= 1e-15 for i in range(s)):
                match = r
        return match

    @property
    def b_matches_a_row(self) -> Optional[int]:
        return self._matching_row(self.b)

    @property
    def b_hat_matches_a_row(self) -> Optional[int]:
        return self._matching_row(self.b_hat)

    @property
    def accumulates_output(self) -> bool:
        return self.b_matches_a_row is None

    @property
    def accumulates_error(self) -> bool:
        return self.b_hat_matches_a_row is None

    @property
    def diagonal(self) -> Tuple[float, ...]:
        return tuple(self.a[i][i] for i in range(self.stage_count))

    @property
    def is_explicit(self) -> bool:
        s = self.stage_count
        return all(self.a[i][j] == 0.0 for i in range(s) for j in range(i, s))

    @property
    def first_stage_is_explicit(self) -> bool:
        return self.a[0][0] == 0.0

    @property
    def stiffly_accurate(self) -> bool:
        return self.a[-1] == self.b

    def dense_prediction_ratio_limit(self, precision) -> float:
        if np.dtype(precision) == np.float32:
            return self.dense_prediction_ratio_float32
        return self.dense_prediction_ratio_float64

    def validate_weight_sums(self) -> None:
        if abs(1.0 - sum(self.b)) > 1e-8:
            raise ValueError("b must sum to one")
        if self.b_hat is not None and abs(1.0 - sum(self.b_hat)) > 1e-8:
            raise ValueError("b_hat must sum to one")

    def __repr__(self) -> str:
        return f"{type(self).__name__}({self.name!r}, stages={self.stage_count}, order={self.order})"


class ERKTableau(ButcherTableau):
    family = "erk"

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.validate_weight_sums()
        if not self.is_explicit:
            raise ValueError("ERK tableau must be strictly lower triangular")


class DIRKTableau(ButcherTableau):
    family = "dirk"

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.validate_weight_sums()


class FIRKTableau(ButcherTableau):
    family = "firk"

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.validate_weight_sums()


class RosenbrockTableau(ButcherTableau):
    """Rosenbrock-W tableau: a, C (Jacobian coupling), gamma, gamma_stages."""

    family = "rosenbrock"

    def __init__(self, a, b, c, order, b_hat=None, C=(), gamma=0.25,
                 gamma_stages=(), name="custom", **kw):
        super().__init__(a, b, c, order, b_hat=b_hat, name=name, **kw)
        s = self.stage_count
        rows = []
        for row in C:
            row = [float(x) for x in row] + [0.0] * (s - len(row))
            rows.append(tuple(row[:s]))
        while len(rows) < s:
            rows.append(tuple([0.0] * s))
        self.C = tuple(rows)
        self.gamma = float(gamma)
        gs = tuple(float(x) for x in gamma_stages)
        self.gamma_stages = gs + (0.0,) * (s - len(gs))


# ---------------------------------------------------------------------------
# Explicit and diagonally implicit tableaus
# ---------------------------------------------------------------------------

HEUN_21 = ERKTableau(
    name="heun_21",
    a=(
        (0.0, 0.0),
        (1.0, 0.0),
    ),
    b=(0.5, 0.5),
    c=(0.0, 1.0),
    order=2,
)

RALSTON_33 = ERKTableau(
    name="ralston_33",
    a=(
        (0.0, 0.0, 0.0),
        (0.5, 0.0, 0.0),
        (0.0, 0.75, 0.0),
    ),
    b=(0.2222222222222222, 0.3333333333333333, 0.4444444444444444),
    c=(0.0, 0.5, 0.75),
    order=3,
)

BOGACKI_SHAMPINE_32 = ERKTableau(
    name="bogacki_shampine_32",
    a=(
        (0.0, 0.0, 0.0, 0.0),
        (0.5, 0.0, 0.0, 0.0),
        (0.0, 0.75, 0.0, 0.0),
        (0.2222222222222222, 0.3333333333333333, 0.4444444444444444, 0.0),
    ),
    b=(0.2222222222222222, 0.3333333333333333, 0.4444444444444444, 0.0),
    b_hat=(0.2916666666666667, 0.25, 0.3333333333333333, 0.125),
    c=(0.0, 0.5, 0.75, 1.0),
    order=3,
)

DORMAND_PRINCE_54 = ERKTableau(
    name="dormand_prince_54",
    a=(
        (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.2, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.075, 0.225, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.9777777777777777, -3.7333333333333334, 3.5555555555555554, 0.0, 0.0, 0.0, 0.0),
        (2.9525986892242035, -11.595793324188385, 9.822892851699436, -0.2908093278463649, 0.0, 0.0, 0.0),
        (2.8462752525252526, -10.757575757575758, 8.906422717743473, 0.2784090909090909, -0.2735313036020583, 0.0, 0.0),
        (0.09114583333333333, 0.0, 0.44923629829290207, 0.6510416666666666, -0.322376179245283, 0.13095238095238096, 0.0),
    ),
    b=(0.09114583333333333, 0.0, 0.44923629829290207, 0.6510416666666666, -0.322376179245283, 0.13095238095238096, 0.0),
    b_hat=(0.08991319444444444, 0.0, 0.4534890685834082, 0.6140625, -0.2715123820754717, 0.08904761904761904, 0.025),
    c=(0.0, 0.2, 0.3, 0.8, 0.8888888888888888, 1.0, 1.0),
    order=5,
)

CLASSICAL_RK4 = ERKTableau(
    name="classical_rk4",
    a=(
        (0.0, 0.0, 0.0, 0.0),
        (0.5, 0.0, 0.0, 0.0),
        (0.0, 0.5, 0.0, 0.0),
        (0.0, 0.0, 1.0, 0.0),
    ),
    b=(0.16666666666666666, 0.3333333333333333, 0.3333333333333333, 0.16666666666666666),
    c=(0.0, 0.5, 0.5, 1.0),
    order=4,
)

CASH_KARP_54 = ERKTableau(
    name="cash_karp_54",
    a=(
        (0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.2, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.075, 0.225, 0.0, 0.0, 0.0, 0.0),
        (0.3, -0.9, 1.2, 0.0, 0.0, 0.0),
        (-0.2037037037037037, 2.5, -2.5925925925925926, 1.2962962962962963, 0.0, 0.0),
        (0.029495804398148147, 0.341796875, 0.041594328703703706, 0.40034541377314814, 0.061767578125, 0.0),
    ),
    b=(0.09788359788359788, 0.0, 0.4025764895330113, 0.21043771043771045, 0.0, 0.2891022021456804),
    b_hat=(0.10217737268518519, 0.0, 0.38390790343915343, 0.24459273726851852, 0.019321986607142856, 0.25),
    c=(0.0, 0.2, 0.3, 0.6, 1.0, 0.875),
    order=5,
)

FEHLBERG_45 = ERKTableau(
    name="fehlberg_45",
    a=(
        (0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.25, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.09375, 0.28125, 0.0, 0.0, 0.0, 0.0),
        (0.8793809740555303, -3.277196176604461, 3.3208921256258535, 0.0, 0.0, 0.0),
        (2.0324074074074074, -8.0, 7.173489278752436, -0.20589668615984405, 0.0, 0.0),
        (-0.2962962962962963, 2.0, -1.3816764132553607, 0.4529727095516569, -0.275, 0.0),
    ),
    b=(0.11851851851851852, 0.0, 0.5189863547758284, 0.5061314903420167, -0.18, 0.03636363636363636),
    b_hat=(0.11574074074074074, 0.0, 0.5489278752436647, 0.5353313840155945, -0.2, 0.0),
    c=(0.0, 0.25, 0.375, 0.9230769230769231, 1.0, 0.5),
    order=5,
)

DORMAND_PRINCE_853 = ERKTableau(
    name="dormand_prince_853",
    a=(
        (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.05260015195876773, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.0197250569845379, 0.0591751709536137, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.02958758547680685, 0.0, 0.08876275643042054, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.2413651341592667, 0.0, -0.8845494793282861, 0.924834003261792, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.037037037037037035, 0.0, 0.0, 0.17082860872947386, 0.12546768756682242, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.037109375, 0.0, 0.0, 0.17025221101954405, 0.06021653898045596, -0.017578125, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.03709200011850479, 0.0, 0.0, 0.17038392571223998, 0.10726203044637328, -0.015319437748624402, 0.008273789163814023, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.6241109587160757, 0.0, 0.0, -3.360892629446941, -0.868219346841726, 27.592099699446706, 20.154067550477894, -43.48988418106996, 0.0, 0.0, 0.0, 0.0),
        (0.47766253643826434, 0.0, 0.0, -2.4881146199716677, -0.590290826836843, 21.230051448181193, 15.279233632882423, -33.28821096898486, -0.020331201708508627, 0.0, 0.0, 0.0),
        (-0.9371424300859873, 0.0, 0.0, 5.186372428844064, 1.0914373489967295, -8.149787010746927, -18.52006565999696, 22.739487099350505, 2.493605552679653, -3.0467644718982196, 0.0, 0.0),
        (2.2733101475165385, 0.0, 0.0, -10.53449546673725, -2.0008720582248625, -17.9589318631188, 27.94888452941996, -2.8589982771350235, -8.87285693353063, 12.360567175794303, 0.6433927460157635, 0.0),
    ),
    b=(0.05429373411656876, 0.0, 0.0, 0.0, 0.0, 4.450312892752409, 1.8915178993145003, -5.801203960010585, 0.3111643669578199, -0.1521609496625161, 0.20136540080403034, 0.04471061572777259),
    b_hat=(0.04117368912237387, 0.0, 0.0, 0.0, 0.0, 5.675469339128613, 2.3872768489717506, -7.465581142465572, 0.6614932157077936, -0.4863400683755336, 0.11944219431891465, 0.06706592359165887),
    c=(0.0, 0.05260015195876773, 0.0789002279381516, 0.1183503419072274, 0.2816496580927726, 0.3333333333333333, 0.25, 0.3076923076923077, 0.6512820512820513, 0.6, 0.8571428571428571, 1.0),
    order=8,
)

TSITOURAS_54 = ERKTableau(
    name="tsitouras_54",
    a=(
        (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.161, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (-0.008480655492356989, 0.335480655492357, 0.0, 0.0, 0.0, 0.0, 0.0),
        (2.8971530571054935, -6.359448489975075, 4.3622954328695815, 0.0, 0.0, 0.0, 0.0),
        (5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0.0, 0.0, 0.0),
        (5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383, 0.0, 0.0),
        (0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774, 0.0),
    ),
    b=(0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774, 0.0),
    b_hat=(0.09824077787029101, 0.010816434459656746, 0.4720087724042376, 1.5237195812770048, -3.872426680888636, 2.7827926300289607, -0.015151515151515152),
    c=(0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0),
    order=5,
)

VERNER_76 = ERKTableau(
    name="verner_76",
    a=(
        (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.005, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (-1.07679012345679, 1.185679012345679, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.04083333333333333, 0.0, 0.1225, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.6389139236255726, 0.0, -2.455672638223657, 2.272258714598084, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (-2.6615773750187572, 0.0, 10.804513886456137, -8.3539146573962, 0.820487594956657, 0.0, 0.0, 0.0, 0.0, 0.0),
        (6.067741434696771, 0.0, -24.711273635911084, 20.427517930788895, -1.9061579788166472, 1.006172249242068, 0.0, 0.0, 0.0, 0.0),
        (12.054670076253203, 0.0, -49.75478495046899, 41.142888638604674, -4.461760149974004, 2.042334822239175, -0.09834843665406108, 0.0, 0.0, 0.0),
        (10.138146522881808, 0.0, -42.6411360317175, 35.76384003992257, -4.3480228403929075, 2.0098622683770357, 0.3487490460338272, -0.27143900510483127, 0.0, 0.0),
        (-45.030072034298676, 0.0, 187.3272437654589, -154.02882369350186, 18.56465306347536, -7.141809679295079, 1.3088085781613785, 0.0, 0.0, 0.0),
    ),
    b=(0.04715561848627222, 0.0, 0.0, 0.25750564298434153, 0.2621665397741262, 0.15216092656738558, 0.49399691700324844, -0.29430311714032503, 0.0813174723249511, 0.0),
    b_hat=(0.044608606606341174, 0.0, 0.0, 0.26716403785713727, 0.2201018300177293, 0.2188431703143157, 0.22898717054112028, 0.0, 0.0, 0.020295184663356284),
    c=(0.0, 0.005, 0.10888888888888888, 0.16333333333333333, 0.4555, 0.6095094489978381, 0.884, 0.925, 1.0, 1.0),
    order=7,
)

IMPLICIT_MIDPOINT = DIRKTableau(
    name="implicit_midpoint",
    a=(
        (0.5,),
    ),
    b=(1.0,),
    c=(0.5,),
    order=2,
    dense_prediction_ratio_float32=1.0,
    dense_prediction_ratio_float64=1.0,
)

TRAPEZOIDAL_DIRK = DIRKTableau(
    name="trapezoidal_dirk",
    a=(
        (0.0, 0.0),
        (0.5, 0.5),
    ),
    b=(0.5, 0.5),
    c=(0.0, 1.0),
    order=2,
    dense_prediction_ratio_float32=0.39,
    dense_prediction_ratio_float64=1.21,
)

KVAERNO3 = DIRKTableau(
    name="kvaerno3",
    a=(
        (0.0, 0.0, 0.0, 0.0),
        (0.4358665215, 0.4358665215, 0.0, 0.0),
        (0.490563388419108, 0.073570090080892, 0.4358665215, 0.0),
        (0.308809969973036, 1.490563388254106, -1.235239879727145, 0.4358665215),
    ),
    b=(0.308809969973036, 1.490563388254106, -1.235239879727145, 0.4358665215),
    b_hat=(0.490563388419108, 0.073570090080892, 0.4358665215, 0.0),
    c=(0.0, 0.871733043, 1.0, 1.0),
    order=3,
    dense_prediction_ratio_float32=0.85,
    dense_prediction_ratio_float64=1.28,
)

KVAERNO5 = DIRKTableau(
    name="kvaerno5",
    a=(
        (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.26, 0.26, 0.0, 0.0, 0.0, 0.0, 0.0),
        (0.13, 0.8403332099679081, 0.26, 0.0, 0.0, 0.0, 0.0),
        (0.22371961478320504, 0.476755323197997, -0.06470895363112615, 0.26, 0.0, 0.0, 0.0),
        (0.16648564323248322, 0.1045001884159172, 0.03631482272098715, -0.13090704451073998, 0.26, 0.0, 0.0),
        (0.13855640231268224, 0.0, -0.04245337201752043, 0.02446657898003141, 0.6194303907248068, 0.26, 0.0),
        (0.13659751177640292, 0.0, -0.05496908796538376, -0.04118626728321046, 0.629933048990164, 0.06962479448202728, 0.26),
    ),
    b=(0.13659751177640292, 0.0, -0.05496908796538376, -0.04118626728321046, 0.629933048990164, 0.06962479448202728, 0.26),
    b_hat=(0.13855640231268224, 0.0, -0.04245337201752043, 0.02446657898003141, 0.6194303907248068, 0.26, 0.0),
    c=(0.0, 0.52, 1.230333209967908, 0.895765984350076, 0.436393609858648, 1.0, 1.0),
    order=5,
)

SDIRK_2_2 = DIRKTableau(
    name="sdirk_2_2",
    a=(
        (0.2928932188134524, 0.0),
        (0.7071067811865476, 0.2928932188134524),
    ),
    b=(0.7071067811865476, 0.2928932188134524),
    c=(0.2928932188134524, 1.0),
    order=2,
    dense_prediction_ratio_float32=1.07,
    dense_prediction_ratio_float64=1.21,
)

L_STABLE_DIRK3 = DIRKTableau(
    name="l_stable_dirk3",
    a=(
        (0.43586652150845895, 0.0, 0.0),
        (0.2820667392457705, 0.43586652150845895, 0.0),
        (1.2084966491760099, -0.6443631706844688, 0.43586652150845895),
    ),
    b=(1.2084966491760099, -0.6443631706844688, 0.43586652150845895),
    c=(0.43586652150845895, 0.7179332607542295, 1.0),
    order=3,
    dense_prediction_ratio_float32=0.85,
    dense_prediction_ratio_float64=1.07,
)

L_STABLE_SDIRK4 = DIRKTableau(
    name="l_stable_sdirk4",
    a=(
        (0.25, 0.0, 0.0, 0.0, 0.0),
        (0.5, 0.25, 0.0, 0.0, 0.0),
        (0.34, -0.04, 0.25, 0.0, 0.0),
        (0.2727941176470588, -0.05036764705882353, 0.027573529411764705, 0.25, 0.0),
        (1.0416666666666667, -1.0208333333333333, 7.8125, -7.083333333333333, 0.25),
    ),
    b=(1.0416666666666667, -1.0208333333333333, 7.8125, -7.083333333333333, 0.25),
    b_hat=(1.2291666666666667, -0.17708333333333334, 7.03125, -7.083333333333333, 0.0),
    c=(0.25, 0.75, 0.55, 0.5, 1.0),
    order=4,
    dense_prediction_ratio_float32=0.79,
    dense_prediction_ratio_float64=0.79,
)


# ---------------------------------------------------------------------------
# Fully implicit (collocation) tableaus
# ---------------------------------------------------------------------------

_SQRT3 = math.sqrt(3.0)
GAUSS_LEGENDRE_2 = FIRKTableau(
    name="firk_gauss_legendre_2",
    a=((0.25, 0.25 - _SQRT3 / 6.0), (0.25 + _SQRT3 / 6.0, 0.25)),
    b=(0.5, 0.5),
    c=(0.5 - _SQRT3 / 6.0, 0.5 + _SQRT3 / 6.0),
    order=4,
    dense_prediction_ratio_float32=8.0,
    dense_prediction_ratio_float64=8.0,
)


def radau_embedded_weights(c: Sequence[float], order: Optional[int] = None) -> Tuple[float, ...]:
    """Weights b* with sum_i b*_i c_i^(k-1) = 1/k for k = 1..order.

    With fewer moment conditions than nodes the minimum-norm solution is
    taken (least squares), the construction behind the Radau IIA(5)
    embedded estimate.
    """
    c = np.asarray(c, dtype=np.float64)
    s = len(c)
    order = s if order is None else order
    if order > s:
        raise ValueError(f"Cannot achieve order {order} with {s} stages")
    m = np.vander(c, N=order, increasing=True).T
    r = np.array([1.0 / k for k in range(1, order + 1)])
    if order == s:
        w = np.linalg.solve(m, r)
    else:
        w = np.linalg.lstsq(m, r, rcond=None)[0]
    return tuple(float(x) for x in w)


_SQRT6 = math.sqrt(6.0)
_RADAU5_C = ((4 - _SQRT6) / 10.0, (4 + _SQRT6) / 10.0, 1.0)
RADAU_IIA_5 = FIRKTableau(
    name="radau_iia_5",
    a=(
        ((88 - 7 * _SQRT6) / 360.0, (296 - 169 * _SQRT6) / 1800.0, (-2 + 3 * _SQRT6) / 225.0),
        ((296 + 169 * _SQRT6) / 1800.0, (88 + 7 * _SQRT6) / 360.0, (-2 - 3 * _SQRT6) / 225.0),
        ((16 - _SQRT6) / 36.0, (16 + _SQRT6) / 36.0, 1.0 / 9.0),
    ),
    b=((16 - _SQRT6) / 36.0, (16 + _SQRT6) / 36.0, 1.0 / 9.0),
    b_hat=radau_embedded_weights(_RADAU5_C, order=2),
    c=_RADAU5_C,
    order=5,
    dense_prediction_ratio_float32=8.0,
    dense_prediction_ratio_float64=8.0,
)

GAUSS_LEGENDRE_4 = FIRKTableau(
    name="firk_gauss_legendre_4",
    a=(
        (0.08696371128436345, -0.026604180084998798, 0.012627462689404742, -0.003555149685795684),
        (0.188118117499868, 0.16303628871563644, -0.027880428602470995, 0.006735500594538164),
        (0.1671919219741886, 0.3539530060337439, 0.16303628871563638, -0.014190694931141147),
        (0.17748257225452202, 0.3134451147418685, 0.3526767575162713, 0.08696371128436325),
    ),
    b=(0.17392742256872662, 0.32607257743127305, 0.3260725774312722, 0.1739274225687269),
    c=(0.06943184420297371, 0.33000947820757187, 0.6699905217924281, 0.9305681557970262),
    order=8,
    dense_prediction_ratio_float32=8.0,
    dense_prediction_ratio_float64=8.0,
)


# ---------------------------------------------------------------------------
# Rosenbrock-W tableaus
# ---------------------------------------------------------------------------

def _ros3p() -> RosenbrockTableau:
    """ROS3P (Rang & Angermann 2005), three stages, order 3."""
    g = 0.5 + math.sqrt(3.0) / 6.0
    ig = 1.0 / g
    b_aux = ig * (2.0 / 3.0 - (1.0 / 6.0) * ig)
    return RosenbrockTableau(
        name="ros3p",
        a=((0.0, 0.0, 0.0), (ig, 0.0, 0.0), (ig, 0.0, 0.0)),
        C=(
            (0.0, 0.0, 0.0),
            (-(ig ** 2), 0.0, 0.0),
            (-ig * (1.0 + ig * (2.0 - 0.5 * ig)), -ig * (2.0 - 0.5 * ig), 0.0),
        ),
        b=(ig * (1.0 + b_aux), b_aux, ig / 3.0),
        b_hat=(2.113248654051871, 1.0, 0.4226497308103742),
        c=(0.0, 1.0, 1.0),
        order=3,
        gamma=g,
        gamma_stages=(g, -0.2113248654051871, 0.5 - 2.0 * g),
    )


def _rodas3p() -> RosenbrockTableau:
    """RODAS3P (Kaps-Rentrop type, p = 3), five stages."""
    return RosenbrockTableau(
        name="rodas3p",
        a=(
            (0.0, 0.0, 0.0, 0.0, 0.0),
            (4.0 / 3.0, 0.0, 0.0, 0.0, 0.0),
            (0.0, 0.0, 0.0, 0.0, 0.0),
            (2.90625, 3.375, 0.40625, 0.0, 0.0),
            (2.90625, 3.375, 0.40625, 0.0, 0.0),
        ),
        C=(
            (0.0, 0.0, 0.0, 0.0, 0.0),
            (-4.0, 0.0, 0.0, 0.0, 0.0),
            (8.25, 6.75, 0.0, 0.0, 0.0),
            (1.21875, -5.0625, -1.96875, 0.0, 0.0),
            (4.03125, -15.1875, -4.03125, 6.0, 0.0),
        ),
        b=(2.90625, 3.375, 0.40625, 0.0, 1.0),
        b_hat=(2.90625, 3.375, 0.40625, 1.0, 0.0),
        c=(0.0, 4.0 / 9.0, 0.0, 1.0, 1.0),
        order=3,
        gamma=1.0 / 3.0,
        gamma_stages=(1.0 / 3.0, -1.0 / 9.0, 1.0, 0.0, 0.0),
    )


def _rosenbrock23() -> RosenbrockTableau:
    """Three-stage Rosenbrock 2(3) in increment form (SciML variant)."""
    s2 = math.sqrt(2.0)
    d = 1.0 / (2.0 + s2)
    inv_d = 2.0 + s2
    return RosenbrockTableau(
        name="rosenbrock23_sciml",
        a=((0.0, 0.0, 0.0), (0.5 * inv_d, 0.0, 0.0), (inv_d, inv_d, 0.0)),
        C=((0.0, 0.0, 0.0), (-inv_d, 0.0, 0.0), (-2.0 * inv_d, -(14.0 + 8.0 * s2), 0.0)),
        b=(inv_d, inv_d, 0.0),
        b_hat=(5.0 * s2 / 6.0 + 5.0 / 3.0, 1.0 / 3.0, -1.0 / 3.0 - s2 / 6.0),
        c=(0.0, 0.5, 1.0),
        order=2,
        gamma=d,
        gamma_stages=(d, 0.0, -d),
    )


ROS3P = _ros3p()
RODAS3P = _rodas3p()
ROSENBROCK_23 = _rosenbrock23()


# ---------------------------------------------------------------------------
# Alias registries (same names as the reference registries)
# ---------------------------------------------------------------------------

ERK_TABLEAU_REGISTRY: Dict[str, ERKTableau] = {
    "heun-21": HEUN_21,
    "ralston-33": RALSTON_33,
    "bogacki-shampine-32": BOGACKI_SHAMPINE_32,
    "rk23": BOGACKI_SHAMPINE_32,
    "ode23": BOGACKI_SHAMPINE_32,
    "dormand-prince-54": DORMAND_PRINCE_54,
    "dopri54": DORMAND_PRINCE_54,
    "rk45": DORMAND_PRINCE_54,
    "ode45": DORMAND_PRINCE_54,
    "classical-rk4": CLASSICAL_RK4,
    "rk4": CLASSICAL_RK4,
    "cash-karp-54": CASH_KARP_54,
    "fehlberg-45": FEHLBERG_45,
    "dormand-prince-853": DORMAND_PRINCE_853,
    "dop853": DORMAND_PRINCE_853,
    "tsit5": TSITOURAS_54,
    "vern7": VERNER_76,
}
DIRK_TABLEAU_REGISTRY: Dict[str, DIRKTableau] = {
    "implicit_midpoint": IMPLICIT_MIDPOINT,
    "trapezoidal_dirk": TRAPEZOIDAL_DIRK,
    "ode23t": TRAPEZOIDAL_DIRK,
    "kvaerno3": KVAERNO3,
    "kvaerno5": KVAERNO5,
    "sdirk_2_2": SDIRK_2_2,
    "l_stable_dirk_3": L_STABLE_DIRK3,
    "l_stable_sdirk_4": L_STABLE_SDIRK4,
}
FIRK_TABLEAU_REGISTRY: Dict[str, FIRKTableau] = {
    "firk_gauss_legendre_2": GAUSS_LEGENDRE_2,
    "firk_gauss_legendre_4": GAUSS_LEGENDRE_4,
    "radau_iia_5": RADAU_IIA_5,
    "radau": RADAU_IIA_5,
}
ROSENBROCK_TABLEAUS: Dict[str, RosenbrockTableau] = {
    "ros3p": ROS3P,
    "rodas3p": RODAS3P,
    "rosenbrock23": ROSENBROCK_23,
    "ode23s": ROSENBROCK_23,
    "rosenbrock23_sciml": ROSENBROCK_23,
}

DEFAULT_ERK_TABLEAU = DORMAND_PRINCE_54
DEFAULT_DIRK_TABLEAU = L_STABLE_DIRK3
DEFAULT_FIRK_TABLEAU = GAUSS_LEGENDRE_2
DEFAULT_ROSENBROCK_TABLEAU = ROS3P
