"""The staggered summation-by-parts pair of the elastic walls, in float64 on the host.

A staggered first derivative maps half points to nodes (`node`) and nodes to
half points (`half`); near a wall the interior rows give way to the closure
rows of Prochnow, O'Reilly, Dunham & Petersson 2017
(https://doi.org/10.1016/j.compfluid.2017.03.015), whose norms make the pair
sum by parts, so the elastic operator stays the exact transpose it is inside
(cuwave/stencils.py:60-213). The kernels read only the low wall's rows that
differ from the interior one; the high wall's are their negated half turn.

`CORNERS` holds the published low-wall corner of each order: the rows of
Q+ over the first half points (the wall point first) and the two norms. The
values are the reference's own (cuwave/stencils.py:216-453); everything else
is computed here.
"""

from __future__ import annotations

from functools import lru_cache
from typing import NamedTuple

import numpy as np

from .stencils import staggered_weights


class Corner(NamedTuple):
    """The low-wall corner of one order on unit spacing."""

    rows: tuple  # (b, b + 1): Q+ rows of the first b nodes over half points 0 .. b
    node_norm: tuple  # (b,) H+ of the first b nodes, the wall node first
    half_norm: tuple  # (b,) H- of the first b half points


CORNERS = {
    2: Corner(
        rows=(
            (-1.0, 1.0),
        ),
        node_norm=(0.5,),
        half_norm=(1.0,),
    ),
    4: Corner(
        rows=(
            (-1.0, 1.0863474151236463, -0.048972800926292664, -0.04940682870347156, 0.012032214506117888),
            (0.0, -1.0912181712967304, 1.020182291667565, 0.106553819443748, -0.035517939814582664),
            (0.0, 0.007306134259626, -1.020182291667565, 1.020182291667565, -0.007306134259626),
            (0.0, -0.002435378086542, 0.048972800926292664, -1.118995949074508, 1.1141251929014242),
        ),
        node_norm=(0.3883101851851179, 1.126736111111313, 0.998263888888687, 0.9866898148148822),
        half_norm=(1.0839120370371043, 0.873263888888687, 1.0434027777779797, 0.999421296296229),
    ),
    6: Corner(
        rows=(
            (-1.0, 1.1394544837977862, -0.11190318831688334, -0.03679307743564517, -0.008806388435058028, 0.024120360922567204, -0.006072190532767),
            (0.0, -1.135642086672916, 1.020723413106962, 0.12901498096987027, 0.04871954028394196, -0.0882462187642217, 0.025430371076363405),
            (0.0, 0.014728358206091937, -1.0628624850408364, 1.1157308563134456, -0.14782738341548576, 0.11609496613710461, -0.03586431220032),
            (0.0, -0.01506154954283925, 0.11105278167363838, -1.1157307986574199, 1.1157308520205425, -0.11105286171832232, 0.015061576224400562),
            (0.0, -0.01631208707591403, 0.08569548827315682, -0.12901506745390884, -1.0716987136279488, 1.1709042547946922, -0.044261374910077415),
            (0.0, 0.012832881287791, -0.042706009696037585, 0.03679310626365804, 0.06856959317400804, -1.1722371680384867, 1.1571642636757338),
        ),
        node_norm=(0.35421198253308606, 1.1990789762234586, 1.010175380886416, 0.8578801746691397, 1.1030043571098747, 0.9756491285780251),
        half_norm=(1.152453960753951, 0.6425913073413562, 1.3188104408728434, 0.8718145591271567, 1.008971192658644, 1.005358539246049),
    ),
    8: Corner(
        rows=(
            (-1.0, 1.159277909708591, -0.10225822161033948, -0.12235496146373043, 0.06474950113711726, 0.018000625588968488, -0.02778298490069253, 0.013963295188980845, -0.0035951636488952684),
            (0.0, -1.08649665225212, 0.7358131866897168, 0.6469041496283849, -0.3324648837009965, -0.0336173293751992, 0.10332516547775163, -0.036946531461095986, 0.003482894993558249),
            (0.0, -0.08235828204281466, -0.736977485988319, 0.31951655484386693, 0.7032314422448357, -0.1703098567519811, -0.03624584851723378, -0.012075871596527, 0.015219347808173),
            (0.0, 0.02341742227992852, -0.147322512660754, 0.16790851042687072, -0.4371788026324864, 0.6774966552970128, -0.39016117877015266, 0.1220854834403491, -0.016245577380768088),
            (0.0, -0.059899186728965506, 0.49388162050978357, -1.4720495494260573, 0.2555966584017443, 0.13077614604961327, 0.7765785014136136, -0.107508816058426, -0.017375374161306),
            (0.0, 0.057499088529918, -0.2740066308298255, 0.4978806316852063, -0.3763835826294999, -0.34419035975624124, 0.5102303779913766, -0.08348589807751411, 0.013153917729437),
            (0.0, -0.006126089943821456, 0.010215000125073587, -0.019155436519664306, 0.1508827394702544, -0.34375282227419424, -0.9464701792134959, 1.1742811192535594, -0.02874709875485436),
            (0.0, -0.005314209550716, 0.020655043764664007, -0.018649899174876958, -0.02843307229096885, 0.06489939657916406, 0.019398914375975886, -1.14119261699885, 1.1595162796051317),
        ),
        node_norm=(0.338981562317581, 1.2170844209197904, 1.1833614859178774, 0.2549326866095341, 1.9558890726497251, 0.3525760140821226, 1.2319317827839134, 0.9652429747194561),
        half_norm=(1.2070182602821766, 0.368740688276087, 1.8546141620235908, 0.39080475586456204, 1.1422350920190358, 1.0954974451192663, 0.9261047151630665, 1.014984881252215),
    ),
}


class Pair(NamedTuple):
    """The staggered pair on an axis of `n` intervals, the wall points dropped."""

    node: np.ndarray  # (n + 1, n): half points -> nodes
    half: np.ndarray  # (n, n + 1): nodes -> half points
    node_norm: np.ndarray  # (n + 1,)
    half_norm: np.ndarray  # (n,)


def closure_norms(space_order: int, n: int) -> tuple[np.ndarray, np.ndarray]:
    """The pair's norms on `n` intervals: (n + 1,) on the nodes, (n,) on the half points,
    the corner's values at either wall (mirrored at the high one), one between."""
    corner = CORNERS[space_order]
    b = len(corner.node_norm)
    node, half = np.ones(n + 1), np.ones(n)
    node[:b] = corner.node_norm
    node[n + 1 - b:] = corner.node_norm[::-1]
    half[:b] = corner.half_norm
    half[n - b:] = corner.half_norm[::-1]
    return node, half


def staggered_pair(space_order: int, n: int, stencil: str = "taylor") -> Pair:
    """The pair on `n` intervals, its interior rows those of `stencil`.

    Q+ (nodes x half points, both wall points included) carries the interior
    weights on its band, the Taylor weights on the rows a corner replaces (the
    corner assumes them past its own columns), and the corners at either end,
    the high one the low one turned by half and negated. Q- is B - Q+^T, B the
    wall differences, so the pair sums by parts whatever the interior rows.
    Dividing by the norms gives the two derivatives; the wall points' columns
    and rows drop.
    """
    if space_order not in CORNERS:
        raise ValueError(f"no staggered closure of order {space_order}: {tuple(CORNERS)}")
    corner, R = CORNERS[space_order], space_order // 2
    b = len(corner.node_norm)
    if n < 2 * b:
        raise ValueError(f"{n} intervals leave no room for two closures of {b} rows")
    inner, taylor = staggered_weights(R, stencil), staggered_weights(R)
    qp = np.zeros((n + 1, n + 2))
    rows = np.arange(n + 1)
    for k in range(1, R + 1):
        right, left = rows + k, rows + 1 - k
        ok = right <= n + 1
        qp[rows[ok], right[ok]] = inner[k - 1]
        ok = left >= 0
        qp[rows[ok], left[ok]] = -inner[k - 1]
    for i in range(b):
        for k in range(1, R + 1):
            qp[i, i + k] = taylor[k - 1]
            qp[n - i, n + 1 - i - k] = -taylor[k - 1]
    low = np.array(corner.rows)
    qp[:b, :b + 1] = low
    qp[n + 1 - b:, n + 1 - b:] = -low[::-1, ::-1]
    walls = np.zeros_like(qp)
    walls[0, 0], walls[n, n + 1] = -1.0, 1.0
    qm = walls.T - qp.T
    node_norm, half_norm = closure_norms(space_order, n)
    return Pair(qp[:, 1:-1] / node_norm[:, None], qm[1:-1] / half_norm[:, None],
                node_norm, half_norm)


@lru_cache
def closure_tables(space_order: int, stencil: str = "taylor") -> dict:
    """The low wall's rows the kernels read and their extents.

    On an axis long enough that the two walls' rows never meet, a row of either
    derivative is a closure row if it differs (beyond 1e-15) from the interior
    row centred on it; the table holds the rows up to the last such one (at
    least one) and the columns up to the last point any of them reads.
    Returns {"node": (rows, span), "half": (rows, span), "counts": (node rows,
    half rows, node span, half span)}.
    """
    R, b = space_order // 2, len(CORNERS[space_order].node_norm)
    n = 4 * (b + 4 * R)
    pair = staggered_pair(space_order, n, stencil)
    c = staggered_weights(R, stencil)

    def count(matrix, shift):
        last = 0
        for i in range(n // 2):
            interior = np.zeros(matrix.shape[1])
            for k in range(1, R + 1):
                if 0 <= i + k - 1 + shift < len(interior):
                    interior[i + k - 1 + shift] = c[k - 1]
                if 0 <= i - k + shift < len(interior):
                    interior[i - k + shift] -= c[k - 1]
            if not np.allclose(matrix[i], interior, rtol=0.0, atol=1e-15):
                last = i + 1
        return max(last, 1)

    nrows, hrows = count(pair.node, 0), count(pair.half, 1)
    nspan = int(np.flatnonzero(np.abs(pair.node[:nrows]).sum(axis=0)).max()) + 1
    hspan = int(np.flatnonzero(np.abs(pair.half[:hrows]).sum(axis=0)).max()) + 1
    return {"node": pair.node[:nrows, :nspan], "half": pair.half[:hrows, :hspan],
            "counts": (nrows, hrows, nspan, hspan)}


def closure_band(space_order: int, stencil: str = "taylor") -> int:
    """Points from a wall within which either derivative's row is a closure row."""
    return max(closure_tables(space_order, stencil)["counts"][:2])
