"""Boundary tables (include/system_files of the reference) and their embedded copy."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "mantiis_amd", "data", "system_files")
EX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
EY = [0, 0, 1, 0, -1, 1, 1, -1, -1]
W = [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4
ORDER = [7, 3, 6, 4, 0, 2, 8, 1, 5]


def tables():
    bt = np.loadtxt(os.path.join(DATA, "bnd_types.txt")).astype(int)
    ifs = np.loadtxt(os.path.join(DATA, "ifstream.txt")).reshape(9, 52).T.astype(bool)
    icsr = np.loadtxt(os.path.join(DATA, "icsr.txt")).reshape(9, 52).T.astype(int) - 1
    return bt, ifs, icsr


def test_ifstream_is_the_neighbourhood_rule():
    """ifStream(t, d) <=> neighbour d fluid and, for a diagonal, both adjacent
    axis neighbours fluid: the rule the kernels' link words rely on."""
    bt, ifs, _ = tables()
    assert len(set(bt)) == 52
    for t in range(52):
        fl = {ORDER[b]: (bt[t] >> (8 - b)) & 1 for b in range(9)}
        assert fl[0] == 1
        for d in range(9):
            if d == 0:
                exp = True
            elif d <= 4:
                exp = bool(fl[d])
            else:
                a, b = {5: (1, 2), 6: (3, 2), 7: (3, 4), 8: (1, 4)}[d]
                exp = bool(fl[d] and fl[a] and fl[b])
            assert ifs[t, d] == exp, (t, d)


def test_icsr_targets_are_fluid_facing():
    """Every specular target of a wall link is a direction whose source is not a
    wall of the same cell (it points back into the fluid)."""
    bt, ifs, icsr = tables()
    for t in range(52):
        for d in range(1, 9):
            if not ifs[t, d]:
                s = icsr[t, d]
                assert 1 <= s <= 8


def test_embedded_tables_match_data():
    inc = open(os.path.join(ROOT, "mantiis_amd", "csrc", "mts_tables.inc")).read()
    bt, ifs, icsr = tables()
    codes = [int(v) for v in re.search(r"MTS_BND_CODE\[52\] = \{([^}]*)\}", inc).group(1).split(",")]
    assert codes == list(bt)
    body = re.search(r"MTS_ICSR\[52\]\[9\] = \{(.*?)\};", inc).group(1)
    rows = [[int(v) for v in r.split(",")] for r in re.findall(r"\{([^{}]*)\}", body)]
    assert (np.array(rows) == icsr).all()
    body = re.search(r"MTS_IFSTREAM\[52\]\[9\] = \{(.*?)\};", inc).group(1)
    rows = [[int(v) for v in r.split(",")] for r in re.findall(r"\{([^{}]*)\}", body)]
    assert (np.array(rows).astype(bool) == ifs).all()
