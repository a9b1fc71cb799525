"""Element tables used by the REPEAT fit path.

The values are data fixed by the reference and kept identical:
- ``VDW_RADII_BOHR``: UFF van der Waals radii in Bohr, indexed by ``Z - 1``
  (``repeat.py:564-579``). Lookup follows ``repeat.py:1936-1939``: ``Z = 0``
  maps to index -1 (the last entry), ``Z > 107`` raises ``IndexError``.
- ``ATOM_SYMBOLS``: element symbols indexed by ``Z`` (``repeat.py:581-592``),
  index 0 is ``"X"``.
- ``DEFAULT_QEQ_PARAMS``: (electronegativity, hardness) in Hartree
  (``repeat.py:611-650``), used by the QEq restraint when no RESP data is given.
"""

import numpy as np

# fmt: off
_VDW = (
    2.72687, 2.23177, 2.31586, 2.59365, 3.85788, 3.63867, 3.45820, 3.30702,  # Z=1..8 H He Li Be B C N O
    3.17852, 3.06419, 2.81853, 2.85443, 4.25094, 4.05819, 3.91835, 3.81252,  # Z=9..16 F Ne Na Mg Al Si P S
    3.72937, 3.65473, 3.60182, 3.21159, 3.11332, 2.99994, 2.97065, 2.85632,  # Z=17..24 Cl Ar K Ca Sc Ti V Cr
    2.79774, 2.75144, 2.71365, 2.67774, 3.30230, 2.61066, 4.14133, 4.04401,  # Z=25..32 Mn Fe Co Ni Cu Zn Ga Ge
    3.99677, 3.97315, 3.95803, 3.91268, 3.88717, 3.44025, 3.16057, 2.95175,  # Z=33..40 As Se Br Kr Rb Sr Y Zr
    2.99049, 2.88372, 2.83270, 2.79963, 2.76750, 2.73916, 2.97443, 2.69097,  # Z=41..48 Nb Mo Tc Ru Rh Pd Ag Cd
    4.21692, 4.14984, 4.17629, 4.22354, 4.25188, 4.16118, 4.26795, 3.49883,  # Z=49..56 In Sn Sb Te I Xe Cs Ba
    3.32781, 3.35993, 3.40718, 3.37789, 3.35143, 3.32592, 3.30041, 3.18230,  # Z=57..64 La Ce Pr Nd Pm Sm Eu Gd
    3.26072, 3.23899, 3.22104, 3.20403, 3.18797, 3.17002, 3.43930, 2.96781,  # Z=65..72 Tb Dy Ho Er Tm Yb Lu Hf
    2.99522, 2.89978, 2.79113, 2.94797, 2.68341, 2.60215, 3.11143, 2.55585,  # Z=73..80 Ta W Re Os Ir Pt Au Hg
    4.10732, 4.06008, 4.12905, 4.44936, 4.48810, 4.50227, 4.62983, 3.47426,  # Z=81..88 Tl Pb Bi Po At Rn Fr Ra
    3.28623, 3.20875, 3.23521, 3.20781, 3.23521, 3.23521, 3.19458, 3.14261,  # Z=89..96 Ac Th Pa U Np Pu Am Cm
    3.15490, 3.13033, 3.11710, 3.10482, 3.09348, 3.06892, 3.05758, 3.04513,  # Z=97..104 Bk Cf Es Fm Md No Lr Rf
    3.03268, 3.02023, 3.00778,  # Z=105..107 Db Sg Bh
)
VDW_RADII_BOHR = np.array(_VDW, dtype=np.float64)

ATOM_SYMBOLS = (
    "X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F",
    "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar", "K",
    "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu",
    "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",
    "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In",
    "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr",
    "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm",
    "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au",
    "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac",
    "Th", "Pa", "U", "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es",
    "Fm", "Md", "No", "Lr", "Rf", "Db", "Sg", "Bh", "Hs", "Mt",
    "Ds", "Rg", "Cn", "Nh", "Fl", "Mc", "Lv", "Ts", "Og",
)

DEFAULT_QEQ_PARAMS = {
    "H": (0.166401, 0.255225),
    "He": (0.449802, 0.899603),
    "Li": (0.110466, 0.175364),
    "Be": (0.179222, 0.326547),
    "B": (0.149493, 0.310892),
    "C": (0.199590, 0.215240),
    "N": (0.245778, 0.243355),
    "O": (0.320234, 0.314869),
    "F": (0.235784, 0.409058),
    "Ne": (0.396149, 0.792298),
    "Na": (0.104476, 0.168749),
    "Mg": (0.145193, 0.271424),
    "Al": (0.111752, 0.216449),
    "Si": (0.153168, 0.256284),
    "P": (0.200757, 0.293988),
    "S": (0.254594, 0.329707),
    "Cl": (0.213918, 0.267278),
    "Ar": (0.289578, 0.579156),
    "K": (0.088968, 0.141114),
    "Ca": (0.118734, 0.211671),
    "Sc": (0.135014, 0.151110),
    "Ti": (0.142033, 0.166471),
    "V": (0.150415, 0.154972),
    "Cr": (0.136447, 0.223945),
    "Mn": (0.169778, 0.182273),
    "Fe": (0.152763, 0.229605),
    "Co": (0.156402, 0.233353),
    "Ni": (0.161252, 0.237542),
    "Cu": (0.199512, 0.127447),
    "Zn": (0.136010, 0.164010),
    "Ga": (0.110209, 0.220564),
    "Ge": (0.148868, 0.252683),
    "As": (0.190651, 0.279950),
    "Se": (0.236219, 0.303616),
    "Br": (0.209178, 0.321925),
    "Kr": (0.257239, 0.514479),
    "Rb": (0.085661, 0.135675),
    "Sr": (0.111127, 0.179333),
    "Y": (0.144017, 0.165515),
    "Zr": (0.141408, 0.208143),
    "Nb": (0.141004, 0.216375),
    "Mo": (0.144164, 0.233500),
    "Tc": (0.153425, 0.236293),
    "Ru": (0.154711, 0.232250),
    "Rh": (0.158349, 0.233132),
    "Pd": (0.166618, 0.222475),
    "Ag": (0.163016, 0.230340),
    "Cd": (0.184992, 0.290828),
    "In": (0.110135, 0.205057),
    "Sn": (0.146516, 0.229605),
    "Sb": (0.180031, 0.245627),
    "Te": (0.213729, 0.259150),
    "I": (0.199586, 0.210207),
    "Xe": (0.222696, 0.445392),
    "Cs": (0.080222, 0.125753),
    "Ba": (0.103410, 0.176099),
    "La": (0.113553, 0.163898),
    "Ce": (0.113920, 0.161693),
    "Pr": (0.113920, 0.161693),
    "Nd": (0.113920, 0.161693),
    "Pm": (0.113920, 0.161693),
    "Sm": (0.113920, 0.161693),
    "Eu": (0.113920, 0.161693),
    "Gd": (0.113920, 0.161693),
    "Th": (0.113920, 0.161693),
}
# fmt: on


def vdw_radius_lookup(atomic_numbers):
    """Raw vdW radius (Bohr) per atom, indexed ``Z - 1`` as ``repeat.py:1936-1939``.

    ``Z = 0`` silently selects the last table entry (Python index -1) and
    ``Z > 107`` raises ``IndexError``, both as in the reference.
    """
    z = np.asarray(atomic_numbers, dtype=np.int64) - 1
    n = VDW_RADII_BOHR.shape[0]
    if np.any(z >= n) or np.any(z < -n):
        raise IndexError("atomic number outside the vdW radius table")
    return VDW_RADII_BOHR[z]


def element_symbol(z):
    """Symbol for atomic number ``z``; ``"X"`` past the table (``repeat.py:2515``)."""
    z = int(z)
    return ATOM_SYMBOLS[z] if z < len(ATOM_SYMBOLS) else "X"
