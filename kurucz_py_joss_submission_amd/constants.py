"""Physical constants and fixed parameters of the synthe hot path.

Values follow the reference so results agree bit for bit:
engine/opacity.py:39-45, engine/radiative.py:14-17, physics/populations.py:60-63,
physics/line_opacity.py:40-49, physics/josh_solver.py:16-22.
"""

H_PLANCK = 6.62607015e-27  # erg s
K_BOLTZ = 1.380649e-16  # erg / K
C_LIGHT_CM = 2.99792458e10  # cm / s
C_LIGHT_NM = 2.99792458e17  # nm / s
C_LIGHT_KM = 299792.458  # km / s
NM_TO_CM = 1e-7
AMU = 1.66054e-24  # g
CGF_CONSTANT = 0.026538 / 1.77245  # rgfall CONGF factor (line_opacity.py:46)
MAX_PROFILE_STEPS = 1_000_000  # MAXPROF (line_opacity.py:49)
JOSH_EPS = 1.0e-38
JOSH_ITER_TOL = 1.0e-5
NXTAU = 51
VOIGT_TABLE_SIZE = 2001
FASTEX_TABLE_SIZE = 1001
MAX_OPACITY_REAL4 = 3.4e38

ELEMENT_SYMBOLS = (
    "", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S",
    "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge",
    "As", "Se", "Br", "Kr", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd",
    "In", "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd",
    "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au", "Hg",
    "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U", "Np", "Pu", "Am", "Cm",
    "Bk", "Cf", "Es",
)
ELEMENT_Z = {sym.upper(): z for z, sym in enumerate(ELEMENT_SYMBOLS) if sym}


def element_atomic_number(symbol: str):
    """Atomic number of an element symbol, or None (engine/opacity.py:634-636)."""
    return ELEMENT_Z.get(str(symbol).strip().upper().replace(" ", ""))
