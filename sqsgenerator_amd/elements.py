"""Chemical element symbols indexed by atomic number (Z = 0 is a vacancy, "0")."""

SYMBOLS = (
    "0 H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge "
    "As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm "
    "Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U "
    "Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og"
).split()

Z_OF = {s: z for z, s in enumerate(SYMBOLS)}


def to_z(value) -> int:
    """Atomic number of a symbol or an integer."""
    if isinstance(value, str):
        if value not in Z_OF:
            raise KeyError(f'An atomic element with name "{value}" is not known to me')
        return Z_OF[value]
    z = int(value)
    if not 0 <= z < len(SYMBOLS):
        raise KeyError(f"Unknown atomic number {z}")
    return z


def symbol(z: int) -> str:
    return SYMBOLS[z]
