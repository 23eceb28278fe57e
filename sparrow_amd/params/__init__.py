"""NDDO parameter tables (<method>.txt, generated by make_params.py) and element data."""
import os

PARAM_DIR = os.path.dirname(os.path.abspath(__file__))


def core_charge(Z):
    """Core charge of element Z (PM6Elements::getCoreCharge, ParameterUtils/PrincipalQuantumNumbers.cpp:138-163)."""
    if Z in (10, 18, 36, 54):
        return 6
    for c in (86, 78, 68, 54, 46, 36, 28, 18, 10, 2):
        if Z > c:
            return Z - c
    return Z
