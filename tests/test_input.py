"""The .dat input reader (hfscf_input_read): hfscf's input format, keyword table and errors."""
import numpy as np
import pytest

import hfscf_amd as H
from conftest import golden

ANG = 1.8897261254535


def test_water_in_angstrom():
    inp = H.read_input(golden("inputs/h2o_angstrom.dat"))
    assert inp["comment"] == "water RHF/cc-pVDZ, geometry in angstrom"
    assert list(inp["Z"]) == [8, 1, 1]
    ref = np.array([[0.0, -0.075791844, 0.0], [0.866811829, 0.601435779, 0.0], [-0.866811829, 0.601435779, 0.0]])
    assert np.abs(inp["xyz"] - ANG * ref).max() < 1e-15
    assert inp["basis_file"] == "cc-pvdz.gbs"
    assert inp["settings"]["energy_tol"] == "1e-10" and inp["settings"]["rms_tol"] == "1.0E-9"
    assert inp["settings"]["diis_range"] == "6" and inp["settings"]["eri_screen"] == "true"
    assert (inp["charge"], inp["multiplicity"]) == (0, 1)
    mol = H.Molecule.from_input(golden("inputs/h2o_angstrom.dat"))
    Z, X = mol.geometry()
    assert list(Z) == [8, 1, 1] and np.abs(X - ANG * ref).max() < 1e-15
    assert mol.nbf == 24 and mol.pure


def test_bohr_charge_dummy_atom_and_cartesian_override():
    inp = H.read_input(golden("inputs/ch2o_bohr.dat"))
    assert list(inp["Z"]) == [6, 8, 1, 1]  # the dummy atom X is dropped
    assert inp["xyz"][1][2] == 1.28 and inp["xyz"][2][1] == 1.77
    assert (inp["charge"], inp["multiplicity"]) == (1, 2)
    assert inp["basis_file"] == "6-31g_st__st_.gbs"
    mol = H.Molecule.from_input(golden("inputs/ch2o_bohr.dat"))
    assert not mol.pure


@pytest.mark.parametrize("body,message", [
    ("basis = cc-pVDZ\n{\n O 0 0 0\n}\n", "Unknown keyword"),
    ("diis_range = 6.5\n{\n O 0 0 0\n}\n", "Invalid keyword specification for diis_range"),
    ("units = nm\n{\n O 0 0 0\n}\n", "Invalid keyword specification for units"),
    ("energy_tol = 1e-x\n{\n O 0 0 0\n}\n", "Invalid keyword specification for energy_tol"),
    ("eri_screen = yes\n{\n O 0 0 0\n}\n", "Invalid keyword specification for eri_screen"),
    ("units = bohr\n{\n O\n H 1 1.0\n}\n", "z-matrix"),
    ("units = bohr\n{\n O 0 0 0\n", "missing an opening"),
    ("units = bohr\n", "missing an opening"),
    ("units = bohr\n{\n Qq 0 0 0\n}\n", "Unsupported element"),
    ("units = bohr\n{\n O 0 0 zero\n}\n", "Bad coordinate format"),
    ("units = bohr\n{\n}\n", "No atoms found"),
])
def test_errors(tmp_path, body, message):
    f = tmp_path / "in.dat"
    f.write_text("# test\n" + body)
    with pytest.raises(ValueError, match=message):
        H.read_input(f)


def test_comment_line_required(tmp_path):
    f = tmp_path / "in.dat"
    f.write_text("basis_set = sto-3g\n{\n O 0 0 0\n}\n")
    with pytest.raises(ValueError, match="comment line"):
        H.read_input(f)


def test_atom_block_limit(tmp_path):
    f = tmp_path / "in.dat"
    f.write_text("# big\n{\n" + "".join(f" H 0 0 {k}.0\n" for k in range(120)) + "}\n")
    with pytest.raises(ValueError, match="Maximum number of supported atoms"):
        H.read_input(f)
