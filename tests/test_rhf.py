"""Host RHF input producer (include/ccsd_rhf.h) and the INCAR front-end, on the CPU.

Pins, each independent of this build's own code:
- published STO-3G water values (T. D. Crawford's programming projects, geometry
  in bohr): E_nuc = 8.002367061810450, E_SCF = -74.942079928192,
  E_MP2 = -0.049149636120, E_CCSD = -0.070680088376;
- the reference's own comment value, H2O/STO-3G at the INCAR~ geometry (bohr),
  E_corr = -0.0501273 (ccsd_memory_optimized.cpp:2003), printed to 7 digits;
- Szabo & Ostlund's H2/STO-3G integrals at R = 1.4 bohr (4 printed digits);
- two-electron systems, where CCSD equals full CI;
- rigid-motion invariance of E_HF.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import ccsd_ref, fci2, rhf_spin
from tests.conftest import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "ccsd_parallel_amd", "lib", "ccsd_incar")

WATER_BOHR = [[0.0, -0.143225816552, 0.0], [1.638036840407, 1.136548822547, 0.0],
              [-1.638036840407, 1.136548822547, 0.0]]
# config 1: O at origin, H at (+-0.7570, 0.5859, 0) angstrom
WATER_ANG = [[0.0, 0.0, 0.0], [0.7570, 0.5859, 0.0], [-0.7570, 0.5859, 0.0]]


def _capi():
    from ccsd_parallel_amd import capi
    return capi


def _oracle_ccsd(r, toler=1e-12, frozen=0):
    eps, mo, nocc = r["eps"], r["mo_eri"], r["nocc"]
    if frozen:
        eps, mo, nocc = eps[frozen:], mo[frozen:, frozen:, frozen:, frozen:], nocc - frozen
    prob = rhf_spin.problem_from_rhf(eps, mo, nocc)
    return ccsd_ref.ccsd_energy(prob, toler=toler, max_iter=500, diis_dim=8)[0]


def test_crawford_sto3g_water():
    r = _capi().rhf([8, 1, 1], WATER_BOHR, "STO-3G")
    assert r["converged"] == 1 and r["commutator"] < 1e-10
    assert abs(r["e_nuc"] - 8.002367061810450) < 1e-9
    assert abs(r["e_hf"] - (-74.942079928192)) < 1e-10
    assert abs(rhf_spin.mp2_closed_shell(r["eps"], r["mo_eri"], r["nocc"]) - (-0.049149636120)) < 1e-11
    assert abs(_oracle_ccsd(r) - (-0.070680088376)) < 1e-10


def test_reference_sto3g_comment_value():
    """INCAR~ reads H at (0, 1.079252, +-1.474611) bohr (its second z has a typo,
    -1.47611; the symmetric reading reproduces the printed value)."""
    R = [[0, 0, 0], [0, 1.079252, 1.474611], [0, 1.079252, -1.474611]]
    r = _capi().rhf([8, 1, 1], R, "STO-3G")
    assert abs(_oracle_ccsd(r, toler=1e-11) - (-0.0501273)) < 5e-8


def test_h2_sto3g_szabo_integrals():
    capi = _capi()
    R = [[0, 0, 0], [0, 0, 1.4]]
    S, T, V = capi.ao_one_electron([1, 1], R, "STO-3G")
    assert abs(S[0, 1] - 0.6593) < 1e-4 and abs(S[0, 0] - 1.0) < 1e-12
    assert abs(T[0, 0] - 0.7600) < 1e-4 and abs(T[0, 1] - 0.2365) < 1e-4
    assert abs(V[0, 0] - (-1.2266 - 0.6538)) < 2e-4 and abs(V[0, 1] - 2 * (-0.5974)) < 2e-4
    r = capi.rhf([1, 1], R, "STO-3G", want_ao_eri=True)
    g = r["ao_eri"]
    for idx, ref in (((0, 0, 0, 0), 0.7746), ((0, 0, 1, 1), 0.5697), ((1, 0, 0, 0), 0.4441), ((1, 0, 1, 0), 0.2970)):
        assert abs(g[idx] - ref) < 1e-4, (idx, g[idx])
    assert abs(r["e_hf"] - (-1.1167)) < 1e-4


def test_h2_631g_ccsd_equals_fci():
    r = _capi().rhf([1, 1], [[0, 0, 0], [0.3, -0.2, 1.45]], "6-31G")
    f, g, O, V = rhf_spin.spin_orbital_tensor(r["eps"], r["mo_eri"], r["nocc"])
    assert abs(_oracle_ccsd(r, toler=1e-13) - fci2.fci_two_electron(f, g)[0]) < 1e-10


def test_rigid_motion_invariance():
    capi = _capi()
    r0 = capi.rhf([8, 1, 1], WATER_ANG, "6-31G", angstrom=True)
    rng = np.random.default_rng(3)
    Q = np.linalg.qr(rng.standard_normal((3, 3)))[0]
    R1 = np.array(WATER_ANG) @ Q.T + np.array([0.4, -1.1, 2.3])
    r1 = capi.rhf([8, 1, 1], R1, "6-31G", angstrom=True)
    assert abs(r0["e_hf"] - r1["e_hf"]) < 1e-10
    assert np.max(np.abs(r0["eps"] - r1["eps"])) < 1e-8


def test_config1_scf_is_tight():
    capi = _capi()
    assert capi.basis_size([8, 1, 1], "6-31G") == 13 and capi.basis_size([8, 1, 1], "STO-3G") == 7
    r = capi.rhf([8, 1, 1], WATER_ANG, "6-31G", angstrom=True)
    assert r["converged"] == 1 and r["commutator"] < 1e-10 and r["nocc"] == 5 and r["n_basis"] == 13
    # MO coefficients orthonormal in the AO metric
    S, _, _ = capi.ao_one_electron([8, 1, 1], WATER_ANG, "6-31G", angstrom=True)
    assert np.max(np.abs(r["C"].T @ S @ r["C"] - np.eye(13))) < 1e-10


def test_rhf_errors():
    capi = _capi()
    with pytest.raises(capi.CCSDError):
        capi.rhf([8, 1, 1], WATER_ANG, "6-31G", n_electrons=9, angstrom=True)  # open shell
    with pytest.raises(capi.CCSDError):
        capi.rhf([6, 1], [[0, 0, 0], [0, 0, 2.0]], "6-31G")  # no basis data for carbon here
    with pytest.raises(capi.CCSDError):
        capi.basis_size([8], "cc-pVDZ")


def _incar(tmp_path, method, extra="", atoms=None, angstrom=True, basis="6-31G"):
    atoms = atoms or [("O", WATER_ANG[0]), ("H", WATER_ANG[1]), ("H", WATER_ANG[2])]
    lines = [f"Basis_Set {basis}", f"Method {method}", "convergance_criteria -10.0", "Relax_Pos false",
             f"use_angstrom {'true' if angstrom else 'false'}", "print_stuffies false", extra, "", "#ATOMS START"]
    lines += [f"{s} {x} {y} {z}" for s, (x, y, z) in atoms] + ["#ATOMS END", ""]
    p = tmp_path / "INCAR"
    p.write_text("\n".join(lines))
    return str(p)


def _run(path):
    return subprocess.run([APP, path], capture_output=True, text=True, timeout=600)


def test_incar_hf_matches_library(tmp_path):
    out = _run(_incar(tmp_path, "HF"))
    assert out.returncode == 0, out.stderr
    e = float(out.stdout.split("E_HF")[1].split()[0])
    r = _capi().rhf([8, 1, 1], WATER_ANG, "6-31G", angstrom=True)
    assert abs(e - r["e_hf"]) < 1e-10


def test_incar_rejects_unsupported(tmp_path):
    assert _run(_incar(tmp_path, "CCSDT")).returncode != 0
    p = _incar(tmp_path, "HF")
    txt = open(p).read().replace("Relax_Pos false", "Relax_Pos true")
    open(p, "w").write(txt)
    assert _run(p).returncode != 0


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_incar_ccsd_without_device_fails_loudly(tmp_path):
    out = _run(_incar(tmp_path, "CCSD"))
    assert out.returncode != 0 and "no HIP device" in out.stderr
