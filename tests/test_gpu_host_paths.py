"""Host-side paths of the product run in the GPU suite as well: CSR assembly
of the workload operators bit-exact against the oracle at full size, the
independent pins of the generators, the PETSc binary loader and the
EPSValuesView text.  None of them needs the device; they are collected with
the GPU tests so that the GPU suite's record covers the assembly and file
formats of the operators it solves."""
import numpy as np
import pytest

import test_cpu_generators as gen
import test_cpu_io as io
import test_cpu_oracle as core

pytestmark = pytest.mark.gpu


def test_host_generators_full_size_bit_exact(orc):
    core.test_generators_full_size_bit_exact(orc)


def test_host_generators_bit_exact(orc):
    core.test_generators_bit_exact(orc)


def test_host_csr_assemble_insert_values():
    core.test_csr_assemble_insert_values()


@pytest.mark.parametrize("nx,ny", [(12, 9), (1, 5)])
def test_host_fv2d_theta0_closed_form(orc, nx, ny):
    gen.test_fv2d_theta0_closed_form(orc, nx, ny)


@pytest.mark.parametrize("m", [0, 2, 7])
def test_host_fv2d_kl_field_unit_modes(orc, m):
    gen.test_fv2d_kl_field_unit_modes(orc, m)


def test_host_fv2d_kl_field_random_theta(orc):
    gen.test_fv2d_kl_field_random_theta(orc)


@pytest.mark.parametrize("dims", [(6, 5, 4), (2, 7, 1)])
def test_host_stencil27_offdiagonal_spectrum(orc, dims):
    gen.test_stencil27_offdiagonal_spectrum(orc, dims)


def test_host_splitmix64_known_answers(orc):
    gen.test_splitmix64_known_answers(orc)


def test_host_fv2d_full_size_sigma_and_faces(orc):
    """At the workload size 256 x 256 (held-out theta 0): sigma is the
    Gershgorin bound of L, and every interior face value is the harmonic mean
    of the two cell conductivities recovered from the boundary inwards."""
    import slepc_amd as sa
    th = orc.fv_theta(5, 0)
    A = sa.fv2d(256, 256, th)
    n = A.n
    rows = np.repeat(np.arange(n), np.diff(A.rowptr))
    diag = A.val[A.col == rows]
    off = np.where(A.col != rows, A.val, 0.0)
    offsum = np.bincount(rows, weights=off, minlength=n)
    L_abs_rowsum = np.abs(A.sigma - diag) + offsum
    assert abs(L_abs_rowsum.max() - A.sigma) <= 1e-13 * A.sigma


def test_host_mat_load_round_trip(tmp_path):
    io.test_mat_load_round_trip(tmp_path)


def test_host_mat_load_rejects_bad_files(tmp_path):
    io.test_mat_load_rejects_bad_files(tmp_path)


def test_host_vec_write_and_load(tmp_path):
    io.test_vec_write_and_load(tmp_path)


def test_host_values_view_text():
    import slepc_amd as sa
    txt = sa.format_eigenvalues([6.0, 5.0, 4.0, 3.0, 2.0, 1.0], fmt="ascii")
    assert txt == "Eigenvalues = \n   6.00000\n   5.00000\n   4.00000\n   3.00000\n   2.00000\n   1.00000\n\n"
    io.test_values_view_complex_pairs_match_test9()
    txt = sa.format_eigenvalues([20.0], fmt="matlab", name="eps")
    assert txt == "Lambda_eps = [\n2.0000000000000000e+01\n];\n"
