"""CPU checks of the boundary: the C ABI library loads and exports every
declared symbol, and the seeded generators build the operators they claim
(against scipy Kronecker constructions and closed forms)."""
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

import utopia_amd as ua
from utopia_amd import _abi, petsc_io


def test_library_exports_every_declared_symbol():
    lib = ua.lib()  # loads without a GPU
    declared = set(_abi.declared_symbols())
    out = subprocess.run(["nm", "-D", "--defined-only", str(_abi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T uhip_" in l or " T isolver_lsolve_" in l}
    assert declared, "no prototypes parsed from include/utopia_hip.h"
    # the isolver plug-out's entry points (utopia_isolver_lsolve.cpp:42-220) under their reference names
    assert {n for n in declared if n.startswith("isolver_lsolve_")} == {
        "isolver_lsolve_" + f for f in ("init", "update_crs", "update_with_preconditioner_crs", "apply",
                                        "set_max_iterations", "set_atol", "set_stol", "set_rtol",
                                        "set_verbosity", "destroy")}
    assert declared <= exported, sorted(declared - exported)
    for name in declared:
        assert getattr(lib, name) is not None


def test_ctypes_structs_match_c_layout():
    import ctypes
    assert ua.lib().uhip_abi_sizeof(0) == ctypes.sizeof(ua.uhip_csr)
    assert ua.lib().uhip_abi_sizeof(1) == ctypes.sizeof(ua.uhip_cheb_step)


def test_library_carries_gfx950_code_object():
    # the offload bundle id of our own build names its target
    data = _abi.LIB_PATH.read_bytes()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
    assert b"amdgcn-amd-amdhsa--gfx942" not in data


@pytest.mark.parametrize("kind,dims,nnz", [
    (ua.STENCIL_2D5, (256, 256, 1), 326_656),
    (ua.STENCIL_3D7, (128, 128, 128), 14_581_760),
    (ua.STENCIL_3D7_VARCOEF, (256, 256, 256), 117_047_296),
    (ua.STENCIL_2D9, (1024, 1024, 1), 9_424_900),
])
def test_nnz_matches_workload_table(kind, dims, nnz):
    assert ua.lib().uhip_gen_nnz(kind, *dims) == nnz


def lap1d(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])


def test_poisson_2d_and_3d_match_kronecker():
    nx, ny = 7, 5
    A = ua.gen_stencil(ua.STENCIL_2D5, nx, ny).to_scipy()
    ref = sp.kron(sp.eye(ny), lap1d(nx)) + sp.kron(lap1d(ny), sp.eye(nx))
    assert abs(A - ref).max() == 0
    nx, ny, nz = 4, 3, 5
    A = ua.gen_stencil(ua.STENCIL_3D7, nx, ny, nz).to_scipy()
    ref = (sp.kron(sp.eye(nz), sp.kron(sp.eye(ny), lap1d(nx))) + sp.kron(sp.eye(nz), sp.kron(lap1d(ny), sp.eye(nx)))
           + sp.kron(lap1d(nz), sp.eye(nx * ny)))
    assert abs(A - ref).max() == 0


def test_nine_point_and_shift():
    nx, ny = 6, 4
    A = ua.gen_stencil(ua.STENCIL_2D9, nx, ny, shift=1e-3).to_scipy()
    one = sp.diags([np.ones(nx - 1), np.ones(nx), np.ones(nx - 1)], [-1, 0, 1])
    onez = sp.diags([np.ones(ny - 1), np.ones(ny), np.ones(ny - 1)], [-1, 0, 1])
    ref = -sp.kron(onez, one) + (9 + 1e-3) * sp.eye(nx * ny)
    assert abs(A - ref).max() < 1e-15


def test_varcoef_is_spd_with_harmonic_faces():
    nx, ny, nz, seed = 5, 4, 3, 7
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, nx, ny, nz, seed=seed).to_scipy().toarray()
    kap = np.exp(ua.gen_normal(seed, nx * ny * nz))
    assert np.allclose(A, A.T, rtol=0, atol=0)
    idx = lambda x, y, z: x + nx * (y + ny * z)
    i, j = idx(1, 1, 1), idx(2, 1, 1)
    assert A[i, j] == -2 * kap[i] * kap[j] / (kap[i] + kap[j])
    corner = idx(0, 0, 0)
    off = -A[corner].sum() + A[corner, corner]  # sum of the three interior faces
    assert np.isclose(A[corner, corner], off + 3 * kap[corner], rtol=1e-15)
    assert np.linalg.eigvalsh(A).min() > 0


def test_block_permuted_operator_is_the_renumbered_varcoef_operator():
    """Kind 5 = P A P^T with A the kind-2 operator and P a permutation that moves whole 64-row
    blocks (the tail block of 56 rows stays): recover P from the rows themselves, then compare."""
    nx, ny, nz, seed = 13, 11, 8, 9  # 1144 rows: 17 full blocks and a tail of 56
    A = ua.gen_stencil(ua.STENCIL_3D7_VARCOEF, nx, ny, nz, seed=seed).to_scipy()
    B = ua.gen_stencil(ua.STENCIL_3D7_BLOCKPERM, nx, ny, nz, seed=seed)
    n = A.shape[0]
    assert np.all(np.diff(B.rowptr) > 0)
    for r in range(n):  # sorted columns within every row
        c = B.col[B.rowptr[r]:B.rowptr[r + 1]]
        assert np.all(np.diff(c) > 0)
    Bs = B.to_scipy()
    assert abs(Bs - Bs.T).max() == 0
    # the diagonal identifies each row: kappa draws are distinct with probability 1
    da, db = A.diagonal(), Bs.diagonal()
    order_a, order_b = np.argsort(da), np.argsort(db)
    perm = np.empty(n, np.int64)
    perm[order_a] = order_b  # perm[old row] = new row
    assert np.array_equal(np.sort(da), np.sort(db)) and len(np.unique(da)) == n
    full = (n // 64) * 64
    assert np.array_equal(perm[full:], np.arange(full, n))
    blocks = perm[:full].reshape(-1, 64)
    assert np.all(blocks % 64 == np.arange(64)) and np.all(np.diff(blocks, axis=1) == 1)
    assert not np.array_equal(blocks[:, 0], np.arange(0, full, 64))  # not the identity
    P = sp.csr_matrix((np.ones(n), (perm, np.arange(n))), shape=(n, n))
    assert abs(P @ A @ P.T - Bs).max() == 0


def test_edge_coefficient_operator_is_symmetric_and_diagonally_dominant():
    A = ua.gen_stencil(ua.STENCIL_2D5_EDGECOEF, 9, 7, seed=3).to_scipy().toarray()
    assert np.array_equal(A, A.T)
    off = np.abs(A).sum(1) - np.abs(np.diag(A))
    # interior rows: diagonal = sum of the edge coefficients; boundary rows add Dirichlet edges
    assert np.all(np.diag(A) >= off - 1e-12) and np.any(np.diag(A) > off + 1e-3)
    assert np.linalg.eigvalsh(A).min() > 0


def p1d(nf):
    nc = nf // 2
    P = np.zeros((nf, nc))
    for f in range(nf):
        if f % 2:
            P[f, (f - 1) // 2] = 1.0
        else:
            for c in (f // 2 - 1, f // 2):
                if 0 <= c < nc:
                    P[f, c] = 0.5
    return P


def test_prolongation_is_tensor_of_linear_interpolation():
    P = ua.gen_prolongation(8, 6, 4).to_scipy().toarray()
    ref = np.kron(p1d(4), np.kron(p1d(6), p1d(8)))
    assert np.array_equal(P, ref)
    P2 = ua.gen_prolongation(8, 8, 1).to_scipy().toarray()
    assert np.array_equal(P2, np.kron(p1d(8), p1d(8)))


def test_uniform_and_normal_streams():
    u = ua.gen_uniform(42, 200_000)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.01
    assert np.array_equal(u, ua.gen_uniform(42, 200_000))
    z = ua.gen_normal(5, 200_000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def test_petsc_binary_roundtrip_and_reference_fixture(tmp_path, golden):
    A = petsc_io.read_mat(golden / "utopia_data" / "f_A")
    assert A.shape == (1681, 1681) and A.nnz == 13385
    b = petsc_io.read_vec(golden / "utopia_data" / "f_rhs")
    assert b.shape == (1681,)
    petsc_io.write_mat(tmp_path / "A", A)
    petsc_io.write_vec(tmp_path / "b", b)
    A2 = petsc_io.read_mat(tmp_path / "A")
    assert np.array_equal(A2.rowptr, A.rowptr) and np.array_equal(A2.col, A.col) and np.array_equal(A2.val, A.val)
    assert np.array_equal(petsc_io.read_vec(tmp_path / "b"), b)
    assert (tmp_path / "A").read_bytes() == (golden / "utopia_data" / "f_A").read_bytes()


def test_matrix_market_reference_fixture_against_scipy(tmp_path, golden):
    """gre_343_343_crg.mm (the reference's MatrixMarket fixture): the reader agrees with
    scipy.io.mmread entry for entry; write -> read round trips exactly; symmetric and
    pattern files mirror / fill as the format says."""
    import scipy.io
    path = golden / "utopia_data" / "gre_343_343_crg.mm"
    A = petsc_io.read_matrix_market(path)
    S = scipy.io.mmread(str(path)).tocsr()
    S.sum_duplicates()
    S.sort_indices()
    assert A.shape == S.shape == (343, 343)
    assert np.array_equal(A.rowptr, S.indptr) and np.array_equal(A.col, S.indices) and np.array_equal(A.val, S.data)
    petsc_io.write_matrix_market(tmp_path / "A.mm", A)
    B = petsc_io.read_matrix_market(tmp_path / "A.mm")
    assert np.array_equal(B.rowptr, A.rowptr) and np.array_equal(B.col, A.col) and np.array_equal(B.val, A.val)
    (tmp_path / "s.mm").write_text("%%MatrixMarket matrix coordinate real symmetric\n% c\n3 3 4\n1 1 2\n2 1 -1\n"
                                   "3 2 -1.5\n2 1 0.25\n")
    C = petsc_io.read_matrix_market(tmp_path / "s.mm").to_scipy().toarray()
    assert np.array_equal(C, np.array([[2, -0.75, 0], [-0.75, 0, -1.5], [0, -1.5, 0]]))
    (tmp_path / "p.mm").write_text("%%MatrixMarket matrix coordinate pattern skew-symmetric\n2 2 1\n2 1\n")
    assert np.array_equal(petsc_io.read_matrix_market(tmp_path / "p.mm").to_scipy().toarray(), [[0, -1], [1, 0]])
