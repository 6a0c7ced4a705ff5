"""Known answers for the numpy preconditioner model (tests/precond_model.py)
that the device's preconditioners (block Jacobi, j-line, multicolour line
Gauss-Seidel, five-colour, natural-order and chunked line ILU) are checked
against.  Each one is the exact inverse on the matrices whose pattern it
captures completely: where the kept pattern holds all of the exact LU's
fill, block ILU(0) is the exact inverse; couplings between chunks are
dropped, which makes one-column chunks a block Jacobi over columns."""
import numpy as np

import precond_model as pm


def _random_blocks(nic, njc, nv, seed, drop_columns=False):
    """Random stencil blocks J (13 slots) with a dominant shift idt, so that
    A = I idt - J is well conditioned.  drop_columns: no coupling between
    different columns (slots with di != 0 are zero)."""
    rng = np.random.default_rng(seed)
    B = 0.1 * rng.standard_normal((nic, njc, len(pm.SLOTS), nv, nv))
    for s, (di, dj) in enumerate(pm.SLOTS):
        if drop_columns and di != 0:
            B[:, :, s] = 0.0
        for i in range(nic):
            for j in range(njc):
                if not (0 <= i + di < nic and 0 <= j + dj < njc):
                    B[i, j, s] = 0.0
    idt = 5.0 + rng.random((nic, njc))
    return B, idt


def _dense(B, idt):
    """A = I idt - J as a dense matrix, unknown (i, j, k) at (i njc + j) nv + k."""
    nic, njc, ns, nv, _ = B.shape
    A = np.zeros((nic * njc * nv, nic * njc * nv))
    for i in range(nic):
        for j in range(njc):
            r = (i * njc + j) * nv
            A[r:r + nv, r:r + nv] += np.eye(nv) * idt[i, j]
            for s, (di, dj) in enumerate(pm.SLOTS):
                ii, jj = i + di, j + dj
                if 0 <= ii < nic and 0 <= jj < njc:
                    c = (ii * njc + jj) * nv
                    A[r:r + nv, c:c + nv] -= B[i, j, s]
    return A


def test_chunked_ilu_is_exact_on_one_column():
    """One column: the matrix is block banded with half-bandwidth 2, its LU
    has no fill outside the band, so ILU(0) on the 13-point pattern is the
    exact LU; M^-1 and M^-T equal A^-1 and A^-T."""
    nic, njc, nv = 1, 14, 5
    B, idt = _random_blocks(nic, njc, nv, seed=1)
    A = _dense(B, idt)
    v = np.random.default_rng(2).standard_normal((nic, njc, nv))
    for tr in (False, True):
        z = pm.ilu_chunked(B, idt, v, 8, transpose=tr, f32=False)
        ref = np.linalg.solve(A.T if tr else A, v.reshape(-1)).reshape(v.shape)
        assert np.abs(z - ref).max() <= 1e-12 * np.abs(ref).max(), tr


def test_one_column_chunks_are_block_jacobi_over_columns():
    """Chunks of one column drop every coupling between columns: the model
    equals the exact solve of the column-block-diagonal part of A, and
    differs from A^-1 when columns are coupled."""
    nic, njc, nv = 4, 9, 5
    B, idt = _random_blocks(nic, njc, nv, seed=3)
    Bd = B.copy()
    for s, (di, dj) in enumerate(pm.SLOTS):
        if di != 0:
            Bd[:, :, s] = 0.0
    A, Ad = _dense(B, idt), _dense(Bd, idt)
    v = np.random.default_rng(4).standard_normal((nic, njc, nv))
    for tr in (False, True):
        z = pm.ilu_chunked(B, idt, v, 1, transpose=tr, f32=False)
        ref = np.linalg.solve(Ad.T if tr else Ad, v.reshape(-1)).reshape(v.shape)
        assert np.abs(z - ref).max() <= 1e-12 * np.abs(ref).max(), tr
        full = np.linalg.solve(A.T if tr else A, v.reshape(-1)).reshape(v.shape)
        assert np.abs(z - full).max() > 1e-4 * np.abs(full).max(), tr


def test_chunks_spanning_the_grid_without_cross_coupling_are_exact():
    """Without couplings between columns the chunk width does not matter:
    every chunk width gives the exact inverse of A (each column is exact,
    and ILU(0) of a block-diagonal matrix of banded columns is exact)."""
    nic, njc, nv = 5, 8, 4
    B, idt = _random_blocks(nic, njc, nv, seed=5, drop_columns=True)
    A = _dense(B, idt)
    v = np.random.default_rng(6).standard_normal((nic, njc, nv))
    ref = np.linalg.solve(A, v.reshape(-1)).reshape(v.shape)
    for w in (1, 2, 8):
        z = pm.ilu_chunked(B, idt, v, w, f32=False)
        assert np.abs(z - ref).max() <= 1e-12 * np.abs(ref).max(), w


def test_float32_factors_round_at_float32_level():
    """Storing the factors in float32 changes the result by about float32
    rounding (not more, and not zero)."""
    nic, njc, nv = 3, 10, 5
    B, idt = _random_blocks(nic, njc, nv, seed=7)
    v = np.random.default_rng(8).standard_normal((nic, njc, nv))
    z64 = pm.ilu_chunked(B, idt, v, 8, f32=False)
    z32 = pm.ilu_chunked(B, idt, v, 8, f32=True)
    rel = np.abs(z32 - z64).max() / np.abs(z64).max()
    assert 1e-10 < rel < 1e-5


def _keep_slots(B, slots):
    Bk = np.zeros_like(B)
    Bk[:, :, slots] = B[:, :, slots]
    return Bk


def _exact(B, idt, v):
    return np.linalg.solve(_dense(B, idt), v.reshape(-1)).reshape(v.shape)


def test_block_jacobi_and_line_models_are_exact_on_their_own_pattern():
    """Block Jacobi (1) is A^-1 when J holds only the diagonal blocks; the
    j-line solve (2) is A^-1 when J holds only the diagonal and the j +- 1
    blocks; with couplings outside those patterns they are not."""
    nic, njc, nv = 4, 7, 5
    B, idt = _random_blocks(nic, njc, nv, seed=9)
    v = np.random.default_rng(10).standard_normal((nic, njc, nv))
    for precond, slots in ((1, [0]), (2, [0, 1, 2])):
        Bk = _keep_slots(B, slots)
        ref = _exact(Bk, idt, v)
        z = pm.apply(Bk, idt, v, precond)
        assert np.abs(z - ref).max() <= 1e-12 * np.abs(ref).max(), precond
        full = _exact(B, idt, v)
        assert np.abs(pm.apply(B, idt, v, precond) - full).max() > 1e-4 * np.abs(full).max(), precond


def test_multicolour_line_gauss_seidel_converges_to_the_inverse():
    """The symmetric three-colour j-line Gauss-Seidel (3): one sweep is exact
    without couplings between lines, and repeated sweeps on a diagonally
    dominant A converge to A^-1 v."""
    nic, njc, nv = 6, 5, 4
    B, idt = _random_blocks(nic, njc, nv, seed=11)
    v = np.random.default_rng(12).standard_normal((nic, njc, nv))
    Bl = _keep_slots(B, [0, 1, 2])
    ref = _exact(Bl, idt, v)
    assert np.abs(pm.apply(Bl, idt, v, 3, sweeps=1) - ref).max() <= 1e-12 * np.abs(ref).max()
    ref = _exact(B, idt, v)
    one = np.abs(pm.apply(B, idt, v, 3, sweeps=1) - ref).max()
    many = np.abs(pm.apply(B, idt, v, 3, sweeps=40) - ref).max()
    assert many <= 1e-10 * np.abs(ref).max() and one > 1e3 * many


def test_five_colour_ilu_is_exact_for_colour_lower_triangular_matrices():
    """Five-colour ILU(0) (4): when every coupling points to a lower colour,
    A is block lower triangular in colour order, its LU has no fill and the
    model is A^-1; with couplings both ways it is not."""
    nic, njc, nv = 6, 6, 4
    B, idt = _random_blocks(nic, njc, nv, seed=13)
    Bl = B.copy()
    for i in range(nic):
        for j in range(njc):
            for s, (di, dj) in enumerate(pm.SLOTS):
                if s and not (0 <= i + di < nic and 0 <= j + dj < njc and pm.colour5(i + di, j + dj) < pm.colour5(i, j)):
                    Bl[i, j, s] = 0.0
    v = np.random.default_rng(14).standard_normal((nic, njc, nv))
    ref = _exact(Bl, idt, v)
    assert np.abs(pm.ilu_mc5(Bl, idt, v) - ref).max() <= 1e-12 * np.abs(ref).max()
    full = _exact(B, idt, v)
    assert np.abs(pm.ilu_mc5(B, idt, v) - full).max() > 1e-4 * np.abs(full).max()


def test_natural_order_ilu_is_exact_on_one_row():
    """Natural-order ILU(0) (5) orders cells j-major; on a grid one cell high
    the matrix is block banded with half-bandwidth 2 and the model is A^-1
    (M^-1 and M^-T)."""
    nic, njc, nv = 12, 1, 5
    B, idt = _random_blocks(nic, njc, nv, seed=15)
    A = _dense(B, idt)
    v = np.random.default_rng(16).standard_normal((nic, njc, nv))
    for tr in (False, True):
        ref = np.linalg.solve(A.T if tr else A, v.reshape(-1)).reshape(v.shape)
        z = pm.ilu_natural(B, idt, v, transpose=tr)
        assert np.abs(z - ref).max() <= 1e-12 * np.abs(ref).max(), tr


def test_nine_point_chunked_ilu_is_exact_on_one_tridiagonal_column():
    """The 3x3-box pattern (nine_point): on one column whose J holds only
    the j +- 1 couplings, the matrix is block tridiagonal, its LU has no
    fill and the model is A^-1; with the j +- 2 couplings present it is not
    (they are outside the kept pattern)."""
    nic, njc, nv = 1, 12, 5
    B, idt = _random_blocks(nic, njc, nv, seed=17)
    v = np.random.default_rng(18).standard_normal((nic, njc, nv))
    Bt = _keep_slots(B, [0, 1, 2])
    for tr in (False, True):
        A = _dense(Bt, idt)
        ref = np.linalg.solve(A.T if tr else A, v.reshape(-1)).reshape(v.shape)
        z = pm.ilu_chunked(Bt, idt, v, 8, transpose=tr, f32=False, nine_point=True)
        assert np.abs(z - ref).max() <= 1e-12 * np.abs(ref).max(), tr
    full = _exact(B, idt, v)
    z = pm.ilu_chunked(B, idt, v, 8, f32=False, nine_point=True)
    assert np.abs(z - full).max() > 1e-4 * np.abs(full).max()


def test_nine_point_chunked_ilu_is_exact_on_one_row_in_a_chunk():
    """The mirror case along i: one row of cells (njc = 1) inside one chunk,
    ordered by i.  With only the i +- 1 couplings the matrix is block
    tridiagonal in that order, its LU has no fill and the box-pattern model
    is A^-1; with the (+-2, 0) couplings present it is not (they lie outside
    the 3x3 box, so the keep rule along i drops them)."""
    nic, njc, nv = 8, 1, 5
    B, idt = _random_blocks(nic, njc, nv, seed=21)
    v = np.random.default_rng(22).standard_normal((nic, njc, nv))
    Bt = _keep_slots(B, [0, 3, 4])
    for tr in (False, True):
        A = _dense(Bt, idt)
        ref = np.linalg.solve(A.T if tr else A, v.reshape(-1)).reshape(v.shape)
        z = pm.ilu_chunked(Bt, idt, v, 8, transpose=tr, f32=False, nine_point=True)
        assert np.abs(z - ref).max() <= 1e-12 * np.abs(ref).max(), tr
    Bw = _keep_slots(B, [0, 3, 4, 11, 12])
    full = _exact(Bw, idt, v)
    z = pm.ilu_chunked(Bw, idt, v, 8, f32=False, nine_point=True)
    assert np.abs(z - full).max() > 1e-4 * np.abs(full).max()
    # the 13-point pattern keeps them and is exact again
    z13 = pm.ilu_chunked(Bw, idt, v, 8, f32=False)
    assert np.abs(z13 - full).max() <= 1e-12 * np.abs(full).max()


def test_float16_scaled_known_answer():
    """A block scaled by its largest magnitude 3: the entry 1 is stored as
    the float16 nearest to 1/3, 0x3555 = 0.333251953125, and read back as
    3 x that = 0.999755859375 (float16 without the scale would keep 1 exact);
    the largest entry and entries exact after scaling (1.5 -> 0.5, -0.75 ->
    -0.25) come back unchanged; an all-zero block stays zero."""
    blk = np.array([[3.0, 1.0], [1.5, -0.75]])
    out = pm.f16_scaled(blk)
    assert out[0, 1] == 0.999755859375
    assert out[0, 0] == 3.0 and out[1, 0] == 1.5 and out[1, 1] == -0.75
    assert np.array_equal(pm.f16_scaled(np.zeros((2, 2))), np.zeros((2, 2)))


def test_float16_scaled_factors_round_at_float16_level():
    """float16 storage with a scale per block moves the result by about
    float16 rounding (2^-11 relative per entry), far more than float32."""
    nic, njc, nv = 3, 10, 5
    B, idt = _random_blocks(nic, njc, nv, seed=19)
    v = np.random.default_rng(20).standard_normal((nic, njc, nv))
    z64 = pm.ilu_chunked(B, idt, v, 8, f32=False)
    z16 = pm.ilu_chunked(B, idt, v, 8, f16=True)
    z32 = pm.ilu_chunked(B, idt, v, 8, f32=True)
    e16 = np.abs(z16 - z64).max() / np.abs(z64).max()
    e32 = np.abs(z32 - z64).max() / np.abs(z64).max()
    assert 1e-6 < e16 < 1e-2 and e16 > 100 * e32
    blk = np.array([[3.0, -1e-3], [2.5e4, 0.0]])
    assert np.abs(pm.f16_scaled(blk) - blk).max() <= 2.0 ** -11 * 2.5e4
